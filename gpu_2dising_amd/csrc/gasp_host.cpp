// gasp_host.cpp -- host-side inputs and outputs of the engine: initial grids
// and fates (src/grid.c), the gridinput.bin / gridstates.bin formats
// (src/io.c:4-130) and the bootstrap error bar (src/bootstrap.c).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <vector>

#include "gasp_internal.h"

using gasp::set_error;

extern "C" {

// src/grid.c:5-21
int *init_grids_uniform(int L, int ngrids, int spin) {
    gasp::clear_error();
    const size_t n = (size_t)L * L * ngrids;
    int *g = (L > 0 && ngrids > 0) ? (int *)malloc(n * sizeof(int)) : nullptr;
    if (!g) {
        set_error(GASP_ERR_NOMEM, "Error allocating memory for Ising grids!");
        return nullptr;
    }
    for (size_t i = 0; i < n; ++i) g[i] = spin;
    return g;
}

// src/grid.c:24-38
int *init_grids_from_file(int L, int ngrids) {
    gasp::clear_error();
    const size_t n = (size_t)L * L * ngrids;
    int *g = (L > 0 && ngrids > 0) ? (int *)malloc(n * sizeof(int)) : nullptr;
    if (!g) {
        set_error(GASP_ERR_NOMEM, "Error allocating memory for Ising grids!");
        return nullptr;
    }
    read_input_grid(L, ngrids, g);
    if (gasp_last_error()) {
        free(g);
        return nullptr;
    }
    return g;
}

// src/grid.c:41-71: input grid m is copied into replicas
// [m*ngrids/ninputs, (m+1)*ngrids/ninputs)
int *init_grids_from_array(int L, int ngrids, int ninputs, int *array) {
    gasp::clear_error();
    if (ninputs < 1 || ngrids % ninputs != 0) {
        set_error(GASP_ERR_ARG, "Error: ngrids must be a multiple of ninputs!");
        return nullptr;
    }
    const size_t N = (size_t)L * L;
    int *g = (L > 0 && ngrids > 0 && array) ? (int *)malloc(N * ngrids * sizeof(int)) : nullptr;
    if (!g) {
        set_error(GASP_ERR_NOMEM, "Error allocating memory for Ising grids!");
        return nullptr;
    }
    const int per = ngrids / ninputs;
    for (int m = 0; m < ninputs; ++m)
        for (int i = 0; i < per; ++i) memcpy(g + ((size_t)m * per + i) * N, array + (size_t)m * N, N * sizeof(int));
    return g;
}

// src/grid.c:74-87
int *init_fates(int ngrids) {
    gasp::clear_error();
    int *f = ngrids > 0 ? (int *)malloc((size_t)ngrids * sizeof(int)) : nullptr;
    if (!f) {
        set_error(GASP_ERR_NOMEM, "Error allocating memory for grid fates");
        return nullptr;
    }
    for (int i = 0; i < ngrids; ++i) f[i] = -1;
    return f;
}

// src/io.c:4-65.  Reads gridinput.bin: int32 L, then ceil(L*L/8) bytes, bit k
// (LSB first) set <=> spin k is +1; the grid is copied into every replica.
// Padding bits of the last byte are ignored (the reference writes them past
// the end of each grid).
void read_input_grid(int L, int ngrids, int *ising_grids) {
    gasp::clear_error();
    const char *filename = "gridinput.bin";
    FILE *f = fopen(filename, "rb");
    if (!f) {
        set_error(GASP_ERR_IO, "Error opening %s for input!", filename);
        return;
    }
    int Lcheck = 0;
    if (fread(&Lcheck, sizeof(int), 1, f) != 1 || Lcheck != L) {
        fclose(f);
        set_error(GASP_ERR_IO, "Error - size of grid in input file does not match L!");
        return;
    }
    const size_t N = (size_t)L * L;
    const size_t nbytes = (N + 7) / 8;
    std::vector<unsigned char> bits(nbytes);
    const size_t got = fread(bits.data(), 1, nbytes, f);
    fclose(f);
    if (got != nbytes) {
        set_error(GASP_ERR_IO, "Error - %s is truncated", filename);
        return;
    }
    for (size_t s = 0; s < N; ++s) {
        const int v = ((bits[s >> 3] >> (s & 7)) & 1) ? 1 : -1;
        for (int g = 0; g < ngrids; ++g) ising_grids[(size_t)g * N + s] = v;
    }
    fprintf(stderr, "Read initial configuration of all grids from gridinput.bin\n");
}

// src/io.c:67-130.  Appends one record: int32 L, int32 ngrids, int32 isweep,
// then ceil(L*L*ngrids/8) bytes packed LSB first over all grids.
int write_ising_grids(int L, int ngrids, int *ising_grids, int isweep, float *magnetisation, float *lclus_size,
                      char *cv, double dn_thr, double up_thr, char *filename) {
    (void)magnetisation;
    (void)lclus_size;
    (void)cv;
    (void)dn_thr;
    (void)up_thr;
    gasp::clear_error();
    const size_t n = (size_t)L * L * ngrids;
    const size_t nbytes = (n + 7) / 8;
    std::vector<unsigned char> bits(nbytes, 0);
    for (size_t i = 0; i < n; ++i)
        if (ising_grids[i] == 1) bits[i >> 3] |= (unsigned char)(1u << (i & 7));
    return gasp::append_gridstates(L, ngrids, isweep, bits.data(), filename);
}


// src/bootstrap.c:6-63.  m resampled committors over ngrids fates, indices
// drawn with libc rand() % ngrids before counting; the population standard
// deviation of the m estimates is returned (-1.0 on allocation failure).
double bootstrap_errbar(int ngrids, int m, int *grid_fate_host) {
    if (ngrids < 1 || m < 1 || !grid_fate_host) return -1.0;
    std::vector<int> ind;
    std::vector<double> store;
    try {
        ind.resize(ngrids);
        store.resize(m);
    } catch (...) {
        fprintf(stderr, "Memory allocation failed for ngrids_rand_ind\n");
        return -1.0;
    }
    for (int im = 0; im < m; ++im) {
        int nA = 0, nB = 0;
        for (int i = 0; i < ngrids; ++i) ind[i] = rand() % ngrids;
        for (int i = 0; i < ngrids; ++i) {
            const int f = grid_fate_host[ind[i]];
            if (f == 0) nA++;
            else if (f == 1) nB++;
        }
        store[im] = (double)nB / (double)(nA + nB);
    }
    double mean = 0.0, num = 0.0;
    for (int im = 0; im < m; ++im) mean = mean + store[im];
    mean = mean / m;
    for (int im = 0; im < m; ++im) num = num + (store[im] - mean) * (store[im] - mean);
    return sqrt(num / m);
}

// bootstrap_errbar for `ninputs` successive inputs (n fates each), exactly as
// `ninputs` calls of bootstrap_errbar in a row after srand(seed), with the
// inputs spread over threads; see libc_window below.
int gasp_bootstrap_errbar_libc(const int *fates, int ninputs, int n, int m, unsigned seed, int nthreads,
                               double *out);

}  // extern "C"

// ------------------------------------------------------------ libc rand() stream
// glibc's rand() is random_r's TYPE_3 additive feedback generator: srand(s)
// sets r[0] = s (1 if s = 0, as a 32-bit signed value) and
// r[i] = 16807 r[i-1] mod (2^31 - 1) for i = 1..30 (Schrage's method),
// r[i] = r[i-31] for i = 31..33, and r[i] = r[i-3] + r[i-31] mod 2^32 for
// i >= 34; call k (k = 0, 1, ...) returns r[k + 344] >> 1.  The recurrence is
// linear, with characteristic polynomial x^31 - x^28 - 1: if
// x^d = sum_j a_j x^j modulo it, then r[t + d] = sum_j a_j r[t + j] for t >= 3,
// so the stream jumps ahead by any number of calls.
namespace {
using Poly = std::array<uint32_t, 31>;

Poly poly_mulmod(const Poly &a, const Poly &b) {
    uint32_t prod[61] = {};
    for (int i = 0; i < 31; ++i)
        for (int j = 0; j < 31; ++j) prod[i + j] += a[i] * b[j];
    for (int k = 60; k >= 31; --k) {  // x^k = x^(k-3) + x^(k-31)
        prod[k - 3] += prod[k];
        prod[k - 31] += prod[k];
    }
    Poly r;
    std::copy(prod, prod + 31, r.begin());
    return r;
}

Poly x_pow(uint64_t d) {
    Poly r{}, b{};
    r[0] = 1u;
    b[1] = 1u;
    for (; d; d >>= 1) {
        if (d & 1) r = poly_mulmod(r, b);
        b = poly_mulmod(b, b);
    }
    return r;
}

// r[s .. s+30] for s = 313 + call, the 31 values before the one call `call` uses
void libc_window(unsigned seed, uint64_t call, uint32_t w[31]) {
    int32_t r0[34];
    int32_t word = seed == 0 ? 1 : (int32_t)seed;
    r0[0] = word;
    for (int i = 1; i < 31; ++i) {
        const int32_t hi = word / 127773, lo = word % 127773;
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r0[i] = word;
    }
    for (int i = 31; i < 34; ++i) r0[i] = r0[i - 31];
    uint32_t base[31];  // r[3 .. 33]
    for (int j = 0; j < 31; ++j) base[j] = (uint32_t)r0[3 + j];
    Poly a = x_pow(310 + call);  // r[313 + call] from the window at t = 3
    for (int k = 0; k < 31; ++k) {
        uint32_t v = 0u;
        for (int j = 0; j < 31; ++j) v += a[j] * base[j];
        w[k] = v;
        // a <- a x, reduced: the x^31 term becomes x^28 + 1
        const uint32_t top = a[30];
        for (int j = 30; j > 0; --j) a[j] = a[j - 1];
        a[0] = top;
        a[28] += top;
    }
}

// The reference's estimator for one input, drawing from the stream in w
// (advanced in place): m resamples of n indices rand() % n.
double bootstrap_one(const int *f, int n, int m, uint32_t w[31], std::vector<uint32_t> &buf,
                     std::vector<double> &store) {
    const size_t need = (size_t)n;
    const uint64_t M = UINT64_MAX / (uint64_t)n + 1u;
    buf.resize(need + 31);
    for (int im = 0; im < m; ++im) {
        std::copy(w, w + 31, buf.begin());
        for (size_t i = 31; i < need + 31; ++i) buf[i] = buf[i - 3] + buf[i - 31];
        std::copy(buf.end() - 31, buf.end(), w);
        // (r >> 1) % n by Lemire's exact 32-bit remainder: M = ceil(2^64 / n),
        // then the high word of (M (r >> 1) mod 2^64) n; counts without branches
        int nA = 0, nB = 0;
        for (size_t i = 0; i < need; ++i) {
            const uint64_t low = M * (uint64_t)(buf[31 + i] >> 1);
            const int g = (int)(((unsigned __int128)low * (uint64_t)n) >> 64);
            const int x = f[g];
            nA += (x == 0);
            nB += (x == 1);
        }
        store[im] = (double)nB / (double)(nA + nB);
    }
    double mean = 0.0, num = 0.0;
    for (int im = 0; im < m; ++im) mean = mean + store[im];
    mean = mean / m;
    for (int im = 0; im < m; ++im) num = num + (store[im] - mean) * (store[im] - mean);
    return sqrt(num / m);
}
}  // namespace

extern "C" int gasp_bootstrap_errbar_libc(const int *fates, int ninputs, int n, int m, unsigned seed,
                                          int nthreads, double *out) {
    gasp::clear_error();
    if (!fates || !out || ninputs < 0 || n < 1 || m < 1)
        return gasp::set_error(GASP_ERR_ARG, "gasp_bootstrap_errbar_libc: bad argument");
    constexpr int kMaxThreads = 256;
    if (nthreads < 1) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::max(1, std::min({nthreads, ninputs, kMaxThreads}));
    const uint64_t per_input = (uint64_t)n * (uint64_t)m;  // rand() calls of one input
    std::atomic<bool> failed{false};
    auto work = [&](int i0, int i1) {
        try {  // an exception must not leave a thread (std::terminate)
            uint32_t w[31];
            libc_window(seed, (uint64_t)i0 * per_input, w);
            std::vector<uint32_t> buf;
            std::vector<double> store((size_t)m);
            for (int i = i0; i < i1; ++i) out[i] = bootstrap_one(fates + (size_t)i * n, n, m, w, buf, store);
        } catch (...) {
            failed = true;
        }
    };
    std::vector<std::thread> pool;
    const int chunk = (ninputs + nthreads - 1) / nthreads;
    try {
        for (int t = 1; t < nthreads; ++t) {
            const int i0 = t * chunk, i1 = std::min(ninputs, i0 + chunk);
            if (i0 < i1) pool.emplace_back(work, i0, i1);
        }
    } catch (...) {
        for (auto &th : pool) th.join();
        return gasp::set_error(GASP_ERR_NOMEM, "gasp_bootstrap_errbar_libc: cannot start threads");
    }
    work(0, std::min(ninputs, chunk));
    for (auto &th : pool) th.join();
    if (failed) return gasp::set_error(GASP_ERR_NOMEM, "gasp_bootstrap_errbar_libc: out of host memory");
    return 0;
}

int gasp::append_gridstates(int L, int ngrids, int isweep, const uint8_t *bits, const char *filename) {
    const char *name = filename ? filename : "gridstates.bin";
    FILE *f = fopen(name, "ab");
    if (!f) f = fopen(name, "wb");
    if (!f) return set_error(GASP_ERR_IO, "Error opening %s for write!", name);
    const int hdr[3] = {L, ngrids, isweep};
    const size_t nbytes = ((size_t)L * L * ngrids + 7) / 8;
    const bool ok = fwrite(hdr, sizeof(int), 3, f) == 3 && fwrite(bits, 1, nbytes, f) == nbytes;
    fclose(f);
    if (!ok) return set_error(GASP_ERR_IO, "Error writing %s", name);
    return 0;
}

