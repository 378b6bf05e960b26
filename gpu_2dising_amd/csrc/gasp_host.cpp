// gasp_host.cpp -- host-side inputs and outputs of the engine: initial grids
// and fates (src/grid.c), the gridinput.bin / gridstates.bin formats
// (src/io.c:4-130) and the bootstrap error bar (src/bootstrap.c).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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

}  // extern "C"

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

