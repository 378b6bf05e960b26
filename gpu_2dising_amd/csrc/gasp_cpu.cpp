// gasp_cpu.cpp -- the CPU fallback of the drop-in: the reference's CPU engine
// (src/mc_cpu.c) with its single process-global MT19937 stream
// (src/mt19937ar.c), run by the command line program and the Python module
// when gpuInitDevice finds no GPU (src/ising.cu:127-155,
// src/ising_ext.cu:637-664,1007-1030,1179-1187).
//
// The arithmetic is the reference's, bit for bit, written in integers:
//  - genrand_real3 is xi = (u + 0.5) 2^-32 for the 32-bit output u, so
//    floor(L xi) = ((2u + 1) L) >> 33 exactly (no rounding in either form);
//  - "xi < Pacc[idx]" with the double table is "u < T[idx]" for the exact
//    integer threshold T = ceil(Pacc 2^32 - 0.5) (2^32 when Pacc >= 1).
// MT19937 itself is std::mt19937: the C++ standard fixes it as the
// Matsumoto-Nishimura generator with the init_genrand seeding
// (x_i = 1812433253 (x_(i-1) ^ (x_(i-1) >> 30)) + i) and the same tempering.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <random>
#include <vector>

#include "gasp_internal.h"

using gasp::set_error;

namespace {

// src/mt19937ar.c:56-57: one stream for the process; before init_genrand it
// behaves as if seeded with 5489, as the reference does (src/mt19937ar.c:113).
std::mt19937 g_mt(5489u);

// src/mc_cpu.c:10: double Pacc[20], indexed 5(s+1) + nsum + 4
double g_pacc[20];
uint64_t g_thr[20];  // exact integer images of g_pacc

// Union-find root with path halving.
inline int uf_find(std::vector<int> &p, int x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

}  // namespace

extern "C" {

// src/mt19937ar.c:60-73
void init_genrand(unsigned long s) { g_mt.seed((std::mt19937::result_type)(s & 0xffffffffUL)); }

// src/mt19937ar.c:105-140
unsigned long genrand_int32(void) { return (unsigned long)g_mt(); }

// src/mt19937ar.c:163-167: (u + 0.5) / 2^32, in (0, 1)
double genrand_real3(void) { return ((double)g_mt() + 0.5) * (1.0 / 4294967296.0); }

// src/mc_cpu.c:14-38: Pacc[5(s+1) + nsum + 4] = exp(-beta 2 s (nsum + h))
void preComputeProbs_cpu(double beta, double h) {
    for (int s = -1; s < 2; s += 2)
        for (int nsum = -4; nsum < 5; nsum += 2) {
            const int idx = 5 * (s + 1) + nsum + 4;
            g_pacc[idx] = exp(-beta * (2.0 * (double)s * ((double)nsum + h)));
            g_thr[idx] = gasp::threshold_of(g_pacc[idx]);
        }
}

// src/mc_cpu.c:41-89.  Per attempt three draws from the global stream: row,
// column, acceptance.  beta and h are unused (the table holds them), as in
// the reference.
void mc_sweep_cpu(int L, int *ising_grids, int grid_index, double beta, double h, int nsweeps) {
    (void)beta;
    (void)h;
    if (L < 1 || nsweeps < 1 || !ising_grids) return;
    const size_t N = (size_t)L * L;
    int *g = ising_grids + (size_t)grid_index * N;
    const uint64_t uL = (uint64_t)L;
    const long long n = (long long)N * nsweeps;
    for (long long a = 0; a < n; ++a) {
        const uint64_t u1 = g_mt(), u2 = g_mt();
        const int row = (int)(((2 * u1 + 1) * uL) >> 33);
        const int col = (int)(((2 * u2 + 1) * uL) >> 33);
        const int up = row == 0 ? L - 1 : row - 1;
        const int dn = row == L - 1 ? 0 : row + 1;
        const int lt = col == 0 ? L - 1 : col - 1;
        const int rt = col == L - 1 ? 0 : col + 1;
        int *c = g + (size_t)row * L + col;
        const int nsum = g[(size_t)up * L + col] + g[(size_t)dn * L + col] + g[(size_t)row * L + lt] +
                         g[(size_t)row * L + rt];
        const int idx = 5 * (*c + 1) + nsum + 4;
        const uint64_t u3 = g_mt();
        if (u3 < g_thr[idx]) *c = -*c;
    }
}

// src/mc_cpu.c:91-104: (float)(sum / (double)(L L))
void compute_magnetisation_cpu(int L, int *ising_grids, int grid_index, float *magnetisation) {
    const size_t N = (size_t)L * L;
    const int *g = ising_grids + (size_t)grid_index * N;
    long long s = 0;
    for (size_t i = 0; i < N; ++i) s += g[i];
    magnetisation[grid_index] = (float)((double)s / (double)N);
}

// src/mc_cpu.c:310-364: size of the largest 4-connected periodic cluster of
// sites holding `spin`, as a float (union-find over right and lower bonds).
void compute_largest_cluster_cpu(int L, int *ising_grids, const int grid_index, int spin, float *lclus_size) {
    const int N = L * L;
    const int *g = ising_grids + (size_t)grid_index * N;
    std::vector<int> p(N), cnt(N, 0);
    for (int i = 0; i < N; ++i) p[i] = i;
    for (int row = 0; row < L; ++row)
        for (int col = 0; col < L; ++col) {
            const int i = row * L + col;
            if (g[i] != spin) continue;
            const int nb[2] = {row * L + (col + 1) % L, ((row + 1) % L) * L + col};
            for (int j : nb)
                if (g[j] == spin) {
                    const int a = uf_find(p, i), b = uf_find(p, j);
                    if (a != b) p[std::max(a, b)] = std::min(a, b);
                }
        }
    int best = 0;
    for (int i = 0; i < N; ++i)
        if (g[i] == spin) best = std::max(best, ++cnt[uf_find(p, i)]);
    lclus_size[grid_index] = (float)best;
}

// src/mc_cpu.c:106-308.  Per iteration (isweep a multiple of spc =
// min(mag, grid)): at a check, the collective variable of every replica, the
// report and the stopping test; the output callback at grid_output_int; then
// spc sweeps of every replica in turn from the one global stream.  Errors set
// gasp_last_error() and return where the reference exits.
void mc_driver_cpu(mc_grids_t grids, double beta, double h, int *grid_fate, mc_sampler_t samples,
                   mc_function_t calc, GridOutputFunc outfunc) {
    gasp::clear_error();
    const int L = grids.L, R = grids.ngrids;
    int *spins = grids.ising_grids;
    const int tot = samples.tot_nsweeps, mag = samples.mag_output_int, gout = samples.grid_output_int;
    const int itask = calc.itask, ninputs = calc.ninputs;
    const double dn = calc.dn_thr, up = calc.up_thr;
    if (ninputs < 1 || R % ninputs != 0) {
        set_error(GASP_ERR_ARG, "Error: ngrids must be divisible by ninputs!");
        return;
    }
    if (itask != 0 && itask != 1) {
        set_error(GASP_ERR_ARG, "Error: itask must be 0 or 1!");
        return;
    }
    if (mag < 1 || gout < 1) {
        set_error(GASP_ERR_ARG, "Error: mag_output_int and grid_output_int must be positive");
        return;
    }
    if (L < 1 || R < 1 || !spins || (itask == 1 && !grid_fate) || !calc.result) {
        set_error(GASP_ERR_ARG, "Error: mc_driver_cpu received a NULL array");
        return;
    }
    const bool use_cluster = calc.cv && strcmp(calc.cv, "largest_cluster") == 0;
    const int sub = R / ninputs;
    const int spc = mag < gout ? mag : gout;
    const int result_size = itask == 0 ? tot / mag : ninputs;
    const bool py = gasp::output_style() == 1;

    std::vector<float> magnetisation(R, 0.0f), lclus(use_cluster ? R : 0, 0.0f);
    std::vector<float> result(std::max(result_size, 1), itask == 0 ? 1.0f : 0.0f);
    const float *colvar = use_cluster ? lclus.data() : magnetisation.data();
    const auto t0 = std::chrono::steady_clock::now();

    int isweep = 0;
    while (isweep < tot) {
        if (isweep % mag == 0) {
            for (int r = 0; r < R; ++r) {
                compute_magnetisation_cpu(L, spins, r, magnetisation.data());
                if (use_cluster) compute_largest_cluster_cpu(L, spins, r, -calc.initial_spin, lclus.data());
            }
            if (itask == 0) {
                int nnuc = 0;
                for (int r = 0; r < R; ++r) nnuc += (double)colvar[r] > up;
                if (py)
                    gasp::emit("\r Sweep : %10d, Reached cv = %6.2f : %4d , Unresolved : %4d", isweep, up, nnuc,
                               R - nnuc);
                else
                    gasp::emit("%10d  %12.6f\n", isweep, (double)nnuc / (double)R);
                const int k = isweep / mag;
                if (k < result_size) result[k] = (float)((double)nnuc / (double)R);
                if (nnuc == R) break;
            } else {
                int nA = 0, nB = 0;
                for (int r = 0; r < R; ++r) {
                    if (grid_fate[r] == 0) {
                        ++nA;
                    } else if (grid_fate[r] == 1) {
                        ++nB;
                    } else if ((double)colvar[r] > up) {
                        grid_fate[r] = 1;
                        ++nB;
                    } else if ((double)colvar[r] < dn) {
                        grid_fate[r] = 0;
                        ++nA;
                    }
                }
                gasp::emit("\r Sweep : %10d, Reached cv = %6.2f : %4d , Reached cv = %6.2f : %4d , Unresolved : %4d",
                           isweep, dn, nA, up, nB, R - nA - nB);
                if (nA + nB == R) break;
            }
        }
        if (isweep % gout == 0 && outfunc)
            outfunc(L, R, spins, isweep, magnetisation.data(), use_cluster ? lclus.data() : nullptr, calc.cv, dn, up,
                    calc.filename);
        for (int r = 0; r < R; ++r) mc_sweep_cpu(L, spins, r, beta, h, spc);
        isweep += spc;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    if (itask == 0) {
        for (int i = 0; i < result_size; ++i) calc.result[i] = result[i];
    } else {
        // src/mc_cpu.c:288-302
        for (int ii = 0; ii < ninputs; ++ii) {
            int nB = 0, nF = 0;
            for (int jj = 0; jj < sub; ++jj) {
                const int f = grid_fate[ii * sub + jj];
                if (f > -1) nB += f;
                else nF += 1;
            }
            calc.result[ii] = (float)nB / (float)(sub - nF);
        }
    }
    if (py) {
        gasp::emit("\n");
    } else {
        gasp::emit("\n# Time taken on CPU = %f seconds\n", secs);
        if (itask == 1) gasp::emit("pB estimate : %10.6f\n", (double)calc.result[0]);
    }
}

}  // extern "C"
