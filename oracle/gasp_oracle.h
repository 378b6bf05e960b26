/*
 * gasp_oracle.h -- CPU oracle for the replica-parallel Metropolis engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (gpu_2dising_amd/) never links, loads or calls anything here.
 *
 * Every function restates one piece of the reference arithmetic and cites
 * the reference file:line it follows (paths relative to the reference root).
 * Pinning (what each function is checked against) is listed beside it; the
 * checks live in tests/test_oracle_*.py.
 */
#ifndef GASP_ORACLE_H
#define GASP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- MT19937 */
/* Per-stream state.  The reference keeps ONE process-global state
 * (src/mt19937ar.c:56-57); the oracle keeps it in a struct so that each
 * replica can own a stream (per-replica seeds base_seed + r).
 * Pinned: bit-exact vs oracle/_ref (compiled src/mt19937ar.c) and vs
 * libstdc++ std::mt19937 (C++ standard: 10000th output from 5489 is
 * 4123659995). */
typedef struct {
    uint32_t mt[624];
    int mti;
} orc_mt_t;

void     orc_mt_seed(orc_mt_t *s, uint32_t seed);   /* src/mt19937ar.c:60-73   */
uint32_t orc_mt_u32(orc_mt_t *s);                   /* src/mt19937ar.c:105-140 */
double   orc_mt_real3(orc_mt_t *s);                 /* src/mt19937ar.c:163-167 */
/* Bulk generation of the next n outputs (used by the draw-stream fixtures). */
void     orc_mt_fill(orc_mt_t *s, uint32_t *out, long long n);

/* ------------------------------------------------------ acceptance tables */
/* idx = 5*(s+1) + nsum + 4 for s in {-1,+1}, nsum in {-4,-2,0,2,4}.       */
void orc_pacc_cpu(double beta, double h, double pacc[20]); /* src/mc_cpu.c:14-38   (double, exp)  */
void orc_pacc_gpu(double beta, double h, float pacc[20]);  /* src/mc_gpu.cu:21-37  (float, expf)  */

/* Exact integer form of "xi < P" for xi = (u + 0.5) * 2^-32 (genrand_real3,
 * src/mt19937ar.c:165): returns T in [0, 2^32] with  xi < P  <=>  u < T.
 * Pinned: brute force around T with the literal double comparison. */
uint64_t orc_threshold(double p);

/* -------------------------------------------------------- Philox4x32-10 */
/* Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11.  Pinned: Random123 known-answer vectors and the rocRAND header
 * implementation (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ----------------------------------------------------- single-grid sweeps */
/* Literal restatement of mc_sweep_cpu (src/mc_cpu.c:41-89): per attempt
 * row = floor(L*real3), col = floor(L*real3), flip, xi = real3, keep the flip
 * iff xi < Pacc[idx].  Runs `nattempts` attempts (nsweeps*L*L in the
 * reference) on the int32 grid `grid` (L*L spins, +-1). */
void orc_sweep_mt(int L, int32_t *grid, orc_mt_t *rng, const double pacc[20],
                  long long nattempts);

/* Production stream of the engine (no counterpart in the reference, whose GPU
 * stream is cuRAND XORWOW): attempt a of replica r draws
 *   w = Philox4x32-10(ctr = {p_lo, p_hi, r, 0x47415350}, key = {seed_lo, seed_hi}),
 *   p = a >> 1, (ws, wa) = (w[2*(a&1)], w[2*(a&1)+1]),
 *   site = (ws * L*L) >> 32, accept iff wa < thr[idx].
 * Dynamics per accepted attempt exactly as src/mc_gpu.cu:104-130. */
void orc_sweep_philox(int L, int32_t *grid, uint64_t seed, uint32_t replica,
                      uint64_t attempt0, long long nattempts, const uint64_t thr[20]);

/* ------------------------------------------------------ collective vars */
float orc_magnetisation_cpu(int L, const int32_t *grid); /* src/mc_cpu.c:91-104   */
float orc_magnetisation_gpu(int L, const int32_t *grid); /* src/mc_gpu.cu:437-455 */
float orc_largest_cluster(int L, const int32_t *grid, int spin); /* src/mc_cpu.c:310-364 */

/* ------------------------------------------------------------ bootstrap */
/* src/bootstrap.c:6-63.  Uses libc rand() exactly as the reference does;
 * callers control the libc state with srand().  Pinned: bit-exact vs
 * oracle/_ref (compiled src/bootstrap.c) after the same srand(). */
double orc_bootstrap_errbar(int ngrids, int m, const int32_t *fates);

/* --------------------------------------------------------- grid formats */
/* gridstates.bin record body (src/io.c:86-115): bit k set <=> spin k is +1,
 * LSB-first, packed continuously over all grids.  out must hold
 * ceil(n/8) bytes. */
void orc_pack_bits(const int32_t *spins, long long n, uint8_t *out);
void orc_unpack_bits(const uint8_t *bits, long long n, int32_t *spins);

/* ------------------------------------------------------- engine driver */
/* Restatement of the mc_driver_gpu schedule (src/mc_gpu.cu:530-831) with CPU
 * sweeps in place of the kernels.  rng_kind 0: per-replica MT19937 seeded
 * seed + r, double table (src/mc_cpu.c arithmetic); rng_kind 1: production
 * Philox stream with the FP32 table.  `grids` holds R*L*L int32 spins and is
 * advanced in place.  cv_kind 0 thresholds the magnetisation, 1 the largest
 * cluster of spin -initial_spin (src/mc_gpu.cu:604-607,670).
 * `result` gets tot/mag floats (itask 0) or ninputs floats (itask 1);
 * `fates` (itask 1) is R ints, -1 on entry.  `cv_out`, if not NULL, receives
 * one row of R floats (the thresholded CV) per check, up to max_checks rows;
 * `check_sweeps` gets the sweep number of each row.  The number of checks
 * performed is returned (-1 on a bad argument).
 * `dumps_out`, if not NULL, receives the host grid copy handed to the output
 * callback at each call (R*L*L int32 per call, up to max_dumps); the sweep
 * numbers go to dump_sweeps and their count to *ndumps. */
int orc_driver(int L, int R, int32_t *grids, int rng_kind, uint64_t seed,
               double beta, double h, int tot_nsweeps, int mag_output_int,
               int grid_output_int, int itask, int cv_kind, int initial_spin,
               double dn_thr, double up_thr,
               int ninputs, int32_t *fates, float *result,
               float *cv_out, int32_t *check_sweeps, int max_checks,
               int32_t *dumps_out, int32_t *dump_sweeps, int max_dumps, int *ndumps,
               int nthreads);

/* ------------------------------------------------------- CPU baseline */
/* Advance R independent replicas by `nattempts` attempts each, replicas
 * partitioned over `nthreads` POSIX threads (no shared state).  rng_kind 0:
 * per-replica MT19937 states in `mt` (R entries, advanced in place) with the
 * double table; rng_kind 1: Philox stream starting at attempt0.  Returns wall
 * seconds (CLOCK_MONOTONIC). */
double orc_sweep_batch(int L, int R, int32_t *grids, int rng_kind, orc_mt_t *mt,
                       uint64_t seed, uint64_t attempt0, long long nattempts,
                       double beta, double h, int nthreads);

/* The engine's bit-sliced multispin stream (gpu_method 3): one group of nrep
 * replicas (wb = 64 or 32 per group, group index `group`) sharing a site
 * sequence, each replica's acceptance draw assembled from random bit planes;
 * see gasp_oracle.c.  frozen: nrep flags (or NULL), frozen replicas do not
 * change.  The batch form runs every group of R replicas on threads with the
 * single-precision table at (beta, h); frozen has R flags or is NULL. */
#define ORC_SLICED_DOMAIN 0x534C4300u
void orc_sweep_sliced(int L, int32_t *g0, int nrep, int wb, uint64_t seed, uint32_t group,
                      uint64_t attempt0, long long nattempts, const uint64_t thr[20],
                      const uint8_t *frozen);
double orc_sweep_sliced_batch(int L, int R, int32_t *grids, int wb, uint64_t seed, uint64_t attempt0,
                              long long nattempts, double beta, double h, const uint8_t *frozen, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
