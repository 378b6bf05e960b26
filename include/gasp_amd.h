/*
 * gasp_amd.h -- C ABI of the MI355X (gfx950) replica-parallel Metropolis engine.
 *
 * Drop-in boundary for GASP's GPU engine.  The types and the entry points in
 * the first two sections keep the reference's names, argument lists and
 * meaning, so that the reference's callers (the `gasp` CLI main and the PyGASP
 * extension) bind them unchanged; each cites the reference declaration it
 * replaces (paths relative to the reference root).  They are exported with C
 * linkage (the reference's GPU functions had C++ linkage).
 *
 * Differences from the reference, all deliberate:
 *  - `curandState` is replaced by the opaque `gasp_rng_state` (per-replica
 *    counter-based Philox4x32-10 streams, or per-replica MT19937 streams that
 *    reproduce src/mc_cpu.c's random stream bit for bit).
 *  - Errors never call exit(): the function prints the reference's message to
 *    stderr, records an error code readable with gasp_last_error(), and
 *    returns.
 *  - gpuInitDevice returns -1 when no GPU is visible; the caller then falls
 *    back to the CPU engine declared below (mc_driver_cpu and its helpers,
 *    the reference's single global MT19937 stream), as the reference's CLI
 *    and Python module do.
 *
 * All device pointers returned by this library are hipMalloc allocations and
 * may be released with hipFree() or gasp_free().
 */
#ifndef GASP_AMD_H
#define GASP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* =================================================================== types */

/* include/mc_cpu.h:5-9 */
typedef struct {
    int L;            /* size of each L x L grid */
    int ngrids;       /* number of grids (replicas) */
    int *ising_grids; /* host spins, ngrids*L*L ints of +-1, grid-major */
} mc_grids_t;

/* include/mc_cpu.h:11-15 */
typedef struct {
    int tot_nsweeps;     /* maximum number of sweeps */
    int mag_output_int;  /* sweeps between collective-variable checks */
    int grid_output_int; /* sweeps between grid output callbacks */
} mc_sampler_t;

/* include/mc_cpu.h:17-26 */
typedef struct {
    int itask;        /* 0: nucleated fraction vs sweep, 1: committor */
    double dn_thr;    /* cv below which a replica has returned to A */
    double up_thr;    /* cv above which a replica has reached B */
    int ninputs;      /* number of distinct input grids (itask 1) */
    int initial_spin; /* majority spin of the parent phase */
    char *cv;         /* "magnetisation" or "largest_cluster" */
    float *result;    /* tot/mag floats (itask 0) or ninputs floats (itask 1) */
    char *filename;   /* passed through to the output callback */
} mc_function_t;

/* include/mc_cpu.h:31 -- called synchronously on the driver's thread. */
typedef int (*GridOutputFunc)(int L, int ngrids, int *grid_data, int isweep,
                              float *magnetisation, float *lclus_size, char *cv,
                              double dn_thr, double up_thr, char *filename);

/* include/mc_gpu.h:12-18 */
typedef struct {
    int L;
    int ngrids;
    int *ising_grids;      /* host copy, refreshed at every grid output */
    int *d_ising_grids;    /* device int32 spins, ngrids*L*L */
    int *d_neighbour_list; /* kept for ABI compatibility (4*L*L ints) */
} mc_gpu_grids_t;

/* Replaces curandState (include/mc_gpu.h:21).  Opaque device allocation. */
typedef struct gasp_rng_state gasp_rng_state;

/* include/mc_gpu.h:20-24 */
typedef struct {
    gasp_rng_state *d_state; /* from gpuInitRand / gasp_init_rand_mt */
    int threadsPerBlock;     /* validated as in the reference (32 or 64 for
                                method 2); launch shapes are chosen by the
                                engine's tuner */
    int gpu_method;          /* 0: HBM spins, 1: 1-bit-per-spin LDS lattice per
                                replica, 2: 64-replica multispin LDS words */
} gpu_run_t;

/* ==================================================== reference entry points */

/* include/gpu_tools.h:16 (src/gpu_tools.cu:18-99).  Prints the device table,
 * selects deviceIndex (-1 keeps the current device), stores the CU count in
 * *gpu_nsms and returns the device index, or -1 if no GPU is visible. */
int gpuInitDevice(int deviceIndex, int *gpu_nsms);

/* include/gpu_tools.h:19 (src/gpu_tools.cu:143-159).  Allocates the device
 * grids (ngrids*L*L int32) and the neighbour table and copies ising_grids to
 * the device.  The caller frees both. */
void gpuInitGrid(int L, int ngrids, int threadsPerBlock, int *ising_grids,
                 int **d_ising_grids, int **d_neighbour_list);

/* include/gpu_tools.h:22 (src/gpu_tools.cu:162-203).  One independent
 * counter-based stream per replica: Philox4x32-10 keyed by rngseed, counter
 * (attempt pair, replica).  The caller frees *d_state. */
void gpuInitRand(int ngrids, int threadsPerBlock, unsigned long rngseed,
                 gasp_rng_state **d_state);

/* include/mc_gpu.h:28 (src/mc_gpu.cu:21-37).  Builds the single-precision
 * acceptance table expf(-beta*2s(nsum+h)) on the host, exactly as the
 * reference, and converts it to exact integer thresholds for the kernels.
 * Process-global, like the reference's __constant__ table. */
void preComputeProbs_gpu(double beta, double h);

/* include/gpu_tools.h:38 (src/gpu_tools.cu:205-232).  The reference's rule:
 * 2 when L <= maxL = floor(sqrt(32 floor(m^2 / 32))), m = floor(sqrt(8 shmem /
 * threadsPerBlock)), shmem = the device's shared memory per block (64 KB is
 * assumed when no device answers), else 0.  gpu_device -1 means the current
 * device. */
int select_gpu_method(int L, int ngrids, int threadsPerBlock, int gpu_device);

/* include/mc_gpu.h:47 (src/mc_gpu.cu:530-831).  Same schedule, outputs and
 * stdout text as the reference driver: itask 0 fills calc.result[isweep/mag]
 * with the nucleated fraction (pre-filled with 1.0) and stops when all have
 * nucleated; itask 1 fixes grid_fate[] (0: reached A, 1: reached B, -1:
 * unresolved) and writes pB per input grid to calc.result.  outfunc receives
 * the host grids every grid_output_int sweeps. */
void mc_driver_gpu(mc_gpu_grids_t grids, double beta, double h, int *grid_fate,
                   mc_sampler_t samples, mc_function_t calc, gpu_run_t gpu_state,
                   GridOutputFunc outfunc);

/* include/grid.h:5-10 (src/grid.c:5-87).  malloc'd; caller frees.  Return
 * NULL (and set the error) where the reference exits. */
int *init_grids_uniform(int L, int ngrids, int spin);
int *init_grids_from_file(int L, int ngrids);
int *init_grids_from_array(int L, int ngrids, int ninputs, int *array);
int *init_fates(int ngrids);

/* include/io.h:8-11 (src/io.c:4-130).  gridinput.bin reader and the
 * gridstates.bin appender (int32 L, int32 ngrids, int32 isweep, then
 * ceil(L*L*ngrids/8) bytes, bit k set <=> spin k is +1, LSB first). */
void read_input_grid(int L, int ngrids, int *ising_grids);
int write_ising_grids(int L, int ngrids, int *ising_grids, int isweep,
                      float *magnetisation, float *lclus_size, char *cv,
                      double dn_thr, double up_thr, char *filename);

/* include/bootstrap.h:6 (src/bootstrap.c:6-63).  m resamples drawn with libc
 * rand(); returns the population standard deviation of the resampled pB. */
double bootstrap_errbar(int ngrids, int m, int *grid_fate_host);

/* ------------------------------------------------------------ CPU fallback */
/* The reference's CPU engine (include/mc_cpu.h, include/mt19937ar.h), used by
 * the callers when gpuInitDevice returns -1.  One process-global MT19937
 * stream (not thread-safe, as the reference's). */

/* include/mt19937ar.h:47 (src/mt19937ar.c:60-73): seeds the global stream. */
void init_genrand(unsigned long s);
/* include/mt19937ar.h:56 (src/mt19937ar.c:105-140): next 32-bit output. */
unsigned long genrand_int32(void);
/* include/mt19937ar.h:69 (src/mt19937ar.c:163-167): (u + 0.5) / 2^32. */
double genrand_real3(void);
/* include/mc_cpu.h:34 (src/mc_cpu.c:14-38): process-global double table. */
void preComputeProbs_cpu(double beta, double h);
/* include/mc_cpu.h:37 (src/mc_cpu.c:41-89): nsweeps*L*L random-site
 * Metropolis attempts on grid grid_index, three draws per attempt (row,
 * column, acceptance) from the global stream; beta and h are unused (the
 * table holds them). */
void mc_sweep_cpu(int L, int *ising_grids, int grid_index, double beta, double h, int nsweeps);
/* include/mc_cpu.h:40 (src/mc_cpu.c:91-104): magnetisation[grid_index] =
 * (float)(sum / (double)(L*L)). */
void compute_magnetisation_cpu(int L, int *ising_grids, int grid_index, float *magnetisation);
/* include/mc_cpu.h:43 (src/mc_cpu.c:310-364): size of the largest
 * 4-connected periodic cluster of `spin` in grid grid_index, as a float. */
void compute_largest_cluster_cpu(int L, int *ising_grids, const int grid_index, int spin, float *lclus_size);
/* include/mc_cpu.h:46 (src/mc_cpu.c:106-308).  The reference's CPU schedule:
 * at each check the collective variable of every replica, the report (same
 * stdout text) and the stopping test, then outfunc at grid_output_int, then
 * min(mag, grid) sweeps of each replica in turn from the one global stream
 * (seed it with init_genrand, and call preComputeProbs_cpu, first).  Same
 * outputs as mc_driver_gpu.  Errors set gasp_last_error() and return. */
void mc_driver_cpu(mc_grids_t grids, double beta, double h, int *grid_fate, mc_sampler_t samples,
                   mc_function_t calc, GridOutputFunc outfunc);

/* ================================================== engine-specific additions */
/* No counterpart in the reference. */

/* Error reporting.  0 = success. */
enum {
    GASP_OK = 0,
    GASP_ERR_ARG = 1,     /* invalid argument (reference: exit(EXIT_FAILURE)) */
    GASP_ERR_HIP = 2,     /* HIP runtime error */
    GASP_ERR_NOMEM = 3,   /* host or device allocation failed */
    GASP_ERR_IO = 4,      /* file could not be opened or read */
    GASP_ERR_NODEV = 5    /* no GPU visible */
};
int gasp_last_error(void);
const char *gasp_error_string(void);
void gasp_clear_error(void);

/* Per-replica MT19937 streams, replica r seeded (rngseed + r) & 0xffffffff,
 * consumed three draws per attempt (row, col, acceptance) exactly as
 * src/mc_cpu.c:51-77, with the double-precision acceptance table of
 * src/mc_cpu.c:14-38.  Grids then match the reference CPU arithmetic bit for
 * bit.  Returns 0 or an error code. */
int gasp_init_rand_mt(int ngrids, unsigned long rngseed, gasp_rng_state **d_state);

/* Release a device allocation made by this library. */
void gasp_free(void *d_ptr);

/* Engine tuning: slots per replica (number of lanes that speculate on
 * consecutive attempts of one replica; 0 = automatic) and waves per workgroup
 * for methods 0/1 (0 = automatic).  For method 2 the second value is the word
 * split: 1 = a workgroup holds the 64 replicas of a 64-bit site word (slots up
 * to 16), 2 = it holds one 32-bit half (32 replicas, slots up to 32), 4 and
 * 8 = a 16- or 8-bit part (16 or 8 replicas, slots up to 32); 0 = automatic:
 * halves, or smaller parts when that leaves fewer workgroups than CUs or
 * when the lattice would not fit the LDS otherwise.
 * Method 0 also takes 128, 256, 512 or 1024 slots: one replica per workgroup,
 * spread over 2, 4, 8 or 16 waves (the waves value is then implied).  Methods
 * 1 and 2: up to 64.  Method 3 (bit-sliced, below): the first value is the number of
 * waves per workgroup (1, 2, 4 or 8) that evaluate one word group's attempts
 * together, the second the word split (1: 64 replicas per word, 2: 32).
 * Applies to subsequent runs. */
void gasp_set_tuning(int slots_per_replica, int waves_per_group);
void gasp_get_tuning(int method, int L, int ngrids, int *slots_per_replica,
                     int *waves_per_group);

/* Conflict filter of method 2 on split words (half words with 2 to 32 slots;
 * quarter and eighth words with 8 to 32 slots and Philox streams), for
 * engines created afterwards: 1 = row and column masks scanned over the
 * lanes of a replica, 2 = a mark plane beside the lattice in LDS (up to L =
 * 90, 128 and 181 for half, quarter and eighth words; larger lattices keep
 * the masks), 0 (or any other value) = automatic: the mark plane where its
 * LDS leaves as many workgroups resident as the launch fills (on quarter and
 * eighth words only with 32 slots or L > 64).  Trajectories do not depend on
 * it.  No counterpart in the reference (a tuning knob of this engine, like
 * gasp_set_tuning). */
void gasp_set_conflict_filter(int filter);

/* Method 0 with 128 to 1024 slots (one replica over 2 to 16 waves), for
 * engines created afterwards: 1 = the byte lattice stays in HBM (read through
 * L2 every round), 2 = it is staged in LDS for each launch where it fits beside
 * the dependence tables (L up to about 330 at one workgroup per CU), 3 = it is
 * staged in LDS as bits (an eighth of the LDS, one more shift and mask per spin
 * read), 0 (or any other value) = automatic: as bits where that leaves the
 * dependence tables more entries than bytes would (large L), else as bytes
 * where the tables keep min(N, 4096) entries, else in HBM.  Trajectories do
 * not depend on it.  No counterpart in the reference (a tuning
 * knob of this engine, like gasp_set_tuning). */
void gasp_set_lattice_staging(int staging);
/* Method 0 with 128 to 1024 slots, for engines created afterwards: at most
 * max_entries entries (the largest power of two not above it) in each of the
 * two LDS dependence tables of a workgroup, which otherwise take up to 16384
 * where the LDS of a CU allows (0 or less = that default).  Fewer entries leave
 * LDS to other kernels on the same CUs (a batched graph) at the price of more
 * deferred attempts; trajectories do not depend on it.  No counterpart in the
 * reference (a tuning knob of this engine, like gasp_set_tuning). */
void gasp_set_dependence_tables(int max_entries);

/* Statistics of the most recent mc_driver_gpu / gasp_engine run on this
 * thread: sweep-kernel time (HIP events), attempts performed per replica,
 * kernel launches. */
typedef struct {
    double sweep_kernel_ms;
    double attempts_per_replica;
    int launches;
    int slots_per_replica;
    int waves_per_group;
    int method;
} gasp_run_stats_t;
void gasp_get_run_stats(gasp_run_stats_t *out);

/* Stream-level engine (used by benchmarks and by callers that drive the
 * schedule themselves).  The engine owns the replicas' lattice in the
 * method's native layout.  All calls are ordered on the engine's stream.
 *
 * gpu_method 3 (engine-level only; mc_driver_gpu, the CLI and the Python
 * module keep the reference's 0/1/2): bit-sliced multispin.  The replicas of
 * a 64-bit word (or of a 32-bit half, see gasp_set_tuning) share one random
 * site sequence; every attempt updates all of them at once with bitwise
 * adders for the neighbour sums and a bit-sliced comparison of each
 * replica's own 32-bit acceptance draw with its class threshold (the
 * single-precision table).  Each replica follows random-site Metropolis;
 * the replicas of a word are correlated through their common sites.
 * Philox states only (gpuInitRand).  Layout, grids, dumps and checks as
 * method 2.  With freezing, a replica with a decided fate keeps its spins
 * while the rest of its word continues. */
typedef struct gasp_engine gasp_engine;
int gasp_engine_create(int L, int ngrids, int gpu_method, gasp_rng_state *d_state,
                       double beta, double h, gasp_engine **out);
int gasp_engine_destroy(gasp_engine *e);
/* int32 +-1 spins, ngrids*L*L; `on_device` selects device or host memory. */
int gasp_engine_set_grids(gasp_engine *e, const int *grids, int on_device);
/* Host int32 input grids [ninputs][L*L]; replica r starts from input
 * r / (ngrids / ninputs), as init_grids_from_array (src/grid.c:41-71) but
 * replicated on the device. */
int gasp_engine_set_inputs(gasp_engine *e, const int *inputs, int ninputs);
/* 1: replicas whose fate is decided stop sweeping (their state no longer
 * changes); pB and fates are unaffected.  Default 0 (the reference keeps
 * sweeping them).  A freezing method-2 engine raises its slots per replica as
 * replicas are decided (counted at gasp_engine_get_counts and
 * gasp_engine_set_fates) unless gasp_set_tuning fixed the slots; trajectories
 * do not depend on the slots. */
int gasp_engine_set_freeze(gasp_engine *e, int freeze);
int gasp_engine_get_grids(gasp_engine *e, int *grids, int on_device);
/* Grids of the replicas listed in replicas[0..n) as int8 +-1, [n][L*L], in
 * host memory: the int8 snapshot copies of the Python module's grid history
 * (src/ising_ext.cu:280-463) without moving every replica's int32 grid. */
int gasp_engine_gather_grids(gasp_engine *e, const int *replicas, int n, int8_t *host_out);
/* All grids as the body of a gridstates.bin record (write_ising_grids,
 * src/io.c:67-130): ceil(R*L*L / 8) bytes, replica-major, site-major, bit k of
 * byte b = spin 8b + k (1 = up), packed on the device so that only the bits
 * cross to the host. */
int gasp_engine_get_gridstates(gasp_engine *e, uint8_t *host_out);
/* Advance every replica by nsweeps sweeps.  check: 0 none, 1 count replicas
 * with m > up_thr, 2 update fates (sticky; 1 if m > up_thr, 0 if
 * m < dn_thr).  The magnetisation after the sweeps is always computed. */
int gasp_engine_sweep(gasp_engine *e, int nsweeps, int check, double dn_thr,
                      double up_thr);
int gasp_engine_sync(gasp_engine *e);
int gasp_engine_get_magnetisation(gasp_engine *e, float *host_out);
/* counts[0] = #(m > up_thr), counts[1] = #fate 0, counts[2] = #fate 1. */
int gasp_engine_get_counts(gasp_engine *e, int *counts3);
int gasp_engine_set_fates(gasp_engine *e, const int *host_fates);
int gasp_engine_get_fates(gasp_engine *e, int *host_fates);
/* Largest 4-connected periodic cluster of `spin` per replica (floats). */
int gasp_engine_largest_cluster(gasp_engine *e, int spin, float *host_out);
/* Stream-ordered sweeps of caller-owned device grids, the granularity of the
 * reference driver's per-interval launches (src/mc_gpu.cu:675-708:
 * mc_sweep_gpu* then compute_magnetisation_gpu) without an engine object.
 * d_ising_grids (device int32 +-1, ngrids*L*L, as gpuInitGrid makes them)
 * advance by nsweeps sweeps of gpu_method's kernel at (beta, h) with the
 * reference's single-precision table; d_mag (device, ngrids floats, or NULL)
 * receives the magnetisation (float)sum/(float)(L*L).  d_state must come from
 * gpuInitRand (Philox) with at least ngrids streams: the kernels read its key
 * and attempt counter from device memory and a kernel advances the counter, so
 * successive calls continue each replica's chain and nothing is read back to
 * the host (a state of the wrong kind leaves the grids alone and writes NaN
 * magnetisations).  The work is queued on `stream` (a hipStream_t; NULL: the
 * null stream) and not waited for.  Allocates (stream-ordered) a workspace
 * of about the method's native lattice for each call. */
int gasp_sweep_grids(int L, int ngrids, int gpu_method, gasp_rng_state *d_state, double beta, double h,
                     int *d_ising_grids, int nsweeps, float *d_mag, void *stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  gasp_sweep_grids with the same arguments and results,
 * but its workspace is a 16 MiB device array inside the library (one per
 * device, held until the process ends) instead of a per-call allocation; it
 * calls no hipMalloc/hipFree.  Sizes whose workspace exceeds 16 MiB
 * (thresholds and counters, the magnetisation when d_mag is NULL, and the
 * method's native lattice: e.g. method 2 at L = 64 up to about 32000
 * replicas, method 0 at L = 64 up to 4095) are refused with GASP_ERR_ARG.  All
 * calls on one device share that array: calls that overlap in time on one
 * device (other streams or threads) race and must be serialised by the
 * caller. */
int gasp_sweep_grids_noalloc(int L, int ngrids, int gpu_method, gasp_rng_state *d_state, double beta, double h,
                             int *d_ising_grids, int nsweeps, float *d_mag, void *stream);

/* Milliseconds spent in sweep kernels on the engine's stream so far. */
double gasp_engine_kernel_ms(gasp_engine *e);
/* The engine's HIP stream (hipStream_t). */
void *gasp_engine_stream(gasp_engine *e);
/* The engine's launch shape: slots per replica and waves per workgroup (the
 * word split for method 2), as gasp_set_tuning takes them. */
int gasp_engine_get_shape(gasp_engine *e, int *slots, int *waves);
/* Moves the engine's replicas to per-replica attempt `attempt` of their
 * Philox streams (the next sweep uses attempts attempt, attempt + 1, ...):
 * rewinding replays a trajectory's random numbers.  Waits for the engine's
 * launches.  Counter-based streams only (MT19937 states are refused).  No
 * counterpart in the reference (curand states do not rewind). */
int gasp_engine_set_attempt(gasp_engine *e, unsigned long long attempt);
/* The engine's conflict filter, as gasp_set_conflict_filter takes it (1 for
 * engines without the choice). */
int gasp_engine_get_conflict_filter(gasp_engine *e, int *filter);
/* The engine's lattice staging, as gasp_set_lattice_staging takes it (1 for
 * engines without the choice). */
int gasp_engine_get_lattice_staging(gasp_engine *e, int *staging);

/* ------------------------------------------------ batched graph, autotuner */
/* One captured hipGraph advancing n engines concurrently (uneven L, methods,
 * temperatures and tasks); each replay runs nsweeps[i] sweeps of engine i in
 * launches[i] kernels (NULL: 1) with threshold check checks[i] at the end.
 * Counter-based (Philox) engines only.  Engines must outlive the graph; their
 * getters are valid after gasp_graph_launch returns. */
typedef struct gasp_graph gasp_graph;
int gasp_graph_create(gasp_engine **engines, int n, const int *nsweeps, const int *launches,
                      const int *checks, const double *dn, const double *up, gasp_graph **out);
int gasp_graph_launch(gasp_graph *g, int replays); /* runs and waits */
double gasp_graph_ms(gasp_graph *g);               /* GPU time of all replays */
int gasp_graph_destroy(gasp_graph *g);
/* bootstrap_errbar (src/bootstrap.c:6-63) for `ninputs` inputs of n fates each
 * (int32 [ninputs][n], -1/0/1; host memory, or device memory if on_device)
 * on the GPU: m resamples per input, the same estimator with Philox4x32-10
 * draws (key `seed`) instead of the process-global libc rand() sequence, so
 * the error bars agree statistically, not bit for bit (the default
 * bootstrap_errbar keeps the reference's sequence).  host_out[ninputs]. */
int gasp_bootstrap_errbar_gpu(const int *fates, int on_device, int ninputs, int n, int m,
                              unsigned long long seed, double *host_out);
/* bootstrap_errbar (src/bootstrap.c:6-63) for `ninputs` inputs of n fates each
 * (host int32 [ninputs][n]), bit for bit as `ninputs` successive calls of
 * bootstrap_errbar(n, m, ...) after srand(seed) with glibc's rand() (seed 1 is
 * glibc's unseeded state): the library generates glibc's rand() sequence
 * itself and jumps it ahead to each input's first draw, so the inputs run on
 * `nthreads` host threads (0: one per core; at most 256, and no more than
 * ninputs).  Does not touch the process's
 * own rand() state.  out[ninputs]. */
int gasp_bootstrap_errbar_libc(const int *fates, int ninputs, int n, int m, unsigned seed, int nthreads,
                               double *out);
/* Times the candidate launch shapes (slots per replica x waves per workgroup,
 * or word split for method 2) for (method, L, ngrids) on probe_sweeps sweeps
 * at (beta, h) from all spins down, and keeps the fastest for later engines
 * of that (method, L, ngrids).  With check_sweeps > 0 it then times that
 * interval as 1, 2, 4, ... equal launches (of at least check_sweeps / 8
 * sweeps) and returns the fastest sweeps per launch in *best_launch_sweeps
 * (check_sweeps when 0 or 1).  *best_rate: attempts/s of the chosen shape. */
int gasp_autotune(int method, int L, int ngrids, double beta, double h, int probe_sweeps, int check_sweeps,
                  int *best_slots, int *best_waves, int *best_launch_sweeps, double *best_rate);

/* ------------------------------------------------ committor surrogate (CNN) */
/* No counterpart in the reference: the learned addition.  A 4-layer circular
 * 3x3 CNN (32 channels) + global average pool + MLP (64) + sigmoid mapping an
 * L x L spin grid (L a power of two >= 16) to pB, trained with binary
 * cross-entropy on shooting labels.  BF16 MFMA kernels with FP32
 * accumulation, FP32 master weights, Adam.  Spins are int8 +-1 [n][L][L];
 * parameters a float32 vector of gasp_cnn_num_params() entries. */
typedef struct gasp_cnn gasp_cnn;
int gasp_cnn_num_params(void);
int gasp_cnn_create(int L, int max_batch, gasp_cnn **out);
int gasp_cnn_destroy(gasp_cnn *c);
int gasp_cnn_set_params(gasp_cnn *c, const float *host_params); /* resets Adam */
int gasp_cnn_get_params(gasp_cnn *c, float *host_params);
int gasp_cnn_predict(gasp_cnn *c, const int8_t *spins, int n, float *pB);
/* gradient of the mean loss over n <= max_batch grids (no update) */
int gasp_cnn_gradient(gasp_cnn *c, const int8_t *spins, const float *labels, int n, float *grad, float *loss);
int gasp_cnn_train_step(gasp_cnn *c, const int8_t *spins, const float *labels, int n, float lr, float *loss);
/* device-resident training set, then epochs over it in a caller-given order */
int gasp_cnn_set_data(gasp_cnn *c, const int8_t *spins, const float *labels, int count);
int gasp_cnn_train_epoch(gasp_cnn *c, const int *order, int count, int batch, float lr, float *mean_loss);

#ifdef __cplusplus
}
#endif
#endif /* GASP_AMD_H */
