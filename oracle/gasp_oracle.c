/*
 * gasp_oracle.c -- CPU oracle for the replica-parallel Metropolis engine.
 *
 * TEST INFRASTRUCTURE ONLY (see gasp_oracle.h): the checker for the HIP
 * engine in gpu_2dising_amd/, never part of it.  Plain C99 + POSIX threads.
 */
#define _POSIX_C_SOURCE 200809L
#include "gasp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ================================================================ MT19937 */
/* Matsumoto & Nishimura, ACM TOMACS 8 (1998) 3-30; seeding as the 2002
 * revision.  Restates src/mt19937ar.c:49-73 (parameters, init_genrand),
 * :105-140 (genrand_int32), :163-167 (genrand_real3). */
#define MT_N 624
#define MT_M 397
#define MT_A 0x9908b0dfu
#define MT_UPPER 0x80000000u
#define MT_LOWER 0x7fffffffu

void orc_mt_seed(orc_mt_t *s, uint32_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < MT_N; i++) {
        uint32_t prev = s->mt[i - 1];
        s->mt[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
    }
    s->mti = MT_N;
}

static void mt_twist(orc_mt_t *s)
{
    uint32_t *mt = s->mt;
    int k;
    for (k = 0; k < MT_N; k++) {
        uint32_t y = (mt[k] & MT_UPPER) | (mt[(k + 1) % MT_N] & MT_LOWER);
        mt[k] = mt[(k + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1u) ? MT_A : 0u);
    }
    s->mti = 0;
}

uint32_t orc_mt_u32(orc_mt_t *s)
{
    if (s->mti >= MT_N) mt_twist(s);
    uint32_t y = s->mt[s->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

double orc_mt_real3(orc_mt_t *s)
{
    return (((double)orc_mt_u32(s)) + 0.5) * (1.0 / 4294967296.0);
}

void orc_mt_fill(orc_mt_t *s, uint32_t *out, long long n)
{
    for (long long i = 0; i < n; i++) out[i] = orc_mt_u32(s);
}

/* ====================================================== acceptance tables */
void orc_pacc_cpu(double beta, double h, double pacc[20])
{
    /* src/mc_cpu.c:28-35 */
    memset(pacc, 0, 20 * sizeof(double));
    for (int s = -1; s < 2; s += 2)
        for (int nsum = -4; nsum < 5; nsum += 2) {
            int idx = 5 * (s + 1) + nsum + 4;
            double de = 2.0 * (double)s * ((double)nsum + h);
            pacc[idx] = exp(-beta * de);
        }
}

void orc_pacc_gpu(double beta, double h, float pacc[20])
{
    /* src/mc_gpu.cu:25-31: evaluated on the host in single precision */
    memset(pacc, 0, 20 * sizeof(float));
    for (int s = -1; s < 2; s += 2)
        for (int nsum = -4; nsum < 5; nsum += 2) {
            int idx = 5 * (s + 1) + nsum + 4;
            float de = 2.0f * (float)s * ((float)nsum + (float)h);
            pacc[idx] = expf(-(float)beta * de);
        }
}

uint64_t orc_threshold(double p)
{
    /* xi = (u + 0.5) 2^-32 is exact in double; p 2^32 - 0.5 is exact for
     * p < 1, so  xi < p  <=>  u < p 2^32 - 0.5  <=>  u < ceil(p 2^32 - 0.5). */
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return 4294967296ull;
    double y = ceil(p * 4294967296.0 - 0.5);
    if (y <= 0.0) return 0;
    return (uint64_t)y;
}

/* ========================================================== Philox4x32-10 */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ======================================================= single-grid sweeps */
void orc_sweep_mt(int L, int32_t *g, orc_mt_t *rng, const double pacc[20],
                  long long nattempts)
{
    /* src/mc_cpu.c:47-86 */
    for (long long imove = 0; imove < nattempts; imove++) {
        int row = (int)floor(L * orc_mt_real3(rng));
        int col = (int)floor(L * orc_mt_real3(rng));
        int my = L * row + col;
        int up = L * ((row + L - 1) % L) + col;
        int dn = L * ((row + 1) % L) + col;
        int rt = L * row + (col + 1) % L;
        int lt = L * row + (col + L - 1) % L;
        int nsum = g[up] + g[dn] + g[lt] + g[rt];
        int idx = 5 * (g[my] + 1) + nsum + 4;
        g[my] = -g[my];
        double xi = orc_mt_real3(rng);
        if (!(xi < pacc[idx])) g[my] = -g[my];
    }
}

#define ORC_PHILOX_DOMAIN 0x47415350u

void orc_sweep_philox(int L, int32_t *g, uint64_t seed, uint32_t replica,
                      uint64_t attempt0, long long nattempts, const uint64_t thr[20])
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint64_t N = (uint64_t)L * (uint64_t)L;
    uint32_t w[4];
    uint64_t have = ~(uint64_t)0;
    for (long long i = 0; i < nattempts; i++) {
        uint64_t a = attempt0 + (uint64_t)i;
        uint64_t p = a >> 1;
        if (p != have) {
            uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), replica, ORC_PHILOX_DOMAIN};
            orc_philox4x32_10(ctr, key, w);
            have = p;
        }
        uint32_t ws = w[2 * (a & 1)], wa = w[2 * (a & 1) + 1];
        int site = (int)(((uint64_t)ws * N) >> 32);
        int row = site / L, col = site % L;
        int up = L * ((row + L - 1) % L) + col;
        int dn = L * ((row + 1) % L) + col;
        int rt = L * row + (col + 1) % L;
        int lt = L * row + (col + L - 1) % L;
        int nsum = g[up] + g[dn] + g[lt] + g[rt];
        int idx = 5 * (g[site] + 1) + nsum + 4;
        if ((uint64_t)wa < thr[idx]) g[site] = -g[site];
    }
}

/* ====================================================== bit-sliced stream */
/* The engine's bit-sliced multispin mode (gpu_method 3, gasp_sliced.h),
 * restated per replica.  Group `group` holds replicas wb*group + b,
 * b = 0 .. nrep-1 (wb = 64 or 32), grids g0 + b*L*L.  Attempt a of the group:
 * Philox4x32-10 block k = philox(ctr = (a_lo, a_hi, group, ORC_SLICED_DOMAIN
 * + k), key = seed) holds words 4k .. 4k+3 of the attempt's word stream; the
 * site of every replica is floor(w_0 N / 2^32) (the production stream's site
 * map); bit 31 - p of replica b's 32-bit acceptance draw u is bit (b & 31)
 * of word (wb/32)(p + 1) + b/32 (plane p); replica b flips iff u < T of its
 * class, the single-precision table's threshold at the reference index
 * 5(s+1)+nsum+4 (src/mc_cpu.c:55-84 arithmetic; T >= 2^32 always flips).
 * Replicas with frozen[b] set do not change.  The whole draw is assembled for
 * every replica; the GPU stops at the first differing bit, which decides the
 * same comparison. */
void orc_sweep_sliced(int L, int32_t *g0, int nrep, int wb, uint64_t seed, uint32_t group,
                      uint64_t attempt0, long long nattempts, const uint64_t thr[20],
                      const uint8_t *frozen)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint64_t N = (uint64_t)L * (uint64_t)L;
    const int q = wb / 32;
    uint32_t blk[17][4];
    for (long long i = 0; i < nattempts; i++) {
        uint64_t a = attempt0 + (uint64_t)i;
        int nblk = 0;
        for (; nblk < 17; nblk++) {
            uint32_t ctr[4] = {(uint32_t)a, (uint32_t)(a >> 32), group, ORC_SLICED_DOMAIN + (uint32_t)nblk};
            orc_philox4x32_10(ctr, key, blk[nblk]);
        }
        int site = (int)(((uint64_t)blk[0][0] * N) >> 32);
        int row = site / L, col = site % L;
        int up = L * ((row + L - 1) % L) + col;
        int dn = L * ((row + 1) % L) + col;
        int rt = L * row + (col + 1) % L;
        int lt = L * row + (col + L - 1) % L;
        for (int b = 0; b < nrep; b++) {
            if (frozen && frozen[b]) continue;
            int32_t *g = g0 + (long long)b * (long long)N;
            int nsum = g[up] + g[dn] + g[lt] + g[rt];
            int idx = 5 * (g[site] + 1) + nsum + 4;
            int flip;
            if (thr[idx] >= ((uint64_t)1 << 32)) {
                flip = 1;
            } else {
                uint32_t u = 0;
                for (int p = 0; p < 32; p++) {
                    int word = q * (p + 1) + b / 32;
                    uint32_t bit = (blk[word >> 2][word & 3] >> (b & 31)) & 1u;
                    u |= bit << (31 - p);
                }
                flip = (uint64_t)u < thr[idx];
            }
            if (flip) g[site] = -g[site];
        }
    }
}

typedef struct {
    int L, R, wb, g0, g1;
    int32_t *grids;
    uint64_t seed, attempt0;
    long long nattempts;
    const uint64_t *thr;
    const uint8_t *frozen;
} sliced_job_t;

static void *sliced_worker(void *arg)
{
    sliced_job_t *j = (sliced_job_t *)arg;
    long long N = (long long)j->L * j->L;
    for (int g = j->g0; g < j->g1; g++) {
        int r0 = g * j->wb;
        int nrep = j->R - r0 < j->wb ? j->R - r0 : j->wb;
        orc_sweep_sliced(j->L, j->grids + (long long)r0 * N, nrep, j->wb, j->seed, (uint32_t)g, j->attempt0,
                         j->nattempts, j->thr, j->frozen ? j->frozen + r0 : NULL);
    }
    return NULL;
}

double orc_sweep_sliced_batch(int L, int R, int32_t *grids, int wb, uint64_t seed, uint64_t attempt0,
                              long long nattempts, double beta, double h, const uint8_t *frozen, int nthreads)
{
    float paccf[20];
    uint64_t thr[20];
    orc_pacc_gpu(beta, h, paccf);
    for (int i = 0; i < 20; i++) thr[i] = orc_threshold((double)paccf[i]);
    int ngroups = (R + wb - 1) / wb;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ngroups) nthreads = ngroups;
    pthread_t tid[256];
    sliced_job_t jobs[256];
    if (nthreads > 256) nthreads = 256;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < nthreads; t++) {
        sliced_job_t *j = &jobs[t];
        j->L = L;
        j->R = R;
        j->wb = wb;
        j->g0 = (int)((long long)ngroups * t / nthreads);
        j->g1 = (int)((long long)ngroups * (t + 1) / nthreads);
        j->grids = grids;
        j->seed = seed;
        j->attempt0 = attempt0;
        j->nattempts = nattempts;
        j->thr = thr;
        j->frozen = frozen;
    }
    if (nthreads == 1) {
        sliced_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; t++) pthread_create(&tid[t], NULL, sliced_worker, &jobs[t]);
        for (int t = 0; t < nthreads; t++) pthread_join(tid[t], NULL);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ========================================================= collective vars */
float orc_magnetisation_cpu(int L, const int32_t *g)
{
    /* src/mc_cpu.c:96-100 */
    double m = 0.0;
    for (int i = 0; i < L * L; i++) m += (double)g[i];
    return (float)(m / (double)(L * L));
}

float orc_magnetisation_gpu(int L, const int32_t *g)
{
    /* src/mc_gpu.cu:445-449: float accumulation (exact while L*L <= 2^24) */
    float m = 0.0f;
    for (int i = 0; i < L * L; i++) m += (float)g[i];
    return m / (float)(L * L);
}

float orc_largest_cluster(int L, const int32_t *grid, int spin)
{
    /* src/mc_cpu.c:310-364: BFS over 4-connected periodic neighbours */
    int N = L * L;
    char *visited = (char *)calloc((size_t)N, 1);
    int *queue = (int *)malloc((size_t)N * sizeof(int));
    const int dx[4] = {-1, 1, 0, 0};
    const int dy[4] = {0, 0, -1, 1};
    int max_size = 0;
    for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x) {
            int idx = y * L + x;
            if (grid[idx] != spin || visited[idx]) continue;
            visited[idx] = 1;
            int front = 0, back = 0, size = 0;
            queue[back++] = idx;
            while (front < back) {
                int cur = queue[front++];
                size++;
                int cx = cur % L, cy = cur / L;
                for (int d = 0; d < 4; ++d) {
                    int nx = (cx + dx[d] + L) % L, ny = (cy + dy[d] + L) % L;
                    int n = ny * L + nx;
                    if (grid[n] == spin && !visited[n]) {
                        visited[n] = 1;
                        queue[back++] = n;
                    }
                }
            }
            if (size > max_size) max_size = size;
        }
    free(visited);
    free(queue);
    return (float)max_size;
}

/* =============================================================== bootstrap */
double orc_bootstrap_errbar(int ngrids, int m, const int32_t *fates)
{
    /* src/bootstrap.c:24-57: indices drawn first, then counted */
    int *ind = (int *)malloc((size_t)ngrids * sizeof(int));
    double *store = (double *)malloc((size_t)m * sizeof(double));
    if (!ind || !store) { free(ind); free(store); return -1.0; }
    for (int im = 0; im < m; im++) {
        int nA = 0, nB = 0;
        for (int i = 0; i < ngrids; i++) ind[i] = rand() % ngrids;
        for (int i = 0; i < ngrids; i++) {
            int f = fates[ind[i]];
            if (f == 0) nA++;
            else if (f == 1) nB++;
        }
        store[im] = (double)nB / (double)(nA + nB);
    }
    double mean = 0.0, num = 0.0;
    for (int im = 0; im < m; im++) mean = mean + store[im];
    mean = mean / m;
    for (int im = 0; im < m; im++) num = num + (store[im] - mean) * (store[im] - mean);
    free(ind);
    free(store);
    return sqrt(num / m);
}

/* ============================================================ grid formats */
void orc_pack_bits(const int32_t *spins, long long n, uint8_t *out)
{
    long long nbytes = (n + 7) / 8;
    memset(out, 0, (size_t)nbytes);
    for (long long i = 0; i < n; i++)
        if (spins[i] == 1) out[i >> 3] |= (uint8_t)(1u << (i & 7));
}

void orc_unpack_bits(const uint8_t *bits, long long n, int32_t *spins)
{
    for (long long i = 0; i < n; i++) spins[i] = ((bits[i >> 3] >> (i & 7)) & 1) ? 1 : -1;
}

/* ===================================================== threaded replica loop */
typedef struct {
    int L, r0, r1, rng_kind;
    int32_t *grids;
    orc_mt_t *mt;
    uint64_t seed, attempt0;
    long long nattempts;
    const double *pacc;
    const uint64_t *thr;
} sweep_job_t;

static void *sweep_worker(void *arg)
{
    sweep_job_t *j = (sweep_job_t *)arg;
    long long N = (long long)j->L * j->L;
    for (int r = j->r0; r < j->r1; r++) {
        int32_t *g = j->grids + (long long)r * N;
        if (j->rng_kind == 0)
            orc_sweep_mt(j->L, g, &j->mt[r], j->pacc, j->nattempts);
        else
            orc_sweep_philox(j->L, g, j->seed, (uint32_t)r, j->attempt0, j->nattempts, j->thr);
    }
    return NULL;
}

static void sweep_all(int L, int R, int32_t *grids, int rng_kind, orc_mt_t *mt,
                      uint64_t seed, uint64_t attempt0, long long nattempts,
                      const double *pacc, const uint64_t *thr, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R > 0 ? R : 1;
    pthread_t *tid = (pthread_t *)malloc((size_t)nthreads * sizeof(pthread_t));
    sweep_job_t *jobs = (sweep_job_t *)malloc((size_t)nthreads * sizeof(sweep_job_t));
    for (int t = 0; t < nthreads; t++) {
        sweep_job_t *j = &jobs[t];
        j->L = L;
        j->r0 = (int)((long long)R * t / nthreads);
        j->r1 = (int)((long long)R * (t + 1) / nthreads);
        j->rng_kind = rng_kind;
        j->grids = grids;
        j->mt = mt;
        j->seed = seed;
        j->attempt0 = attempt0;
        j->nattempts = nattempts;
        j->pacc = pacc;
        j->thr = thr;
    }
    if (nthreads == 1) {
        sweep_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; t++) pthread_create(&tid[t], NULL, sweep_worker, &jobs[t]);
        for (int t = 0; t < nthreads; t++) pthread_join(tid[t], NULL);
    }
    free(tid);
    free(jobs);
}

double orc_sweep_batch(int L, int R, int32_t *grids, int rng_kind, orc_mt_t *mt,
                       uint64_t seed, uint64_t attempt0, long long nattempts,
                       double beta, double h, int nthreads)
{
    double pacc[20];
    float paccf[20];
    uint64_t thr[20];
    orc_pacc_cpu(beta, h, pacc);
    orc_pacc_gpu(beta, h, paccf);
    for (int i = 0; i < 20; i++) thr[i] = orc_threshold((double)paccf[i]);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    sweep_all(L, R, grids, rng_kind, mt, seed, attempt0, nattempts, pacc, thr, nthreads);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ================================================================= driver */
int orc_driver(int L, int R, int32_t *grids, int rng_kind, uint64_t seed,
               double beta, double h, int tot_nsweeps, int mag_output_int,
               int grid_output_int, int itask, int cv_kind, int initial_spin,
               double dn_thr, double up_thr,
               int ninputs, int32_t *fates, float *result,
               float *cv_out, int32_t *check_sweeps, int max_checks,
               int32_t *dumps_out, int32_t *dump_sweeps, int max_dumps, int *ndumps,
               int nthreads)
{
    if (L < 1 || R < 1 || ninputs < 1 || R % ninputs != 0) return -1;
    if (mag_output_int < 1 || grid_output_int < 1) return -1;
    if (itask != 0 && itask != 1) return -1;
    const long long N = (long long)L * L;
    const int sub = R / ninputs;
    /* src/mc_gpu.cu:625-626 */
    const int spc = mag_output_int < grid_output_int ? mag_output_int : grid_output_int;
    const int result_size = itask == 0 ? tot_nsweeps / mag_output_int : ninputs;

    double pacc[20];
    float paccf[20];
    uint64_t thr[20];
    orc_pacc_cpu(beta, h, pacc);
    orc_pacc_gpu(beta, h, paccf);
    for (int i = 0; i < 20; i++) thr[i] = orc_threshold((double)paccf[i]);

    orc_mt_t *mt = NULL;
    if (rng_kind == 0) {
        mt = (orc_mt_t *)malloc((size_t)R * sizeof(orc_mt_t));
        for (int r = 0; r < R; r++) orc_mt_seed(&mt[r], (uint32_t)(seed + (uint64_t)r));
    }
    float *colvar = (float *)malloc((size_t)R * sizeof(float));
    int32_t *host = (int32_t *)malloc((size_t)(R * N) * sizeof(int32_t));
    float *res = (float *)malloc((size_t)(result_size > 0 ? result_size : 1) * sizeof(float));
    for (int i = 0; i < R; i++) colvar[i] = 0.0f;
    memcpy(host, grids, (size_t)(R * N) * sizeof(int32_t));
    if (itask == 0)
        for (int i = 0; i < result_size; i++) res[i] = 1.0f; /* src/mc_gpu.cu:645-650 */

    int nchecks = 0, nd = 0;
    int isweep = 0;
    while (isweep < tot_nsweeps) {
        /* src/mc_gpu.cu:658-661: host copy of the grids before this launch */
        if (isweep % grid_output_int == 0)
            memcpy(host, grids, (size_t)(R * N) * sizeof(int32_t));
        /* src/mc_gpu.cu:664-675: collective variable before this launch */
        if (isweep % mag_output_int == 0) {
            for (int r = 0; r < R; r++) {
                const int32_t *g = grids + (long long)r * N;
                colvar[r] = cv_kind == 1 ? orc_largest_cluster(L, g, -initial_spin)
                                         : orc_magnetisation_gpu(L, g);
            }
            if (cv_out && nchecks < max_checks) {
                memcpy(cv_out + (long long)nchecks * R, colvar, (size_t)R * sizeof(float));
                if (check_sweeps) check_sweeps[nchecks] = isweep;
            }
            nchecks++;
        }
        /* src/mc_gpu.cu:679-706: the sweep launch (spc sweeps) */
        sweep_all(L, R, grids, rng_kind, mt, seed, (uint64_t)isweep * (uint64_t)N,
                  (long long)spc * N, pacc, thr, nthreads);
        /* src/mc_gpu.cu:709-711: output callback */
        if (isweep % grid_output_int == 0 || isweep == tot_nsweeps - 1) {
            if (dumps_out && nd < max_dumps) {
                memcpy(dumps_out + (long long)nd * R * N, host, (size_t)(R * N) * sizeof(int32_t));
                if (dump_sweeps) dump_sweeps[nd] = isweep;
            }
            nd++;
        }
        /* src/mc_gpu.cu:714-769: thresholds */
        if (isweep % mag_output_int == 0 || isweep == tot_nsweeps - 1) {
            if (itask == 0) {
                int nnuc = 0;
                for (int r = 0; r < R; r++)
                    if ((double)colvar[r] > up_thr) nnuc++;
                int k = isweep / mag_output_int;
                if (k < result_size) res[k] = (float)((double)nnuc / (double)R);
                if (nnuc == R) break;
            } else {
                int nA = 0, nB = 0;
                for (int r = 0; r < R; r++) {
                    if (fates[r] == 0) nA++;
                    else if (fates[r] == 1) nB++;
                    else if ((double)colvar[r] > up_thr) { fates[r] = 1; nB++; }
                    else if ((double)colvar[r] < dn_thr) { fates[r] = 0; nA++; }
                }
                if (nA + nB == R) break;
            }
        }
        isweep += spc;
    }

    if (itask == 0) {
        for (int i = 0; i < result_size; i++) result[i] = res[i];
    } else {
        /* src/mc_gpu.cu:812-826 */
        for (int ii = 0; ii < ninputs; ii++) {
            int nB = 0, nF = 0;
            for (int jj = 0; jj < sub; jj++) {
                int f = fates[ii * sub + jj];
                if (f > -1) nB += f;
                else nF += 1;
            }
            result[ii] = (float)nB / (float)(sub - nF);
        }
    }
    if (ndumps) *ndumps = nd;
    free(mt);
    free(colvar);
    free(host);
    free(res);
    return nchecks;
}
