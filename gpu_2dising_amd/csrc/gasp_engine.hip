// gasp_engine.hip -- host side of the MI355X replica-parallel Metropolis engine:
// device set-up, random streams, the engine object that owns the replicas'
// lattices in a method's native layout, and mc_driver_gpu.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "gasp_internal.h"
#include "gasp_kernels.h"
#include "gasp_sliced.h"

using namespace gasp;

// =================================================================== errors
namespace {
thread_local int t_err = GASP_OK;
thread_local std::string t_msg;
void (*g_print_cb)(const char *) = nullptr;
int g_style = 0;
}  // namespace

namespace gasp {
int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = code;
    t_msg = buf;
    fprintf(stderr, "%s\n", buf);
    fflush(stderr);
    return code;
}
void clear_error() {
    t_err = GASP_OK;
    t_msg.clear();
}
void emit(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (g_print_cb) {
        g_print_cb(buf);
    } else {
        fputs(buf, stdout);
        fflush(stdout);
    }
}
int output_style() { return g_style; }
uint64_t threshold_of(double p) {
    // xi = (u + 0.5) 2^-32 is exact in double, and so is p 2^32 - 0.5 for
    // p < 1:  xi < p  <=>  u < ceil(p 2^32 - 0.5).
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return 4294967296ull;
    const double y = ceil(p * 4294967296.0 - 0.5);
    return y <= 0.0 ? 0ull : (uint64_t)y;
}
}  // namespace gasp

extern "C" int gasp_last_error(void) { return t_err; }
extern "C" const char *gasp_error_string(void) { return t_msg.c_str(); }
extern "C" void gasp_clear_error(void) { clear_error(); }
extern "C" void gasp_set_output(void (*cb)(const char *), int style) {
    g_print_cb = cb;
    g_style = style;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                         \
    } while (0)

// ============================================================ process state
namespace {

// Single-precision acceptance table of preComputeProbs_gpu and its integer
// image (process-global, as the reference's __constant__ d_Pacc).
float g_pacc_f32[20];
uint64_t g_thr_f32[10];

int g_tune_slots = 0;
int g_tune_waves = 0;
int g_tune_filter = 0;  // method 2 on half words: 1 row/column masks, 2 mark plane (0: automatic)
int g_tune_stage = 0;   // method 0 over several waves: 1 lattice in HBM, 2 staged in LDS as bytes, 3 as bits (0: automatic)
int g_tune_tables = 0;  // method 0 over several waves: most entries per dependence table (0: 16384)
thread_local gasp_run_stats_t t_stats{};

// Thresholds in kernel order j = 5*s01 + (number of up neighbours).  With
// s = 2 s01 - 1 and nsum = 2 c - 4 the reference index 5(s+1) + nsum + 4 is
// 10 s01 + 2 c.
void thr_from_table_f(const float pacc[20], uint64_t out[10]) {
    for (int s01 = 0; s01 < 2; ++s01)
        for (int c = 0; c <= 4; ++c) out[5 * s01 + c] = threshold_of((double)pacc[10 * s01 + 2 * c]);
}
void thr_from_table_d(const double pacc[20], uint64_t out[10]) {
    for (int s01 = 0; s01 < 2; ++s01)
        for (int c = 0; c <= 4; ++c) out[5 * s01 + c] = threshold_of(pacc[10 * s01 + 2 * c]);
}
void table_f32(double beta, double h, float pacc[20]) {
    // src/mc_gpu.cu:25-31, evaluated on the host in single precision
    memset(pacc, 0, 20 * sizeof(float));
    for (int s = -1; s < 2; s += 2)
        for (int nsum = -4; nsum < 5; nsum += 2) {
            const int idx = 5 * (s + 1) + nsum + 4;
            const float de = 2.0f * (float)s * ((float)nsum + (float)h);
            pacc[idx] = expf(-(float)beta * de);
        }
}
void table_f64(double beta, double h, double pacc[20]) {
    // src/mc_cpu.c:28-35
    memset(pacc, 0, 20 * sizeof(double));
    for (int s = -1; s < 2; s += 2)
        for (int nsum = -4; nsum < 5; nsum += 2) {
            const int idx = 5 * (s + 1) + nsum + 4;
            const double de = 2.0 * (double)s * ((double)nsum + h);
            pacc[idx] = exp(-beta * de);
        }
}

struct DeviceInfo {
    int ok = 0;
    int ncu = 0;
    size_t lds_block = 65536;
    size_t lds_cu = 65536;
};
DeviceInfo device_info() {
    DeviceInfo d;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return d;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return d;
    d.ok = 1;
    d.ncu = p.multiProcessorCount;
    d.lds_block = p.sharedMemPerBlock;
    d.lds_cu = p.maxSharedMemoryPerMultiProcessor ? p.maxSharedMemoryPerMultiProcessor : p.sharedMemPerBlock;
    return d;
}

// src/gpu_tools.cu:212-224 with the shared memory per block as an argument.
int select_method_for(int L, long long shmem, int tpb) {
    const int m = (int)sqrt((8.0 * (double)shmem) / (double)tpb);
    const int maxL = (int)sqrt((double)(32 * ((m * m) / 32)));
    return L <= maxL ? 2 : 0;
}

}  // namespace

// ============================================================ random streams
// gasp_rng_state: a device allocation whose first 64 bytes are an RngHeader
// (gasp_kernels.h); per-replica MT19937 states (interleaved [624][ngrids])
// follow at offset 256.
static constexpr size_t kRngMtOffset = 256;

static int rng_alloc(int ngrids, uint64_t seed, int kind, gasp_rng_state **out) {
    if (!out) return set_error(GASP_ERR_ARG, "gpuInitRand: NULL output pointer");
    *out = nullptr;
    if (ngrids < 1) return set_error(GASP_ERR_ARG, "gpuInitRand: ngrids must be positive");
    size_t bytes = kRngMtOffset;
    if (kind == KIND_MT) bytes += (size_t)624 * 4 * (size_t)ngrids;
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, bytes));
    RngHeader h{};
    h.magic = kRngMagic;
    h.kind = (uint32_t)kind;
    h.seed = seed;
    h.ngrids = ngrids;
    h.attempts = 0;
    hipError_t e = hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice);
    if (e == hipSuccess && kind == KIND_MT) {
        uint32_t *st = (uint32_t *)((char *)d + kRngMtOffset);
        hipLaunchKernelGGL(mt_seed_kernel, dim3((ngrids + 255) / 256), dim3(256), 0, 0, st, ngrids, seed);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e != hipSuccess) {
        (void)hipFree(d);
        return set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(e), __FILE__, __LINE__);
    }
    *out = (gasp_rng_state *)d;
    return GASP_OK;
}

extern "C" void gpuInitRand(int ngrids, int threadsPerBlock, unsigned long rngseed,
                            gasp_rng_state **d_state) {
    (void)threadsPerBlock;
    clear_error();
    rng_alloc(ngrids, (uint64_t)rngseed, KIND_PHILOX, d_state);
}

extern "C" int gasp_init_rand_mt(int ngrids, unsigned long rngseed, gasp_rng_state **d_state) {
    clear_error();
    return rng_alloc(ngrids, (uint64_t)rngseed, KIND_MT, d_state);
}

extern "C" void gasp_free(void *d_ptr) {
    if (d_ptr) (void)hipFree(d_ptr);
}

// ============================================================ tuning
extern "C" void gasp_set_tuning(int slots_per_replica, int waves_per_group) {
    g_tune_slots = slots_per_replica;
    g_tune_waves = waves_per_group;
}

extern "C" void gasp_set_conflict_filter(int filter) { g_tune_filter = filter; }
extern "C" void gasp_set_lattice_staging(int staging) { g_tune_stage = staging; }
extern "C" void gasp_set_dependence_tables(int max_entries) { g_tune_tables = max_entries; }

static int pow2_ceil(long long x) {
    int p = 1;
    while (p < x && p < 1024) p <<= 1;
    return p;
}

// The native layout of a method: method 3 (bit-sliced) keeps method 2's.
static int layout_of(int method) { return method == 3 ? 2 : method; }

// Native layout strides.
static int layout_stride(int method, int N) {
    if (method == 2) return (N + 3) & ~3;         // dwords per plane, two planes per 64-replica group
    if (method == 1) return (N + 31) / 32;        // dwords per replica
    return ((N + 15) / 16) * 16;                  // bytes per replica
}

// Entries of each of sweep_hbm_waves' two LDS dependence tables (a power of
// two): one per site, up to what the LDS of a CU holds for the workgroups it
// should run at once (R workgroups over the CUs, at most four per CU), and at
// most 16384 (both tables 128 KB) or gasp_set_dependence_tables' limit.
// Fewer entries than sites only add
// deferrals (sites share entries): at the batched L = 256 shape four sites
// share each entry, and a 20-sweep launch there takes 3.98 ms against 4.33
// with 8192 entries; the batched replay, whose other groups then find less
// LDS on those CUs, still went from 5.5-5.8 to 5.4-5.5 ms.
// `stage` (the kernel's STAGE): 1 or 2, the replica's spins are staged in LDS
// below the tables as bytes or bits (their budget shrinks by the lattice; 0
// entries when the lattice leaves no room); 0, they stay in HBM.
static size_t stage_bytes(int stage, int N) {
    if (stage == 1) return (size_t)layout_stride(0, N);
    if (stage == 2) return (((size_t)N + 127) / 128) * 16;  // N bits in 16-byte units
    return 0;
}
static int hbm_table_entries(int N, int R, int stage) {
    DeviceInfo d = device_info();
    const int ncu = d.ok ? d.ncu : 256;
    const size_t lds = d.ok ? d.lds_block : 65536;
    const int per_cu = std::min(4, std::max(1, (R + ncu - 1) / std::max(ncu, 1)));
    const size_t fixed = (size_t)kLdsLattice + stage_bytes(stage, N);
    const size_t per_wg = lds / (size_t)per_cu;  // at least 16 KB (at most four per CU)
    if (per_wg < fixed + 16) return 0;            // only a staged lattice can leave no room
    const size_t budget = (per_wg - fixed) / 8;
    int M = 1;
    const int cap = g_tune_tables > 0 ? std::min(g_tune_tables, 16384) : 16384;
    while (M < N && M * 2 <= cap && (size_t)M * 2 <= budget) M *= 2;
    return M;
}

// Method 0 over several waves stages the replica's spins in LDS (the rounds
// then wait on LDS instead of L2); returns the kernel's STAGE.  As bits a
// workgroup's lattice takes an eighth of the LDS of bytes, for one more shift
// and mask per spin read, and leaves the dependence tables more entries
// (fewer spurious deferrals).  Automatic: bits where they leave the tables
// more entries than bytes would (from L = 179 at one workgroup per CU, from
// smaller L at several), else bytes where the tables keep min(N, 4096)
// entries, else HBM.  20 sweeps of 256 replicas at L = 256 over eight waves
// take 2.57 ms as bits (16384 entries) against 2.65 as bytes (8192; 2.79 as
// bits with 8192); 5 sweeps at L = 512, 2.84 ms as bits against 9.55 in HBM;
// 8 at L = 384, 2.41 against 6.69 in HBM and 8.72 as bytes beside 1024
// entries.
static int hbm_stage(int N, int R) {
    if (g_tune_stage == 1) return 0;
    if (g_tune_stage == 3) return hbm_table_entries(N, R, 2) > 0 ? 2 : 0;
    const int M = hbm_table_entries(N, R, 1);
    if (g_tune_stage == 2) return M > 0 ? 1 : 0;
    const int M1 = hbm_table_entries(N, R, 2);
    if (M1 > M) return 2;
    return M > 0 && M >= std::min(N, 4096) ? 1 : 0;
}

// Method 2: `waves` is the word split (1, 2, 4, 8: 64, 32, 16, 8 replicas per
// workgroup, 8, 4, 2, 1 bytes per site in LDS).
static size_t lds_bytes_for(int method, int N, int S, int waves, int R) {
    const int stride = layout_stride(method, N);
    if (method == 2) return (size_t)kLdsLattice + (size_t)(8 / waves) * (size_t)stride;
    if (method == 1) {
        const int rpw = std::max(1, waves * 64 / S);
        return (size_t)kLdsLattice + (size_t)rpw * stride * 4;
    }
    if (S > 64) {  // sweep_hbm_waves: the tables, after the staged lattice
        const int stage = hbm_stage(N, R);
        return (size_t)kLdsLattice + stage_bytes(stage, N) + 8 * (size_t)hbm_table_entries(N, R, stage);
    }
    return (size_t)kLdsLattice;
}

// Chooses slots per replica S and waves per workgroup.  Aim: about four
// resident waves per SIMD (16 per CU) over the whole device, within the LDS
// budget; explicit settings from gasp_set_tuning win when they are valid.
// Method 3 (bit-sliced): LDS of a workgroup (16 bytes per site and a header).
static size_t sliced_lds_bytes(int N) { return (size_t)kSlLat + (size_t)kSlEntry * (size_t)N; }

// Shapes picked by gasp_autotune, keyed by (method, L, ngrids); used when no
// explicit gasp_set_tuning override is active.
static std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tuned;

static void choose_tuning(int method, int L, int R, int *S_out, int *W_out) {
    if (g_tune_slots <= 0 && g_tune_waves <= 0) {
        auto it = g_tuned.find(std::make_tuple(method, L, R));
        if (it != g_tuned.end()) {
            *S_out = it->second.first;
            *W_out = it->second.second;
            return;
        }
    }
    DeviceInfo d = device_info();
    const int ncu = d.ok ? d.ncu : 256;
    const size_t lds_block = d.ok ? d.lds_block : 65536;
    const int N = L * L;
    int W;
    // Method 2 splits the 64-bit words (32 replicas per workgroup, a dword
    // per site in LDS) unless whole words are asked for: at L=64 with 16384
    // replicas and 16 slots a launch takes 21.5 ms split against 21.8 ms
    // whole, with 2^19 replicas and 8 slots 4.44e11 attempts/s against 4.29e11,
    // and split words fit lattices up to L=201 instead of 142.
    // With few replicas the words split further (16 or 8 replicas per
    // workgroup), until there is at least one workgroup per CU: 2048 replicas
    // at 32 per workgroup would run on 64 of the 256 CUs.
    if (method == 2) {
        if (g_tune_waves == 1 && g_tune_slots <= 16) {
            W = 1;
        } else if (g_tune_waves == 2 || g_tune_waves == 4 || g_tune_waves == 8) {
            W = g_tune_waves;
        } else {
            const long long groups = (R + 63) / 64;
            W = 2;
            while (W < 8 && groups * W < ncu) W *= 2;
            // ... and until the lattice fits the LDS (eighth words: a byte
            // per site, L up to about 400 on 160 KB)
            while (W < 8 && lds_bytes_for(2, N, 1, W, R) > lds_block) W *= 2;
        }
    } else {
        W = g_tune_waves > 0 ? std::min(g_tune_waves, 16) : 1;
    }
    // method 0 may spread a replica over 2, 4, 8 or 16 waves (sweep_hbm_waves)
    int Smax = (method == 2) ? (W >= 2 ? 32 : 16) : (method == 0 ? 1024 : 64);
    int S;
    if (g_tune_slots > 0) {
        S = std::min(pow2_ceil(g_tune_slots), Smax);
    } else {
        const long long target_lanes = 16LL * ncu * 64;
        S = std::min(pow2_ceil((target_lanes + R - 1) / std::max(R, 1)), Smax);
        if (method == 0 && N >= 4096) {
            // Each replica's byte lattice is read through L2.  A whole wave
            // per replica keeps fewer replicas resident at a time, so their
            // lattices fit the L2s.  At L=64 with 16384 replicas this takes
            // a launch from 92 ms (16 lanes) to 39 ms.
            S = std::max(S, 64);
        }
        // several waves per replica only where conflicts are rare
        if (method == 0 && N < 16384) S = std::min(S, 64);
        // at most eight waves per replica: at L=256 with 256 replicas a
        // 20-sweep launch took 5.32 ms on four waves and 3.97 on eight; sixteen
        // (3.41 ms alone) crowd other kernels' workgroups off the CUs
        if (method == 0) S = std::min(S, 512);
        // With replicas to spare, several slots beat one: at L=64 with 2^19
        // replicas one slot per replica (64-lane workgroups, four per CU for
        // LDS) ran at 2.61e11 attempts/s; split words with 8 slots run at
        // 4.44e11 (4 slots: 3.88e11, 16 slots: 4.09e11).
        if (method == 2) S = std::max(S, std::min(8, Smax));
        // ... and at most 16: 32-lane groups conflict more than their extra
        // lanes gain (at L=64 with 16384 replicas 35.9 against 21.6 ms per
        // 100-sweep launch); gasp_set_tuning can still ask for 32
        if (method == 2) S = std::min(S, 16);
        // Method 1 likewise: 8 slots ran at 3.34e11 against 2.44e11 with one.
        if (method == 1) S = std::max(S, 8);
    }
    if (method == 0 && S > 64) W = S / 64;  // one replica per workgroup
    if (method == 1) {
        // enough slots so that the replica regions of one workgroup fit
        while (S < 64 && lds_bytes_for(1, N, S, W, R) > lds_block) S <<= 1;
        while (W > 1 && lds_bytes_for(1, N, S, W, R) > lds_block) W >>= 1;
    }
    *S_out = S;
    *W_out = W;
}

// Method 3: waves per workgroup (*S_out: 1, 2, 4, 8) and word split
// (*W_out: 1 = 64 replicas per word, 2 = 32).  Default: whole words, and
// waves per workgroup until there are about eight waves per CU: a round
// waits on LDS round trips and two barriers, which one wave per SIMD does not
// hide (L = 64, 16384 replicas, 100-sweep launches: 16.5 ms with one wave
// per workgroup, 9.5 with two, 6.2 with four, 4.8 with eight).
static void choose_sliced(int L, int R, int *S_out, int *W_out) {
    if (g_tune_slots <= 0 && g_tune_waves <= 0) {
        auto it = g_tuned.find(std::make_tuple(3, L, R));
        if (it != g_tuned.end()) {
            *S_out = it->second.first;
            *W_out = it->second.second;
            return;
        }
    }
    DeviceInfo d = device_info();
    const int ncu = d.ok ? d.ncu : 256;
    const int split = (g_tune_waves == 2) ? 2 : 1;
    const long long groups = (R + 64 / split - 1) / (64 / split);
    int S = 1;
    if (g_tune_slots > 0) {
        S = std::min(pow2_ceil(g_tune_slots), 8);
    } else {
        while (S < 8 && groups * S < 8LL * ncu) S *= 2;
    }
    *S_out = S;
    *W_out = split;
}

extern "C" void gasp_get_tuning(int method, int L, int ngrids, int *slots_per_replica,
                                int *waves_per_group) {
    int S = 1, W = 1;
    if (method == 3) {
        choose_sliced(L, ngrids, &S, &W);
        if (slots_per_replica) *slots_per_replica = S;
        if (waves_per_group) *waves_per_group = W;
        return;
    }
    choose_tuning(method, L, ngrids, &S, &W);
    if (slots_per_replica) *slots_per_replica = S;
    if (waves_per_group) *waves_per_group = W;
}

extern "C" void gasp_get_run_stats(gasp_run_stats_t *out) {
    if (out) *out = t_stats;
}

// ============================================================ kernel table
using SweepFn = void (*)(SweepArgs);

// Mark-plane kernels: half words with 2 to 32 slots and both streams;
// quarter and eighth words (few replicas) with 8 to 32 slots, Philox only.
template <int LAT, int RNG, bool P2, int MOFF>
static SweepFn pick_marks(int S) {
    if constexpr (LAT == LAT_MULTIHALF) {
        switch (S) {
            case 2: return sweep_kernel<LAT, RNG, 2, P2, MOFF>;
            case 4: return sweep_kernel<LAT, RNG, 4, P2, MOFF>;
            default: break;
        }
    }
    if constexpr (LAT == LAT_MULTIHALF || RNG == RNG_PHILOX) {
        switch (S) {
            case 8: return sweep_kernel<LAT, RNG, 8, P2, MOFF>;
            case 16: return sweep_kernel<LAT, RNG, 16, P2, MOFF>;
            case 32: return sweep_kernel<LAT, RNG, 32, P2, MOFF>;
            default: break;
        }
    }
    return nullptr;
}

// variant: for the split-word layouts 0 (row/column filter) or the mark
// plane's offset above the lattice (4096 .. 32768 bytes); for the byte
// lattice over several waves 1 if it is staged in LDS as bytes, 2 as bits.
template <int LAT, int RNG, bool P2>
static SweepFn pick_sweep(int S, int variant) {
    if constexpr (LAT == LAT_MULTIHALF || LAT == LAT_MULTI16 || LAT == LAT_MULTI8) {
        switch (variant) {
            case 4096: return pick_marks<LAT, RNG, P2, 4096>(S);
            case 8192: return pick_marks<LAT, RNG, P2, 8192>(S);
            case 16384: return pick_marks<LAT, RNG, P2, 16384>(S);
            case 32768: return pick_marks<LAT, RNG, P2, 32768>(S);
            default: break;
        }
    }
    switch (S) {
        case 1: return sweep_kernel<LAT, RNG, 1, P2>;
        case 2: return sweep_kernel<LAT, RNG, 2, P2>;
        case 4: return sweep_kernel<LAT, RNG, 4, P2>;
        case 8: return sweep_kernel<LAT, RNG, 8, P2>;
        case 16: return sweep_kernel<LAT, RNG, 16, P2>;
        default: break;
    }
    if constexpr (LAT != LAT_MULTISPIN) {
        if (S == 32) return sweep_kernel<LAT, RNG, 32, P2>;
    }
    if constexpr (LAT != LAT_MULTISPIN && LAT != LAT_MULTIHALF && LAT != LAT_MULTI16 && LAT != LAT_MULTI8) {
        if (S == 64) return sweep_kernel<LAT, RNG, 64, P2>;
    }
    if constexpr (LAT == LAT_HBM) {
        if (variant == 1) {
            if (S == 128) return sweep_hbm_waves<RNG, 2, P2, 1>;
            if (S == 256) return sweep_hbm_waves<RNG, 4, P2, 1>;
            if (S == 512) return sweep_hbm_waves<RNG, 8, P2, 1>;
            if (S == 1024) return sweep_hbm_waves<RNG, 16, P2, 1>;
        } else if (variant == 2) {
            if (S == 128) return sweep_hbm_waves<RNG, 2, P2, 2>;
            if (S == 256) return sweep_hbm_waves<RNG, 4, P2, 2>;
            if (S == 512) return sweep_hbm_waves<RNG, 8, P2, 2>;
            if (S == 1024) return sweep_hbm_waves<RNG, 16, P2, 2>;
        } else {
            if (S == 128) return sweep_hbm_waves<RNG, 2, P2, 0>;
            if (S == 256) return sweep_hbm_waves<RNG, 4, P2, 0>;
            if (S == 512) return sweep_hbm_waves<RNG, 8, P2, 0>;
            if (S == 1024) return sweep_hbm_waves<RNG, 16, P2, 0>;
        }
    }
    return nullptr;
}

template <int LAT, int RNG>
static SweepFn pick_sweep2(int S, bool pow2, int variant = 0) {
    return pow2 ? pick_sweep<LAT, RNG, true>(S, variant) : pick_sweep<LAT, RNG, false>(S, variant);
}

template <int WB, bool P2>
static SweepFn pick_sliced(int waves) {
    switch (waves) {
        case 1: return sweep_sliced<WB, 1, P2>;
        case 2: return sweep_sliced<WB, 2, P2>;
        case 4: return sweep_sliced<WB, 4, P2>;
        case 8: return sweep_sliced<WB, 8, P2>;
        default: return nullptr;
    }
}

// Method 3: `waves` per workgroup, `split` 1 (64 replicas per word) or 2 (32).
static SweepFn sliced_fn(int waves, int split, bool pow2) {
    if (split == 2) return pow2 ? pick_sliced<32, true>(waves) : pick_sliced<32, false>(waves);
    return pow2 ? pick_sliced<64, true>(waves) : pick_sliced<64, false>(waves);
}

// variant: see pick_sweep (method 2 on split words: the mark plane's offset;
// method 0 over several waves: 1 = lattice staged in LDS).
static SweepFn sweep_fn(int method, int rng, int S, int W, bool pow2, int variant = 0) {
    if (method == 3) return rng == 0 ? sliced_fn(S, W, pow2) : nullptr;
    if (method == 2 && W == 2)
        return rng == 0 ? pick_sweep2<LAT_MULTIHALF, RNG_PHILOX>(S, pow2, variant)
                        : pick_sweep2<LAT_MULTIHALF, RNG_DRAWS>(S, pow2, variant);
    if (method == 2 && W == 4)
        return rng == 0 ? pick_sweep2<LAT_MULTI16, RNG_PHILOX>(S, pow2, variant)
                        : pick_sweep2<LAT_MULTI16, RNG_DRAWS>(S, pow2, variant);
    if (method == 2 && W == 8)
        return rng == 0 ? pick_sweep2<LAT_MULTI8, RNG_PHILOX>(S, pow2, variant)
                        : pick_sweep2<LAT_MULTI8, RNG_DRAWS>(S, pow2, variant);
    if (method == 2)
        return rng == 0 ? pick_sweep2<LAT_MULTISPIN, RNG_PHILOX>(S, pow2) : pick_sweep2<LAT_MULTISPIN, RNG_DRAWS>(S, pow2);
    if (method == 1)
        return rng == 0 ? pick_sweep2<LAT_BITS, RNG_PHILOX>(S, pow2) : pick_sweep2<LAT_BITS, RNG_DRAWS>(S, pow2);
    return rng == 0 ? pick_sweep2<LAT_HBM, RNG_PHILOX>(S, pow2, variant)
                    : pick_sweep2<LAT_HBM, RNG_DRAWS>(S, pow2, variant);
}

static bool is_pow2(int L) { return L > 0 && (L & (L - 1)) == 0; }
static int ilog2(int L) {
    int g = 0;
    while ((1 << g) < L) ++g;
    return g;
}

// ============================================================ launch shape
// Slots per replica, waves per workgroup (word split for method 2), the
// native layout's stride and size, and the workgroup geometry of the sweep
// kernel for (method, L, R).
struct LaunchShape {
    int S = 1, W = 1, stride = 0, rpw = 0, nblocks = 0, block = 0;
    int tbl = 1;         // sweep_hbm_waves: conflict-table entries
    int stage = 0;       // sweep_hbm_waves: STAGE (0 HBM, 1 LDS bytes, 2 LDS bits)
    int ncu = 256;
    size_t lds = 0, lat_bytes = 0;
    bool sliced = false;  // method 3: sweep_sliced, S = waves per workgroup, W = word split (1: 64, 2: 32)
};

static int launch_shape(int method, int L, int R, LaunchShape *sh) {
    const int N = L * L;
    if (method == 3) {
        // bit-sliced multispin: the method-2 layout in HBM; one workgroup of
        // S waves per word group of 64 (W = 1) or 32 (W = 2) replicas
        choose_sliced(L, R, &sh->S, &sh->W);
        sh->sliced = true;
        sh->stride = layout_stride(2, N);
        DeviceInfo di = device_info();
        if (di.ok) sh->ncu = di.ncu;
        sh->lds = sliced_lds_bytes(N);
        sh->rpw = 64 / sh->W;
        sh->block = 64 * sh->S;
        sh->nblocks = (R + sh->rpw - 1) / sh->rpw;
        sh->lat_bytes = (size_t)((R + 63) / 64) * sh->stride * 8;
        if (di.ok && sh->lds > di.lds_block)
            return set_error(GASP_ERR_ARG,
                             "Error: gpu_method 3 needs %zu bytes of LDS per workgroup at L=%d (device limit %zu)",
                             sh->lds, L, (size_t)di.lds_block);
        return GASP_OK;
    }
    choose_tuning(method, L, R, &sh->S, &sh->W);
    sh->stride = layout_stride(method, N);
    DeviceInfo di = device_info();
    if (di.ok) sh->ncu = di.ncu;
    sh->lds = lds_bytes_for(method, N, sh->S, sh->W, R);
    sh->stage = (method == 0 && sh->S > 64) ? hbm_stage(N, R) : 0;
    sh->tbl = (method == 0 && sh->S > 64) ? hbm_table_entries(N, R, sh->stage) : 1;
    if (method == 2) {
        const int groups = (R + 63) / 64;
        sh->rpw = 64 / sh->W;
        sh->block = sh->rpw * sh->S;
        sh->nblocks = groups * (64 / sh->rpw);
        sh->lat_bytes = (size_t)groups * sh->stride * 8;
    } else {
        sh->block = 64 * sh->W;
        sh->rpw = std::max(1, sh->W * 64 / sh->S);
        sh->nblocks = (R + sh->rpw - 1) / sh->rpw;
        sh->lat_bytes = method == 1 ? (size_t)R * sh->stride * 4 : (size_t)R * sh->stride;
    }
    if (di.ok && sh->lds > di.lds_block)
        return set_error(GASP_ERR_ARG,
                         "Error: gpu_method %d needs %zu bytes of LDS per workgroup at L=%d (device "
                         "limit %zu); use select_gpu_method",
                         method, sh->lds, L, (size_t)di.lds_block);
    return GASP_OK;
}

// LDS of a split-word workgroup (W = 2, 4, 8: 4, 2, 1 bytes per site) with the mark plane
static size_t marks_lds_bytes(int moff, int W, int stride) {
    return (size_t)kLdsLattice + (size_t)moff + (size_t)(8 / W) * (size_t)stride;
}

// Conflict filter of method 2 on split words with 2 to 32 slots (quarter and
// eighth words: 8 to 32, Philox only): the mark plane (returns its offset
// above the lattice, the first of 4, 8, 16 or 32 KB that holds a lattice-sized
// plane, within the 64 KB reach of a ds instruction's offset: up to L = 90 on
// half words) where its LDS leaves as many workgroups resident per CU as the
// launch can use; else 0, the row/column masks.  At L = 64 a 100-sweep launch
// of 16384 replicas (two half-word workgroups per CU) takes 17.85 ms with the
// marks against 21.28 ms, but one of 524288 replicas takes 453 ms against
// 412: there the doubled LDS halves the resident workgroups.
static int choose_filter(int method, int L, const LaunchShape &sh, bool mt) {
    if (method != 2 || sh.W == 1 || sh.S < 2 || sh.S > 32 || g_tune_filter == 1) return 0;
    if (sh.W != 2 && (mt || sh.S < 8)) return 0;
    const size_t plane = (size_t)(8 / sh.W) * (size_t)sh.stride;
    int moff = 4096;
    while ((size_t)moff < plane) moff *= 2;
    if (moff > 32768) return 0;
    if (g_tune_filter == 2) return moff;
    // quarter and eighth words: only where the row/column masks are loose (32
    // slots, 64 attempts per group, or rows and columns folded for L > 64): 20
    // sweeps of 2048 replicas at L = 96 on eighth words, 32 slots, take 2.85 ms
    // with the marks against 3.88, but 4096 at L = 64 on quarter words, 16
    // slots (one workgroup of four waves per CU), 2.36 against 2.10
    if (sh.W != 2 && sh.S < 32 && L <= 64) return 0;
    DeviceInfo d = device_info();
    const size_t lds_cu = d.ok ? d.lds_block : 65536;
    const int waves = std::max(1, (sh.block + 63) / 64);
    auto resident = [&](size_t lds) {
        return std::max<long long>(1, std::min<long long>((long long)(lds_cu / lds), 32 / waves));
    };
    const long long need = (sh.nblocks + sh.ncu - 1) / std::max(sh.ncu, 1);
    const long long plain = std::min(need, resident(sh.lds));
    const long long marks = std::min(need, resident(marks_lds_bytes(moff, sh.W, sh.stride)));
    return marks >= plain ? moff : 0;
}

// ============================================================ engine object
struct gasp_engine {
    int L = 0, N = 0, R = 0, method = 0;
    int S = 1, W = 1, stride = 0, rpw = 0, nblocks = 0, block = 0;
    int tbl = 1;  // sweep_hbm_waves: conflict-table entries
    size_t lds = 0;
    void *d_lat = nullptr;
    size_t lat_bytes = 0;
    float *d_mag = nullptr;
    float *d_lclus = nullptr;
    int *d_fate = nullptr;
    int *d_counts = nullptr;
    uint64_t *d_thr = nullptr;
    int8_t *d_spins8 = nullptr;
    int *d_tmp32 = nullptr;
    size_t tmp32_bytes = 0;
    int freeze = 0;          // skip replicas whose fate is decided
    int S0 = 1;              // slots per replica chosen at creation
    int moff = 0;            // mark plane offset above the lattice (0: row/column filter)
    int stage = 0;           // method 0 over several waves: 1 lattice staged in LDS as bytes, 2 as bits
    bool adapt = false;      // raise the slots as replicas are decided (see engine_adapt_slots)
    int decided = 0;         // decided replicas at the last counts read (or set_fates)
    int ncu = 256;
    uint32_t *d_draws = nullptr;
    uint64_t draws_cap = 0;  // attempts per replica that fit in d_draws
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev_pending;  // start/stop pairs not yet accumulated
    double kernel_ms = 0.0;
    int launches = 0;
    gasp_rng_state *d_rng = nullptr;
    RngHeader rng{};
    uint64_t thr[10]{};
};

static void dfree(void *p) {
    if (p) (void)hipFree(p);
}

static void engine_free(gasp_engine *e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (hipEvent_t ev : e->ev_pending) (void)hipEventDestroy(ev);
    dfree(e->d_lat);
    dfree(e->d_mag);
    dfree(e->d_lclus);
    dfree(e->d_fate);
    dfree(e->d_counts);
    dfree(e->d_thr);
    dfree(e->d_spins8);
    dfree(e->d_tmp32);
    dfree(e->d_draws);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

static int engine_new(int L, int R, int method, gasp_rng_state *d_state, const uint64_t thr[10],
                      int thr_from_double, double beta, double h, gasp_engine **out) {
    *out = nullptr;
    if (L < 2 || L > 8192) return set_error(GASP_ERR_ARG, "Error: L must lie in [2, 8192]");
    if (R < 1) return set_error(GASP_ERR_ARG, "Error: ngrids must be positive");
    if (method < 0 || method > 3) return set_error(GASP_ERR_ARG, "Unknown gpu_method in ising.cu");
    if (!d_state) return set_error(GASP_ERR_ARG, "Error: RNG state is NULL (call gpuInitRand)");
    gasp_engine *e = new gasp_engine();
    e->L = L;
    e->N = L * L;
    e->R = R;
    e->method = method;
    e->d_rng = d_state;
    hipError_t he = hipMemcpy(&e->rng, d_state, sizeof(RngHeader), hipMemcpyDeviceToHost);
    if (he != hipSuccess) {
        engine_free(e);
        return set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(he), __FILE__, __LINE__);
    }
    if (e->rng.magic != kRngMagic) {
        engine_free(e);
        return set_error(GASP_ERR_ARG, "Error: RNG state was not made by gpuInitRand");
    }
    if (e->rng.ngrids < R) {
        engine_free(e);
        return set_error(GASP_ERR_ARG, "Error: RNG state holds %d streams, need %d", e->rng.ngrids, R);
    }
    if (e->rng.kind == KIND_MT && method == 3) {
        engine_free(e);
        return set_error(GASP_ERR_ARG, "Error: gpu_method 3 (bit-sliced) draws from Philox streams (gpuInitRand)");
    }
    if (e->rng.kind == KIND_MT && e->rng.ngrids != R) {
        engine_free(e);
        return set_error(GASP_ERR_ARG, "Error: MT19937 state was made for %d replicas, run has %d",
                         e->rng.ngrids, R);
    }
    if (e->rng.kind == KIND_MT || thr_from_double) {
        double pd[20];
        table_f64(beta, h, pd);
        thr_from_table_d(pd, e->thr);
    } else {
        memcpy(e->thr, thr, sizeof e->thr);
    }

    LaunchShape sh;
    if (launch_shape(method, L, R, &sh)) {
        engine_free(e);
        return gasp_last_error();
    }
    e->S = sh.S;
    e->W = sh.W;
    e->S0 = e->S;
    e->adapt = method == 2 && g_tune_slots <= 0;
    e->stride = sh.stride;
    e->ncu = sh.ncu;
    e->lds = sh.lds;
    // the mark plane (if any) lies above the lattice, in the same allocation
    e->moff = choose_filter(method, L, sh, e->rng.kind == KIND_MT);
    if (e->moff) e->lds = marks_lds_bytes(e->moff, e->W, e->stride);
    e->tbl = sh.tbl;
    e->stage = sh.stage;
    e->rpw = sh.rpw;
    e->block = sh.block;
    e->nblocks = sh.nblocks;
    e->lat_bytes = sh.lat_bytes;
    SweepFn fn = sweep_fn(method, e->rng.kind == KIND_MT ? 1 : 0, e->S, e->W, is_pow2(L), e->stage ? e->stage : e->moff);
    if (!fn) {
        engine_free(e);
        return set_error(GASP_ERR_ARG, "Error: no kernel for S=%d", e->S);
    }
#define ETRY(x)                                                                                   \
    do {                                                                                          \
        hipError_t q_ = (x);                                                                      \
        if (q_ != hipSuccess) {                                                                   \
            engine_free(e);                                                                       \
            return set_error(q_ == hipErrorOutOfMemory ? GASP_ERR_NOMEM : GASP_ERR_HIP,           \
                             "GPUassert: %s %s %d", hipGetErrorString(q_), __FILE__, __LINE__);   \
        }                                                                                         \
    } while (0)
    ETRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    ETRY(hipMalloc(&e->d_lat, e->lat_bytes));
    ETRY(hipMemsetAsync(e->d_lat, 0, e->lat_bytes, e->stream));
    ETRY(hipMalloc(&e->d_mag, (size_t)R * sizeof(float)));
    ETRY(hipMalloc(&e->d_fate, (size_t)R * sizeof(int)));
    ETRY(hipMalloc(&e->d_counts, 4 * sizeof(int)));
    ETRY(hipMalloc(&e->d_thr, 10 * sizeof(uint64_t)));
    ETRY(hipMemcpyAsync(e->d_thr, e->thr, sizeof e->thr, hipMemcpyHostToDevice, e->stream));
    ETRY(hipMemsetAsync(e->d_fate, 0xff, (size_t)R * sizeof(int), e->stream));  // -1
    ETRY(hipMemsetAsync(e->d_counts, 0, 4 * sizeof(int), e->stream));
    ETRY(hipMemsetAsync(e->d_mag, 0, (size_t)R * sizeof(float), e->stream));
    if (e->lds > 65536) ETRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->lds));
    if (e->rng.kind == KIND_MT) {
        // draw stream for up to ~512 MiB per batch (12 bytes per attempt per replica)
        const uint64_t budget = 512ull << 20;
        e->draws_cap = std::max<uint64_t>(1, budget / (12ull * (uint64_t)R));
        // GASP_MT_BATCH_ATTEMPTS lowers the batch (exercises batch boundaries in tests)
        if (const char *env = getenv("GASP_MT_BATCH_ATTEMPTS")) {
            const unsigned long long v = strtoull(env, nullptr, 10);
            if (v > 0 && v < e->draws_cap) e->draws_cap = v;
        }
        ETRY(hipMalloc(&e->d_draws, (size_t)e->draws_cap * 12 * (size_t)R));
    }
    ETRY(hipStreamSynchronize(e->stream));
#undef ETRY
    *out = e;
    return GASP_OK;
}

static int engine_collect_events(gasp_engine *e) {
    for (size_t i = 0; i + 1 < e->ev_pending.size(); i += 2) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e->ev_pending[i], e->ev_pending[i + 1]));
        e->kernel_ms += ms;
        (void)hipEventDestroy(e->ev_pending[i]);
        (void)hipEventDestroy(e->ev_pending[i + 1]);
    }
    e->ev_pending.clear();
    return GASP_OK;
}

static int engine_sync(gasp_engine *e) {
    HIP_TRY(hipStreamSynchronize(e->stream));
    return engine_collect_events(e);
}

// An engine that freezes decided replicas (gasp_engine_set_freeze) raises its
// slots per replica as replicas are decided, keeping about 16 waves per CU of
// lanes on the undecided ones: a wave then holds fewer replicas, waves whose
// replicas are all decided exit at once, and each undecided replica gets more
// lanes.  Trajectories do not depend on the slots, so no result changes.
// Method 2 only (its LDS does not depend on the slots); off when
// gasp_set_tuning fixed the slots at creation.
static int engine_adapt_slots(gasp_engine *e) {
    if (!e->adapt || !e->freeze) return GASP_OK;
    const long long target = 16LL * e->ncu * 64;
    const long long active = std::max(1, e->R - e->decided);
    // up to 16: with 32 the committor leg's shooting (524288 replicas) took
    // 4.27 s against 3.99 and the surrogate's fallback shooting (16256)
    // 0.48 s against 0.32, the straggler rounds conflicting more
    const int smax = 16;
    const int S = std::max(std::min(pow2_ceil((target + active - 1) / active), smax), e->S0);
    if (S == e->S) return GASP_OK;
    SweepFn fn = sweep_fn(e->method, e->rng.kind == KIND_MT ? 1 : 0, S, e->W, is_pow2(e->L), e->moff);
    if (!fn) return GASP_OK;
    if (e->lds > 65536)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->lds));
    e->S = S;
    e->block = e->rpw * S;
    return GASP_OK;
}

static int launch_sweep(gasp_engine *e, uint64_t n_attempts, const uint32_t *draws, int check,
                        double dn, double up, bool capture = false) {
    if (!capture) {
        const int rc = engine_adapt_slots(e);
        if (rc) return rc;
    }
    SweepArgs a{};
    a.lattice = e->d_lat;
    a.L = e->L;
    a.N = e->N;
    a.R = e->R;
    a.lg = ilog2(e->L);
    a.rpw = e->rpw;
    a.stride = e->stride;
    a.a_begin = e->rng.attempts;
    // in a captured graph the attempt counter lives in the RNG header on the
    // device and is advanced by a kernel after the sweep
    a.rng_dev = capture ? (const RngHeader *)e->d_rng : nullptr;
    a.n_attempts = n_attempts;
    a.seed = e->rng.seed;
    a.draws = draws;
    a.thr = e->d_thr;
    a.mag = e->d_mag;
    a.fate = e->d_fate;
    a.counts = e->d_counts;
    a.check = check;
    a.freeze = e->freeze;
    a.tbl_mask = (unsigned)(e->tbl - 1);
    a.dn_thr = dn;
    a.up_thr = up;
    SweepFn fn = sweep_fn(e->method, e->rng.kind == KIND_MT ? 1 : 0, e->S, e->W, is_pow2(e->L),
                          e->stage ? e->stage : e->moff);
    if (check != CHECK_NONE) {
        hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, e->stream, e->d_counts, 4);
        HIP_TRY(hipGetLastError());
    }
    if (capture) {
        hipLaunchKernelGGL(fn, dim3(e->nblocks), dim3(e->block), e->lds, e->stream, a);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(advance_counter, dim3(1), dim3(64), 0, e->stream, (RngHeader *)e->d_rng, e->R,
                           n_attempts);
        HIP_TRY(hipGetLastError());
        return GASP_OK;
    }
    hipEvent_t ev0, ev1;
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventCreate(&ev1));
    HIP_TRY(hipEventRecord(ev0, e->stream));
    hipLaunchKernelGGL(fn, dim3(e->nblocks), dim3(e->block), e->lds, e->stream, a);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev1, e->stream));
    e->ev_pending.push_back(ev0);
    e->ev_pending.push_back(ev1);
    e->launches++;
    e->rng.attempts += n_attempts;
    return GASP_OK;
}

// Advance all replicas by n attempts; `check` applies to the state at the end.
// Launches carry at most 2^30 attempts per replica (32-bit launch-local
// counters); MT19937 draws are generated in batches that fit the buffer.
static int engine_advance(gasp_engine *e, uint64_t n, int check, double dn, double up) {
    if (n == 0) return launch_sweep(e, 0, nullptr, check, dn, up);
    const uint64_t cap = e->rng.kind == KIND_MT ? std::min<uint64_t>(e->draws_cap, 1ull << 30) : (1ull << 30);
    uint64_t left = n;
    while (left > 0) {
        const uint64_t nb = std::min(left, cap);
        if (e->rng.kind == KIND_MT) {
            uint32_t *st = (uint32_t *)((char *)e->d_rng + kRngMtOffset);
            hipLaunchKernelGGL(mt_draws_kernel, dim3(e->R), dim3(64), 0, e->stream, st, e->R,
                               3ull * e->rng.attempts, 3ull * nb, e->d_draws);
            HIP_TRY(hipGetLastError());
        }
        const int rc = launch_sweep(e, nb, e->rng.kind == KIND_MT ? e->d_draws : nullptr,
                                    left == nb ? check : CHECK_NONE, dn, up);
        if (rc) return rc;
        left -= nb;
    }
    return GASP_OK;
}

// Packs int32 grids into the native layout.  `ninputs` grids are given;
// replica r takes grid r / (R / ninputs) (ninputs == R: one grid per replica).
static int engine_set_grids_impl(gasp_engine *e, const int *grids, int on_device, int ninputs) {
    const size_t bytes = (size_t)ninputs * e->N * sizeof(int);
    const int *src = grids;
    if (!on_device) {
        if (!e->d_tmp32 || e->tmp32_bytes < bytes) {
            dfree(e->d_tmp32);
            e->d_tmp32 = nullptr;
            HIP_TRY(hipMalloc(&e->d_tmp32, bytes));
            e->tmp32_bytes = bytes;
        }
        HIP_TRY(hipMemcpyAsync(e->d_tmp32, grids, bytes, hipMemcpyHostToDevice, e->stream));
        src = e->d_tmp32;
    }
    const int N = e->N, R = e->R, per = R / ninputs;
    const dim3 grid(1024), blk(256);
    if (layout_of(e->method) == 2) {
        hipLaunchKernelGGL(pack_multispin, grid, blk, 0, e->stream, src, (unsigned *)e->d_lat, R, N, e->stride, per);
    } else if (e->method == 1) {
        hipLaunchKernelGGL(pack_bits, grid, blk, 0, e->stream, src, (unsigned *)e->d_lat, R, N, e->stride, per);
    } else {
        hipLaunchKernelGGL(pack_hbm, grid, blk, 0, e->stream, src, (unsigned char *)e->d_lat, R, N, e->stride, per);
    }
    HIP_TRY(hipGetLastError());
    return GASP_OK;
}

template <typename T>
static int engine_unpack(gasp_engine *e, T *dst) {
    const int N = e->N, R = e->R;
    const dim3 grid(1024), blk(256);
    if (layout_of(e->method) == 2) {
        hipLaunchKernelGGL(unpack_multispin<T>, grid, blk, 0, e->stream, (const unsigned *)e->d_lat, dst, R, N,
                           e->stride);
    } else if (e->method == 1) {
        hipLaunchKernelGGL(unpack_bits<T>, grid, blk, 0, e->stream, (const unsigned *)e->d_lat, dst, R, N, e->stride);
    } else {
        hipLaunchKernelGGL(unpack_hbm<T>, grid, blk, 0, e->stream, (const unsigned char *)e->d_lat, dst, R, N,
                           e->stride);
    }
    HIP_TRY(hipGetLastError());
    return GASP_OK;
}

static int engine_get_grids_impl(gasp_engine *e, int *grids, int on_device) {
    const size_t bytes = (size_t)e->R * e->N * sizeof(int);
    if (on_device) return engine_unpack<int>(e, grids);
    if (!e->d_tmp32 || e->tmp32_bytes < bytes) {
        dfree(e->d_tmp32);
        e->d_tmp32 = nullptr;
        HIP_TRY(hipMalloc(&e->d_tmp32, bytes));
        e->tmp32_bytes = bytes;
    }
    int rc = engine_unpack<int>(e, e->d_tmp32);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(grids, e->d_tmp32, bytes, hipMemcpyDeviceToHost, e->stream));
    return engine_sync(e);
}

// ------------------------------------------------------------ largest cluster
// One workgroup per replica.  Labels and per-root sizes live in LDS when
// 8*N bytes fit, else in a global workspace.  Lock-free union-find: every
// site of the target spin starts as its own root; each site unites with its
// right and lower neighbours (periodic) by hooking the larger root under the
// smaller with atomicMin and retrying from the returned value when another
// lane won the race; a final pass points every site at its root, sizes are
// counted with atomics and the largest is reduced.
template <bool USE_LDS>
__global__ void __launch_bounds__(256) cluster_kernel(const int8_t *spins, int N, int L, int spin,
                                                      int *work, float *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int wmax[4];
    const int r = blockIdx.x;
    const int8_t *g = spins + (size_t)r * N;
    int *lab, *cnt;
    if constexpr (USE_LDS) {
        lab = (int *)smem;
        cnt = lab + N;
    } else {
        lab = work + (size_t)r * 2 * N;
        cnt = lab + N;
    }
    const int t = threadIdx.x;
    for (int i = t; i < N; i += blockDim.x) {
        lab[i] = (g[i] == spin) ? i : -1;
        cnt[i] = 0;
    }
    __syncthreads();
    auto find = [&](int x) {
        int y = lab[x];
        while (y != x) {
            x = y;
            y = lab[x];
        }
        return x;
    };
    for (int i = t; i < N; i += blockDim.x) {
        if (lab[i] < 0) continue;
        const int row = i / L, col = i - row * L;
        const int nb[2] = {row * L + (col == L - 1 ? 0 : col + 1), (row == L - 1 ? 0 : row + 1) * L + col};
        for (int k = 0; k < 2; ++k) {
            if (g[nb[k]] != spin) continue;
            int a = find(i), b = find(nb[k]);
            while (a != b) {
                if (a < b) {
                    const int tmp = a;
                    a = b;
                    b = tmp;
                }
                // a is the larger root: hook it under b
                const int old = atomicMin(&lab[a], b);
                if (old == a) break;
                a = find(old);
                b = find(b);
            }
        }
    }
    __syncthreads();
    for (int i = t; i < N; i += blockDim.x)
        if (lab[i] >= 0) atomicAdd(&cnt[find(i)], 1);
    __syncthreads();
    int best = 0;
    for (int i = t; i < N; i += blockDim.x) best = max(best, cnt[i]);
    for (int off = 32; off >= 1; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
    if ((t & 63) == 0) wmax[t >> 6] = best;
    __syncthreads();
    if (t == 0) {
        int m = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = max(m, wmax[w]);
        out[r] = (float)m;
    }
}

static int engine_cluster(gasp_engine *e, int spin) {
    const size_t n8 = (size_t)e->R * e->N;
    if (!e->d_spins8) HIP_TRY(hipMalloc(&e->d_spins8, n8));
    if (!e->d_lclus) HIP_TRY(hipMalloc(&e->d_lclus, (size_t)e->R * sizeof(float)));
    int rc = engine_unpack<int8_t>(e, e->d_spins8);
    if (rc) return rc;
    DeviceInfo di = device_info();
    const size_t lds = 8 * (size_t)e->N;
    if (lds + 64 <= (di.ok ? di.lds_block : 65536)) {
        if (lds > 65536)
            HIP_TRY(hipFuncSetAttribute((const void *)cluster_kernel<true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(cluster_kernel<true>, dim3(e->R), dim3(256), lds, e->stream, e->d_spins8, e->N, e->L,
                           spin, (int *)nullptr, e->d_lclus);
    } else {
        int *work = nullptr;
        HIP_TRY(hipMallocAsync((void **)&work, (size_t)e->R * 2 * e->N * sizeof(int), e->stream));
        hipLaunchKernelGGL(cluster_kernel<false>, dim3(e->R), dim3(256), 0, e->stream, e->d_spins8, e->N, e->L,
                           spin, work, e->d_lclus);
        HIP_TRY(hipFreeAsync(work, e->stream));
    }
    HIP_TRY(hipGetLastError());
    return GASP_OK;
}

static int engine_check_cv(gasp_engine *e, const float *d_cv, int check, double dn, double up) {
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, e->stream, e->d_counts, 4);
    hipLaunchKernelGGL(check_cv_kernel, dim3((e->R + 255) / 256), dim3(256), 0, e->stream, d_cv, e->R, check,
                       dn, up, e->d_fate, e->d_counts);
    HIP_TRY(hipGetLastError());
    return GASP_OK;
}

static int engine_store_rng(gasp_engine *e) {
    HIP_TRY(hipMemcpyAsync(e->d_rng, &e->rng, sizeof(RngHeader), hipMemcpyHostToDevice, e->stream));
    return engine_sync(e);
}

// ------------------------------------------------------------ engine C API
extern "C" int gasp_engine_create(int L, int ngrids, int gpu_method, gasp_rng_state *d_state, double beta,
                                  double h, gasp_engine **out) {
    clear_error();
    if (!out) return set_error(GASP_ERR_ARG, "gasp_engine_create: NULL output pointer");
    float pf[20];
    uint64_t thr[10];
    table_f32(beta, h, pf);
    thr_from_table_f(pf, thr);
    return engine_new(L, ngrids, gpu_method, d_state, thr, 0, beta, h, out);
}

extern "C" int gasp_engine_destroy(gasp_engine *e) {
    if (!e) return GASP_OK;
    int rc = engine_store_rng(e);
    engine_free(e);
    return rc;
}

extern "C" int gasp_engine_set_grids(gasp_engine *e, const int *grids, int on_device) {
    clear_error();
    if (!e || !grids) return set_error(GASP_ERR_ARG, "gasp_engine_set_grids: NULL argument");
    int rc = engine_set_grids_impl(e, grids, on_device, e->R);
    return rc ? rc : engine_sync(e);
}

extern "C" int gasp_engine_set_inputs(gasp_engine *e, const int *inputs, int ninputs) {
    clear_error();
    if (!e || !inputs) return set_error(GASP_ERR_ARG, "gasp_engine_set_inputs: NULL argument");
    if (ninputs < 1 || e->R % ninputs != 0)
        return set_error(GASP_ERR_ARG, "Error: ngrids must be a multiple of ninputs!");
    int rc = engine_set_grids_impl(e, inputs, 0, ninputs);
    return rc ? rc : engine_sync(e);
}

extern "C" int gasp_engine_set_freeze(gasp_engine *e, int freeze) {
    if (!e) return set_error(GASP_ERR_ARG, "gasp_engine_set_freeze: NULL engine");
    e->freeze = freeze ? 1 : 0;
    return GASP_OK;
}

extern "C" int gasp_engine_get_grids(gasp_engine *e, int *grids, int on_device) {
    clear_error();
    if (!e || !grids) return set_error(GASP_ERR_ARG, "gasp_engine_get_grids: NULL argument");
    int rc = engine_get_grids_impl(e, grids, on_device);
    return rc ? rc : engine_sync(e);
}

extern "C" int gasp_engine_get_gridstates(gasp_engine *e, uint8_t *host_out) {
    clear_error();
    if (!e || !host_out) return set_error(GASP_ERR_ARG, "gasp_engine_get_gridstates: NULL argument");
    const size_t nbytes = ((size_t)e->R * e->N + 7) / 8;
    uint8_t *d = nullptr;
    hipError_t he = hipMalloc(&d, nbytes);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(pack_gridstates, dim3(2048), dim3(256), 0, e->stream, (const void *)e->d_lat, layout_of(e->method),
                           e->stride, e->R, e->N, d, nbytes);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipMemcpyAsync(host_out, d, nbytes, hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    dfree(d);
    if (he != hipSuccess)
        return set_error(he == hipErrorOutOfMemory ? GASP_ERR_NOMEM : GASP_ERR_HIP, "GPUassert: %s %s %d",
                         hipGetErrorString(he), __FILE__, __LINE__);
    return engine_collect_events(e);
}

extern "C" int gasp_engine_gather_grids(gasp_engine *e, const int *replicas, int n, int8_t *host_out) {
    clear_error();
    if (!e || (n > 0 && (!replicas || !host_out)) || n < 0)
        return set_error(GASP_ERR_ARG, "gasp_engine_gather_grids: bad argument");
    if (n == 0) return GASP_OK;
    for (int k = 0; k < n; ++k)
        if (replicas[k] < 0 || replicas[k] >= e->R)
            return set_error(GASP_ERR_ARG, "gasp_engine_gather_grids: replica %d out of range [0, %d)", replicas[k],
                             e->R);
    int *d_idx = nullptr;
    int8_t *d_out = nullptr;
    const size_t bytes = (size_t)n * e->N;
    hipError_t he = hipMalloc(&d_idx, (size_t)n * sizeof(int));
    if (he == hipSuccess) he = hipMalloc(&d_out, bytes);
    if (he == hipSuccess) he = hipMemcpyAsync(d_idx, replicas, (size_t)n * sizeof(int), hipMemcpyHostToDevice, e->stream);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(gather_spins8, dim3(1024), dim3(256), 0, e->stream, (const void *)e->d_lat, layout_of(e->method),
                           e->stride, d_idx, n, e->N, d_out);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipMemcpyAsync(host_out, d_out, bytes, hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    dfree(d_idx);
    dfree(d_out);
    if (he != hipSuccess)
        return set_error(he == hipErrorOutOfMemory ? GASP_ERR_NOMEM : GASP_ERR_HIP, "GPUassert: %s %s %d",
                         hipGetErrorString(he), __FILE__, __LINE__);
    return engine_collect_events(e);
}

extern "C" int gasp_engine_sweep(gasp_engine *e, int nsweeps, int check, double dn_thr, double up_thr) {
    clear_error();
    if (!e) return set_error(GASP_ERR_ARG, "gasp_engine_sweep: NULL engine");
    if (nsweeps < 0 || check < 0 || check > 2) return set_error(GASP_ERR_ARG, "gasp_engine_sweep: bad argument");
    return engine_advance(e, (uint64_t)nsweeps * (uint64_t)e->N, check, dn_thr, up_thr);
}

extern "C" int gasp_engine_sync(gasp_engine *e) {
    if (!e) return set_error(GASP_ERR_ARG, "gasp_engine_sync: NULL engine");
    return engine_sync(e);
}

extern "C" int gasp_engine_get_magnetisation(gasp_engine *e, float *host_out) {
    if (!e || !host_out) return set_error(GASP_ERR_ARG, "gasp_engine_get_magnetisation: NULL argument");
    HIP_TRY(hipMemcpyAsync(host_out, e->d_mag, (size_t)e->R * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    return engine_sync(e);
}

extern "C" int gasp_engine_get_counts(gasp_engine *e, int *counts3) {
    if (!e || !counts3) return set_error(GASP_ERR_ARG, "gasp_engine_get_counts: NULL argument");
    HIP_TRY(hipMemcpyAsync(counts3, e->d_counts, 3 * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    const int rc = engine_sync(e);
    if (!rc) e->decided = counts3[1] + counts3[2];
    return rc;
}

extern "C" int gasp_engine_set_fates(gasp_engine *e, const int *host_fates) {
    if (!e || !host_fates) return set_error(GASP_ERR_ARG, "gasp_engine_set_fates: NULL argument");
    HIP_TRY(hipMemcpyAsync(e->d_fate, host_fates, (size_t)e->R * sizeof(int), hipMemcpyHostToDevice, e->stream));
    e->decided = 0;
    for (int r = 0; r < e->R; ++r) e->decided += host_fates[r] != -1;
    return engine_sync(e);
}

extern "C" int gasp_engine_get_fates(gasp_engine *e, int *host_fates) {
    if (!e || !host_fates) return set_error(GASP_ERR_ARG, "gasp_engine_get_fates: NULL argument");
    HIP_TRY(hipMemcpyAsync(host_fates, e->d_fate, (size_t)e->R * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    return engine_sync(e);
}

extern "C" int gasp_engine_largest_cluster(gasp_engine *e, int spin, float *host_out) {
    clear_error();
    if (!e || !host_out) return set_error(GASP_ERR_ARG, "gasp_engine_largest_cluster: NULL argument");
    int rc = engine_cluster(e, spin);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(host_out, e->d_lclus, (size_t)e->R * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    return engine_sync(e);
}

extern "C" double gasp_engine_kernel_ms(gasp_engine *e) {
    if (!e) return 0.0;
    engine_sync(e);
    return e->kernel_ms;
}

extern "C" void *gasp_engine_stream(gasp_engine *e) { return e ? (void *)e->stream : nullptr; }

extern "C" int gasp_engine_set_attempt(gasp_engine *e, unsigned long long attempt) {
    clear_error();
    if (!e) return set_error(GASP_ERR_ARG, "gasp_engine_set_attempt: NULL engine");
    if (e->rng.kind != KIND_PHILOX)
        return set_error(GASP_ERR_ARG, "gasp_engine_set_attempt: only counter-based (Philox) streams rewind");
    const int rc = engine_sync(e);  // launches in flight used the old position
    if (rc) return rc;
    e->rng.attempts = (uint64_t)attempt;
    HIP_TRY(hipMemcpy(e->d_rng, &e->rng, sizeof(RngHeader), hipMemcpyHostToDevice));
    return GASP_OK;
}

extern "C" int gasp_engine_get_shape(gasp_engine *e, int *slots, int *waves) {
    if (!e || !slots || !waves) return set_error(GASP_ERR_ARG, "gasp_engine_get_shape: NULL argument");
    *slots = e->S;
    *waves = e->W;
    return GASP_OK;
}

extern "C" int gasp_engine_get_conflict_filter(gasp_engine *e, int *filter) {
    if (!e || !filter) return set_error(GASP_ERR_ARG, "gasp_engine_get_conflict_filter: NULL argument");
    *filter = e->moff ? 2 : 1;
    return GASP_OK;
}

extern "C" int gasp_engine_get_lattice_staging(gasp_engine *e, int *staging) {
    if (!e || !staging) return set_error(GASP_ERR_ARG, "gasp_engine_get_lattice_staging: NULL argument");
    *staging = e->stage + 1;
    return GASP_OK;
}

// ============================================================ caller grids
// Stream-ordered sweeps of caller-owned int32 device grids (the layout
// gpuInitGrid makes): the granularity of the reference driver's per-interval
// launches (src/mc_gpu.cu:675-708: mc_sweep_gpu*, then
// compute_magnetisation_gpu) without an engine object.  The Philox key and
// the attempt counter come from the RNG header in device memory and the
// counter is advanced there by a kernel, so nothing is read back to the host
// and consecutive calls continue one chain per replica.  A workspace holds,
// at 256-byte aligned offsets: the 10 thresholds (0), 4 counters (128), the
// magnetisation when the caller passes none (256: R floats) and the native
// lattice.

// Stores the thresholds and clears the counters (a kernel rather than a host
// copy or a memset, so that the sequence can be captured and replayed).
struct Thresholds10 {
    uint64_t v[10];
};
__global__ void set_thresholds(Thresholds10 t, uint64_t *thr, int *counts) {
    const int i = threadIdx.x;
    if (i < 10) thr[i] = t.v[i];
    if (i < 4) counts[i] = 0;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static size_t grids_ws_lattice_offset(int R, bool own_mag) {
    return own_mag ? align256(256 + (size_t)R * sizeof(float)) : 256;
}

static int grids_check_args(int L, int R, int method, gasp_rng_state *d_state, const int *d_grids, int nsweeps) {
    if (L < 2 || L > 8192) return set_error(GASP_ERR_ARG, "Error: L must lie in [2, 8192]");
    if (R < 1) return set_error(GASP_ERR_ARG, "Error: ngrids must be positive");
    if (method < 0 || method > 3) return set_error(GASP_ERR_ARG, "Unknown gpu_method in ising.cu");
    if (!d_state) return set_error(GASP_ERR_ARG, "Error: RNG state is NULL (call gpuInitRand)");
    if (!d_grids) return set_error(GASP_ERR_ARG, "Error: device grids are NULL");
    if (nsweeps < 0) return set_error(GASP_ERR_ARG, "Error: nsweeps must not be negative");
    return GASP_OK;
}

// The workspace must hold grids_ws_lattice_offset(R, !d_mag) + sh.lat_bytes.
static int sweep_grids_run(int L, int R, int method, gasp_rng_state *d_state, double beta, double h, int *d_grids,
                           int nsweeps, float *d_mag, hipStream_t st, const LaunchShape &sh, unsigned char *ws) {
    const int N = L * L;
    float pf[20];
    Thresholds10 t;
    table_f32(beta, h, pf);
    thr_from_table_f(pf, t.v);
    uint64_t *d_thr = (uint64_t *)ws;
    int *d_counts = (int *)(ws + 128);
    float *mag = d_mag ? d_mag : (float *)(ws + 256);
    void *lat = ws + grids_ws_lattice_offset(R, !d_mag);
    hipLaunchKernelGGL(set_thresholds, dim3(1), dim3(64), 0, st, t, d_thr, d_counts);
    HIP_TRY(hipGetLastError());
    const dim3 grid(1024), blk(256);
    const int lay = layout_of(method);
    if (lay == 2)
        hipLaunchKernelGGL(pack_multispin, grid, blk, 0, st, d_grids, (unsigned *)lat, R, N, sh.stride, 1);
    else if (lay == 1)
        hipLaunchKernelGGL(pack_bits, grid, blk, 0, st, d_grids, (unsigned *)lat, R, N, sh.stride, 1);
    else
        hipLaunchKernelGGL(pack_hbm, grid, blk, 0, st, d_grids, (unsigned char *)lat, R, N, sh.stride, 1);
    HIP_TRY(hipGetLastError());
    SweepFn fn = sweep_fn(method, 0, sh.S, sh.W, is_pow2(L), sh.stage);
    if (!fn) return set_error(GASP_ERR_ARG, "Error: no kernel for S=%d", sh.S);
    if (sh.lds > 65536)
        HIP_TRY(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds));
    SweepArgs a{};
    a.lattice = lat;
    a.L = L;
    a.N = N;
    a.R = R;
    a.lg = ilog2(L);
    a.rpw = sh.rpw;
    a.stride = sh.stride;
    a.rng_dev = (const RngHeader *)d_state;
    a.thr = d_thr;
    a.mag = mag;
    a.counts = d_counts;
    a.check = CHECK_NONE;
    a.tbl_mask = (unsigned)(sh.tbl - 1);
    // launches of at most 2^30 attempts per replica (32-bit launch-local
    // counters); with no sweeps one launch still computes the magnetisation
    uint64_t left = (uint64_t)nsweeps * (uint64_t)N;
    do {
        const uint64_t nb = std::min<uint64_t>(left, 1ull << 30);
        a.n_attempts = nb;
        hipLaunchKernelGGL(fn, dim3(sh.nblocks), dim3(sh.block), sh.lds, st, a);
        HIP_TRY(hipGetLastError());
        if (nb) {
            hipLaunchKernelGGL(advance_counter, dim3(1), dim3(64), 0, st, (RngHeader *)d_state, R, nb);
            HIP_TRY(hipGetLastError());
        }
        left -= nb;
    } while (left > 0);
    if (lay == 2)
        hipLaunchKernelGGL(unpack_multispin<int>, grid, blk, 0, st, (const unsigned *)lat, d_grids, R, N, sh.stride);
    else if (lay == 1)
        hipLaunchKernelGGL(unpack_bits<int>, grid, blk, 0, st, (const unsigned *)lat, d_grids, R, N, sh.stride);
    else
        hipLaunchKernelGGL(unpack_hbm<int>, grid, blk, 0, st, (const unsigned char *)lat, d_grids, R, N, sh.stride);
    HIP_TRY(hipGetLastError());
    return GASP_OK;
}

extern "C" int gasp_sweep_grids(int L, int ngrids, int gpu_method, gasp_rng_state *d_state, double beta, double h,
                                int *d_ising_grids, int nsweeps, float *d_mag, void *stream) {
    clear_error();
    int rc = grids_check_args(L, ngrids, gpu_method, d_state, d_ising_grids, nsweeps);
    if (rc) return rc;
    LaunchShape sh;
    if ((rc = launch_shape(gpu_method, L, ngrids, &sh))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t bytes = grids_ws_lattice_offset(ngrids, !d_mag) + sh.lat_bytes;
    unsigned char *ws = nullptr;
    hipError_t he = hipMallocAsync((void **)&ws, bytes, st);
    if (he != hipSuccess)
        return set_error(he == hipErrorOutOfMemory ? GASP_ERR_NOMEM : GASP_ERR_HIP, "GPUassert: %s %s %d",
                         hipGetErrorString(he), __FILE__, __LINE__);
    rc = sweep_grids_run(L, ngrids, gpu_method, d_state, beta, h, d_ising_grids, nsweeps, d_mag, st, sh, ws);
    he = hipFreeAsync(ws, st);
    if (!rc && he != hipSuccess)
        rc = set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(he), __FILE__, __LINE__);
    return rc;
}

// The same sweeps with the workspace in a file-scope device array instead of
// a per-call allocation, for callers that cannot allocate.  Each device holds
// its own copy of the array for the life of the process; its device address
// is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to kernels.  Calls on one device
// share the array: calls that overlap in time on one device race.  Nothing in
// the workspace needs to start at zero except the counters, which
// set_thresholds clears; the lattice is written whole by the pack kernels.
constexpr size_t kSweepScratchBytes = (size_t)16 << 20;
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kSweepScratchBytes];

static int scratch_for_current_device(unsigned char **out) {
    static std::mutex mu;
    static unsigned char *addr[kScratchMaxDevices] = {};
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchMaxDevices)
        return set_error(GASP_ERR_ARG, "gasp_sweep_grids_noalloc: device %d beyond the %d supported", dev,
                         kScratchMaxDevices);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[dev] = (unsigned char *)p;
    }
    *out = addr[dev];
    return GASP_OK;
}

extern "C" int gasp_sweep_grids_noalloc(int L, int ngrids, int gpu_method, gasp_rng_state *d_state, double beta,
                                        double h, int *d_ising_grids, int nsweeps, float *d_mag, void *stream) {
    clear_error();
    int rc = grids_check_args(L, ngrids, gpu_method, d_state, d_ising_grids, nsweeps);
    if (rc) return rc;
    LaunchShape sh;
    if ((rc = launch_shape(gpu_method, L, ngrids, &sh))) return rc;
    const size_t bytes = grids_ws_lattice_offset(ngrids, !d_mag) + sh.lat_bytes;
    if (bytes > kSweepScratchBytes)
        return set_error(GASP_ERR_ARG,
                         "gasp_sweep_grids_noalloc: gpu_method %d with L=%d, ngrids=%d needs %zu bytes of "
                         "workspace, more than the %zu this entry point holds; use gasp_sweep_grids",
                         gpu_method, L, ngrids, bytes, kSweepScratchBytes);
    unsigned char *ws = nullptr;
    if ((rc = scratch_for_current_device(&ws))) return rc;
    return sweep_grids_run(L, ngrids, gpu_method, d_state, beta, h, d_ising_grids, nsweeps, d_mag,
                           (hipStream_t)stream, sh, ws);
}

// ============================================================ GPU bootstrap
// bootstrap_errbar (src/bootstrap.c:6-63) for many inputs at once, with
// counter-based draws in place of the process-global libc rand() sequence:
// the same estimator (m resamples of n indices with replacement, pB* =
// nB / (nA + nB) over resolved fates, population standard deviation of pB*),
// a different random sequence.  One workgroup per input, its fates in LDS as
// bytes (1 = A, 2 = B, 0 = unresolved); thread t takes resamples t, t + 256,
// ...; index draw k of resample j is word k % 4 of the Philox4x32-10 block
// ctr = (k / 4, j, input, 0x424f4f54), key = seed, mapped as floor(w n / 2^32).
// Per-thread Welford accumulators are combined across the block (Chan et al.).
constexpr uint32_t kBootDomain = 0x424f4f54u;  // "BOOT"

__global__ void __launch_bounds__(256) bootstrap_kernel(const int *__restrict__ fates, int n, int m, uint64_t seed,
                                                        double *__restrict__ out) {
    extern __shared__ __align__(16) unsigned char f[];
    __shared__ double s_cnt[256], s_mean[256], s_m2[256];
    const int tid = threadIdx.x, input = blockIdx.x;
    const int *src = fates + (size_t)input * n;
    for (int i = tid; i < n; i += blockDim.x) f[i] = src[i] == 0 ? 1 : (src[i] == 1 ? 2 : 0);
    __syncthreads();
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    double cnt = 0.0, mean = 0.0, m2 = 0.0;
    for (int j = tid; j < m; j += blockDim.x) {
        int nA = 0, nB = 0;
        for (int i = 0; i < n; i += 4) {
            uint32_t c0 = (uint32_t)(i >> 2), c1 = (uint32_t)j, c2 = (uint32_t)input, c3 = kBootDomain;
            philox4x32_10(c0, c1, c2, c3, k0, k1);
            const uint32_t w[4] = {c0, c1, c2, c3};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (i + q < n) {
                    const unsigned char v = f[((uint64_t)w[q] * (uint32_t)n) >> 32];
                    nA += v == 1;
                    nB += v == 2;
                }
            }
        }
        const double p = (double)nB / (double)(nA + nB);  // NaN without resolved draws, as the reference
        cnt += 1.0;
        const double d = p - mean;
        mean += d / cnt;
        m2 += d * (p - mean);
    }
    s_cnt[tid] = cnt;
    s_mean[tid] = mean;
    s_m2[tid] = m2;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (tid < off) {
            const double na = s_cnt[tid], nb = s_cnt[tid + off];
            const double nt = na + nb;
            if (nb > 0.0) {
                const double d = s_mean[tid + off] - s_mean[tid];
                s_mean[tid] = na > 0.0 ? s_mean[tid] + d * nb / nt : s_mean[tid + off];
                s_m2[tid] = s_m2[tid] + s_m2[tid + off] + (na > 0.0 ? d * d * na * nb / nt : 0.0);
                s_cnt[tid] = nt;
            }
        }
        __syncthreads();
    }
    if (tid == 0) out[input] = sqrt(s_m2[0] / (double)m);
}

extern "C" int gasp_bootstrap_errbar_gpu(const int *fates, int on_device, int ninputs, int n, int m,
                                         unsigned long long seed, double *host_out) {
    clear_error();
    if (!fates || !host_out || ninputs < 1 || n < 1 || m < 1)
        return set_error(GASP_ERR_ARG, "gasp_bootstrap_errbar_gpu: bad argument");
    DeviceInfo di = device_info();
    if (di.ok && (size_t)n > di.lds_block - 3 * 256 * sizeof(double))
        return set_error(GASP_ERR_ARG, "gasp_bootstrap_errbar_gpu: %d fates per input exceed the LDS", n);
    const int *d_f = fates;
    int *d_own = nullptr;
    double *d_out = nullptr;
    hipStream_t st = nullptr;
    hipError_t he = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (he == hipSuccess && !on_device) {
        he = hipMalloc(&d_own, (size_t)ninputs * n * sizeof(int));
        if (he == hipSuccess)
            he = hipMemcpyAsync(d_own, fates, (size_t)ninputs * n * sizeof(int), hipMemcpyHostToDevice, st);
        d_f = d_own;
    }
    if (he == hipSuccess) he = hipMalloc(&d_out, (size_t)ninputs * sizeof(double));
    if (he == hipSuccess && (size_t)n + 3 * 256 * sizeof(double) > 65536)
        he = hipFuncSetAttribute((const void *)bootstrap_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, n);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(bootstrap_kernel, dim3(ninputs), dim3(256), (size_t)n, st, d_f, n, m, (uint64_t)seed,
                           d_out);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipMemcpyAsync(host_out, d_out, (size_t)ninputs * sizeof(double), hipMemcpyDeviceToHost, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    dfree(d_own);
    dfree(d_out);
    if (st) (void)hipStreamDestroy(st);
    if (he != hipSuccess)
        return set_error(he == hipErrorOutOfMemory ? GASP_ERR_NOMEM : GASP_ERR_HIP, "GPUassert: %s %s %d",
                         hipGetErrorString(he), __FILE__, __LINE__);
    return GASP_OK;
}

// ============================================================ batched graph
// One captured hipGraph advancing several engines (different L, method,
// temperature, task) concurrently: the capture stream forks to every engine's
// stream; each branch clears its counters (if it checks), runs `launches`
// sweep kernels of nsweeps/launches sweeps each and advances its attempt
// counter on the device; the branches join back.  A replay is one check
// interval of every engine.
struct gasp_graph {
    std::vector<gasp_engine *> eng;
    std::vector<uint64_t> per_replay;  // attempts per replica per replay
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double ms = 0.0;
    int replays = 0;
};

extern "C" int gasp_graph_destroy(gasp_graph *g) {
    if (!g) return GASP_OK;
    if (g->st) (void)hipStreamSynchronize(g->st);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (hipEvent_t e : g->ev) (void)hipEventDestroy(e);
    if (g->t0) (void)hipEventDestroy(g->t0);
    if (g->t1) (void)hipEventDestroy(g->t1);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
    return GASP_OK;
}

extern "C" int gasp_graph_create(gasp_engine **engines, int n, const int *nsweeps, const int *launches,
                                 const int *checks, const double *dn, const double *up, gasp_graph **out) {
    clear_error();
    if (!engines || n < 1 || !nsweeps || !checks || !dn || !up || !out)
        return set_error(GASP_ERR_ARG, "gasp_graph_create: bad argument");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!engines[i]) return set_error(GASP_ERR_ARG, "gasp_graph_create: NULL engine");
        if (engines[i]->rng.kind != KIND_PHILOX)
            return set_error(GASP_ERR_ARG, "gasp_graph_create: graphs need counter-based (Philox) streams");
        const int k = launches ? launches[i] : 1;
        if (nsweeps[i] < 1 || k < 1 || nsweeps[i] % k != 0)
            return set_error(GASP_ERR_ARG, "gasp_graph_create: nsweeps must be a positive multiple of launches");
        const uint64_t per = (uint64_t)(nsweeps[i] / k) * engines[i]->N;
        if (per >= (1ull << 30)) return set_error(GASP_ERR_ARG, "gasp_graph_create: launch too long");
    }
    gasp_graph *g = new gasp_graph();
#define GTRY(x)                                                                                 \
    do {                                                                                        \
        hipError_t q_ = (x);                                                                    \
        if (q_ != hipSuccess) {                                                                 \
            gasp_graph_destroy(g);                                                              \
            return set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(q_), __FILE__, __LINE__); \
        }                                                                                       \
    } while (0)
    GTRY(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
    GTRY(hipEventCreate(&g->t0));
    GTRY(hipEventCreate(&g->t1));
    for (int i = 0; i < n; ++i) {
        GTRY(engine_sync(engines[i]) == GASP_OK ? hipSuccess : hipErrorUnknown);
        // the device header must hold the host's attempt counter before capture
        GTRY(hipMemcpy(engines[i]->d_rng, &engines[i]->rng, sizeof(RngHeader), hipMemcpyHostToDevice));
        g->eng.push_back(engines[i]);
        g->per_replay.push_back((uint64_t)nsweeps[i] * engines[i]->N);
    }
    hipEvent_t fork;
    GTRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    g->ev.push_back(fork);
    GTRY(hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
    hipError_t ce = hipEventRecord(fork, g->st);
    int rc = GASP_OK;
    for (int i = 0; i < n && ce == hipSuccess && !rc; ++i) {
        gasp_engine *e = engines[i];
        ce = hipStreamWaitEvent(e->stream, fork, 0);
        if (ce != hipSuccess) break;
        const int k = launches ? launches[i] : 1;
        const uint64_t per = (uint64_t)(nsweeps[i] / k) * e->N;
        for (int j = 0; j < k && !rc; ++j) {
            const int chk = (j == k - 1) ? checks[i] : CHECK_NONE;
            rc = launch_sweep(e, per, nullptr, chk, dn[i], up[i], true);
        }
        hipEvent_t join;
        if (ce == hipSuccess) ce = hipEventCreateWithFlags(&join, hipEventDisableTiming);
        if (ce == hipSuccess) {
            g->ev.push_back(join);
            ce = hipEventRecord(join, e->stream);
        }
        if (ce == hipSuccess) ce = hipStreamWaitEvent(g->st, join, 0);
    }
    hipError_t ee = hipStreamEndCapture(g->st, &g->graph);
    if (rc) {
        gasp_graph_destroy(g);
        return rc;
    }
    GTRY(ce);
    GTRY(ee);
    GTRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
#undef GTRY
    *out = g;
    return GASP_OK;
}

extern "C" int gasp_graph_launch(gasp_graph *g, int replays) {
    clear_error();
    if (!g || replays < 0) return set_error(GASP_ERR_ARG, "gasp_graph_launch: bad argument");
    // Direct sweeps on an engine since the last replay advanced only its host
    // counter, and may still be running on its stream: wait for them, then
    // hand the counter to the device header that the graph's kernels read.
    for (gasp_engine *e : g->eng) {
        const int rc = engine_sync(e);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(e->d_rng, &e->rng, sizeof(RngHeader), hipMemcpyHostToDevice, g->st));
    }
    HIP_TRY(hipEventRecord(g->t0, g->st));
    for (int r = 0; r < replays; ++r) HIP_TRY(hipGraphLaunch(g->exec, g->st));
    HIP_TRY(hipEventRecord(g->t1, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, g->t0, g->t1));
    g->ms += ms;
    g->replays += replays;
    for (size_t i = 0; i < g->eng.size(); ++i) g->eng[i]->rng.attempts += (uint64_t)replays * g->per_replay[i];
    return GASP_OK;
}

extern "C" double gasp_graph_ms(gasp_graph *g) { return g ? g->ms : 0.0; }

// ============================================================ autotuner
// Times each candidate launch shape on a throw-away engine (all spins -1,
// Philox, the caller's beta and h, probe_sweeps sweeps after as many warm-up
// sweeps) and keeps the fastest for (method, L, ngrids).  Candidates: slots
// per replica S (lanes that speculate on one replica) and, for methods 0/1,
// waves per workgroup W (for method 2, the word split); for method 0 also
// 128, 256 and 512 slots (a replica over 2, 4 or 8 waves).  Then, with the
// chosen shape and check_sweeps > 0, the sweeps per launch: check_sweeps
// sweeps as one launch or as 2, 4, ... equal launches (launches of at least
// check_sweeps / 8 sweeps), each timed over one check interval after one of
// warm-up; the fastest gives *best_launch_sweeps.
static double probe_rate(gasp_engine *e, const std::vector<int> &grids, int ngrids, int sweeps, int per_launch) {
    auto run = [&]() {
        for (int done = 0; done < sweeps; done += per_launch)
            if (engine_advance(e, (uint64_t)per_launch * e->N, 0, 0, 0)) return false;
        return engine_sync(e) == GASP_OK;
    };
    // warm-up as long as the timed part: from the all-down start the first
    // sweeps accept least, which flatters wide speculation groups
    if (engine_set_grids_impl(e, grids.data(), 0, ngrids) || !run()) return -1.0;
    const double ms0 = e->kernel_ms;
    if (!run()) return -1.0;
    const double ms = e->kernel_ms - ms0;
    return ms > 0 ? (double)sweeps * e->N * ngrids / (ms * 1e-3) : -1.0;
}

extern "C" int gasp_autotune(int method, int L, int ngrids, double beta, double h, int probe_sweeps,
                             int check_sweeps, int *best_slots, int *best_waves, int *best_launch_sweeps,
                             double *best_rate) {
    clear_error();
    if (method < 0 || method > 3 || L < 2 || ngrids < 1 || probe_sweeps < 1 || check_sweeps < 0)
        return set_error(GASP_ERR_ARG, "gasp_autotune: bad argument");
    const int save_s = g_tune_slots, save_w = g_tune_waves;
    std::vector<std::pair<int, int>> cand;
    const int ncu_t = device_info().ok ? device_info().ncu : 256;
    // quarter and eighth words only for few replicas (fewer than two
    // half-word workgroups per CU)
    const bool few = (long long)((ngrids + 63) / 64) * 2 < 2LL * ncu_t;
    for (int S = 1; S <= 64; S <<= 1) {
        if (method == 3) {
            // waves per workgroup x word split
            if (S <= 8) {
                cand.push_back({S, 1});
                cand.push_back({S, 2});
            }
        } else if (method == 2) {
            if (S <= 16) cand.push_back({S, 1});
            if (S >= 2 && S <= 32) cand.push_back({S, 2});
            if (few && S >= 8 && S <= 32) {
                cand.push_back({S, 4});
                cand.push_back({S, 8});
            }
        } else {
            // method 0 with N >= 4096: fewer than 16 lanes per replica read every
            // replica's bytes through L2 at once (at L = 256 with 256 replicas
            // the probes of 1, 2 and 4 lanes took 260, 125 and 30 ms against
            // 2-4 ms for the multi-wave shapes); not probed
            if (method == 0 && L * L >= 4096 && S < 16) continue;
            for (int W = 1; W <= 4; W <<= 1) cand.push_back({S, W});
        }
    }
    if (method == 0) {
        cand.push_back({128, 2});
        cand.push_back({256, 4});
        cand.push_back({512, 8});
        // not 1024 (sixteen waves): alone, the batched L = 256 group's 20 sweeps
        // take 2.60 ms against 2.63 on eight waves (lattice staged in LDS; 3.41
        // against 3.97 from HBM), and in the batched graph its 1024-lane
        // workgroups crowd the other groups' (7.0-7.2 ms per replay against
        // 5.35-5.38 from HBM), while the tuner probes a group alone
    }
    gasp_rng_state *rng = nullptr;
    if (rng_alloc(ngrids, 12345, KIND_PHILOX, &rng)) return gasp_last_error();
    std::vector<int> grids((size_t)ngrids * L * L, -1);
    float pf[20];
    uint64_t thr[10];
    table_f32(beta, h, pf);
    thr_from_table_f(pf, thr);
    double best = -1.0;
    int bs = 1, bw = 1;
    DeviceInfo di = device_info();
    for (auto [S, W] : cand) {
        const size_t lds = method == 3 ? sliced_lds_bytes(L * L) : lds_bytes_for(method, L * L, S, W, ngrids);
        if (di.ok && lds > di.lds_block) continue;
        g_tune_slots = S;
        g_tune_waves = W;
        gasp_engine *e = nullptr;
        if (engine_new(L, ngrids, method, rng, thr, 0, beta, h, &e)) {
            clear_error();
            continue;
        }
        if (e->S != S || e->W != W) {  // shape adjusted by the LDS limit
            engine_free(e);
            continue;
        }
        const double rate = probe_rate(e, grids, ngrids, probe_sweeps, probe_sweeps);
        engine_free(e);
        if (rate > best) {
            best = rate;
            bs = S;
            bw = W;
        }
    }
    int blaunch = check_sweeps;
    if (best > 0 && check_sweeps > 1) {
        g_tune_slots = bs;
        g_tune_waves = bw;
        gasp_engine *e = nullptr;
        if (!engine_new(L, ngrids, method, rng, thr, 0, beta, h, &e)) {
            double best_l = -1.0;
            for (int k = 1; check_sweeps / k >= std::max(1, check_sweeps / 8); k *= 2) {
                if (check_sweeps % k) continue;
                const double rate = probe_rate(e, grids, ngrids, check_sweeps, check_sweeps / k);
                if (rate > best_l) {
                    best_l = rate;
                    blaunch = check_sweeps / k;
                }
            }
            engine_free(e);
        } else {
            clear_error();
        }
    }
    g_tune_slots = save_s;
    g_tune_waves = save_w;
    (void)hipFree(rng);
    if (best < 0) return set_error(GASP_ERR_ARG, "gasp_autotune: no candidate ran");
    g_tuned[std::make_tuple(method, L, ngrids)] = {bs, bw};
    if (best_slots) *best_slots = bs;
    if (best_waves) *best_waves = bw;
    if (best_launch_sweeps) *best_launch_sweeps = blaunch;
    if (best_rate) *best_rate = best;
    return GASP_OK;
}

// ============================================================ device set-up
extern "C" int gpuInitDevice(int deviceIndex, int *gpu_nsms) {
    clear_error();
    int count = 0;
    hipError_t err = hipGetDeviceCount(&count);
    if (count == 0 || err != hipSuccess) {
        emit("No HIP supported devices are available in this system.\n");
        set_error(GASP_ERR_NODEV, "gpuInitDevice: no GPU visible");
        return -1;
    }
    emit("Found %d HIP devices in this system\n", count);
    for (int idev = 0; idev < count; ++idev) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, idev) != hipSuccess) continue;
        emit("Device %d : %s (%s)\n", idev, prop.name, prop.gcnArchName);
        emit("================================\n");
        emit("Number of CUs       : %d\n", prop.multiProcessorCount);
        emit("Max SHMEM per block : %ld KB\n", (long)(prop.sharedMemPerBlock / 1024));
        emit("Max LDS per CU      : %ld KB\n", (long)(prop.maxSharedMemoryPerMultiProcessor / 1024));
        emit("Recommended ngrids  : %d\n", 4 * prop.warpSize * prop.multiProcessorCount);
        emit("\n");
        const long avail = (long)prop.sharedMemPerBlock - kLdsLattice;
        const int maxL = avail > 0 ? (int)floor(sqrt((double)avail / 8.0)) : 0;
        emit("Estimated largest L for shared memory code path : %d\n", maxL);
        emit("(64 replicas per 64-bit multispin word)\n");
        emit("\n");
    }
    if (deviceIndex != -1) {
        err = hipSetDevice(deviceIndex);
        if (err != hipSuccess) {
            set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(err), __FILE__, __LINE__);
            return -1;
        }
    }
    int idev = 0;
    if (hipGetDevice(&idev) != hipSuccess) return -1;
    emit("Using HIP device : %d\n", idev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, idev) == hipSuccess && gpu_nsms) *gpu_nsms = prop.multiProcessorCount;
    return idev;
}

extern "C" void gpuInitGrid(int L, int ngrids, int threadsPerBlock, int *ising_grids, int **d_ising_grids,
                            int **d_neighbour_list) {
    (void)threadsPerBlock;
    clear_error();
    if (!d_ising_grids || !d_neighbour_list || !ising_grids || L < 1 || ngrids < 1) {
        set_error(GASP_ERR_ARG, "gpuInitGrid: bad argument");
        return;
    }
    *d_ising_grids = nullptr;
    *d_neighbour_list = nullptr;
    const size_t n = (size_t)L * L * ngrids;
    hipError_t e = hipMalloc((void **)d_ising_grids, n * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(*d_ising_grids, ising_grids, n * sizeof(int), hipMemcpyHostToDevice);
    // neighbour table of src/mc_gpu.cu:39-58 (down, up, right, left per site)
    std::vector<int> nb((size_t)4 * L * L);
    for (int s = 0; s < L * L; ++s) {
        const int row = s / L, col = s % L;
        nb[4 * s + 0] = L * ((row + 1) % L) + col;
        nb[4 * s + 1] = L * ((row + L - 1) % L) + col;
        nb[4 * s + 2] = L * row + (col + 1) % L;
        nb[4 * s + 3] = L * row + (col + L - 1) % L;
    }
    if (e == hipSuccess) e = hipMalloc((void **)d_neighbour_list, nb.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(*d_neighbour_list, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(e), __FILE__, __LINE__);
}

extern "C" void preComputeProbs_gpu(double beta, double h) {
    table_f32(beta, h, g_pacc_f32);
    thr_from_table_f(g_pacc_f32, g_thr_f32);
}

// src/gpu_tools.cu:205-232: the largest L whose multispin lattice the
// reference fits in one block's shared memory, maxL = floor(sqrt(32 *
// floor(m^2 / 32))) with m = floor(sqrt(8 shmem / tpb)) (201 at 32 threads
// per block and 142 at 64 on a 160 KB device); 2 up to maxL, else 0.  The
// engine's method 2 itself runs lattices beyond maxL (split words), so the
// choice is only the reference's.
extern "C" int select_gpu_method(int L, int ngrids, int threadsPerBlock, int gpu_device) {
    (void)ngrids;
    if (threadsPerBlock < 1) {
        set_error(GASP_ERR_ARG, "select_gpu_method: threadsPerBlock must be positive");
        return 0;
    }
    int dev = gpu_device;
    if (dev < 0 && hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t prop;
    size_t shmem = 65536;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) shmem = prop.sharedMemPerBlock;
    return select_method_for(L, (long long)shmem, threadsPerBlock);
}

// ============================================================ driver
// Schedule of src/mc_gpu.cu:654-778.  Per iteration (isweep a multiple of
// spc = min(mag, grid)): grids to the host if isweep % grid == 0; the
// collective variable of the state at isweep (computed in the epilogue of the
// previous sweep launch, or by the initial measurement) if isweep % mag == 0;
// the next sweep launch; the output callback; the threshold report, which
// may stop the run.  The reference launches one more sweep before it stops;
// that sweep is not observable through the API and is skipped here.
extern "C" void mc_driver_gpu(mc_gpu_grids_t grids, double beta, double h, int *grid_fate, mc_sampler_t samples,
                              mc_function_t calc, gpu_run_t gpu_state, GridOutputFunc outfunc) {
    clear_error();
    const int L = grids.L, R = grids.ngrids;
    const int tot = samples.tot_nsweeps, mag = samples.mag_output_int, gout = samples.grid_output_int;
    const int itask = calc.itask;
    const double dn = calc.dn_thr, up = calc.up_thr;
    const int ninputs = calc.ninputs;
    if (ninputs < 1 || R % ninputs != 0) {
        set_error(GASP_ERR_ARG, "Error: ngrids must be divisible by ninputs!");
        return;
    }
    if (mag < 1 || gout < 1) {
        set_error(GASP_ERR_ARG, "Error: mag_output_int and grid_output_int must be positive");
        return;
    }
    if (itask != 0 && itask != 1) {
        set_error(GASP_ERR_ARG, "Error: itask must be 0 or 1!");
        return;
    }
    if (gpu_state.gpu_method < 0 || gpu_state.gpu_method > 2) {
        set_error(GASP_ERR_ARG, "Unknown gpu_method in ising.cu");
        return;
    }
    if (gpu_state.gpu_method == 2 && gpu_state.threadsPerBlock != 32 && gpu_state.threadsPerBlock != 64) {
        set_error(GASP_ERR_ARG, "Invalid threadsPerBlock for gpu_method=2");
        return;
    }
    if (!grids.d_ising_grids || !grids.ising_grids || (itask == 1 && !grid_fate) || !calc.result) {
        set_error(GASP_ERR_ARG, "Error: mc_driver_gpu received a NULL array");
        return;
    }
    const bool use_cluster = calc.cv && strcmp(calc.cv, "largest_cluster") == 0;
    const int sub = R / ninputs;
    const int spc = mag < gout ? mag : gout;
    const int result_size = itask == 0 ? tot / mag : ninputs;

    gasp_engine *e = nullptr;
    if (engine_new(L, R, gpu_state.gpu_method, gpu_state.d_state, g_thr_f32, 0, beta, h, &e)) return;

    std::vector<float> magnetisation(R, 0.0f), lclus(use_cluster ? R : 0, 0.0f);
    std::vector<float> result(std::max(result_size, 1), 0.0f);
    if (itask == 0)
        for (int i = 0; i < result_size; ++i) result[i] = 1.0f;
    const int check_mode = itask == 0 ? CHECK_COUNT : CHECK_FATES;
    int counts[4] = {0, 0, 0, 0};
    int rc = engine_set_grids_impl(e, grids.d_ising_grids, 1, R);
    if (!rc && itask == 1)
        rc = hipMemcpyAsync(e->d_fate, grid_fate, (size_t)R * sizeof(int), hipMemcpyHostToDevice, e->stream) ==
                     hipSuccess
                 ? GASP_OK
                 : set_error(GASP_ERR_HIP, "GPUassert: fate upload");
    // collective variable of the initial state (isweep = 0 is always a check)
    if (!rc) rc = launch_sweep(e, 0, nullptr, use_cluster ? CHECK_NONE : check_mode, dn, up);
    // With the library's own gridstates.bin writer as the output function, grid
    // dumps are packed to bits on the device and only the bits are copied and
    // appended (the same records); the caller's host array receives the last
    // dumped state at the end, which is what the reference leaves in it.
    const bool packed_out = outfunc == write_ising_grids;
    const size_t pbytes = ((size_t)R * L * L + 7) / 8;
    uint8_t *d_bits = nullptr, *h_bits = nullptr;
    bool dumped = false;
    if (!rc && packed_out) {
        if (hipMalloc(&d_bits, pbytes) != hipSuccess || hipHostMalloc(&h_bits, pbytes) != hipSuccess)
            rc = set_error(GASP_ERR_NOMEM, "Error: no memory for grid dumps");
    }
    const double t_ms0 = e->kernel_ms;
    const auto t0 = std::chrono::steady_clock::now();
    const bool py = output_style() == 1;

    int isweep = 0;
    while (!rc && isweep < tot) {
        const bool do_grid = isweep % gout == 0;
        const bool do_mag = isweep % mag == 0;
        if ((rc = engine_sync(e))) break;
        if (do_mag && use_cluster) {
            if ((rc = engine_cluster(e, -calc.initial_spin))) break;
            if ((rc = engine_check_cv(e, e->d_lclus, check_mode, dn, up))) break;
        }
        if (do_grid && packed_out) {
            hipLaunchKernelGGL(pack_gridstates, dim3(2048), dim3(256), 0, e->stream, (const void *)e->d_lat, layout_of(e->method),
                               e->stride, R, e->N, d_bits, pbytes);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(h_bits, d_bits, pbytes, hipMemcpyDeviceToHost, e->stream) != hipSuccess) {
                rc = set_error(GASP_ERR_HIP, "GPUassert: grid download");
                break;
            }
        } else if (do_grid) {
            if ((rc = engine_unpack<int>(e, grids.d_ising_grids))) break;
            if (hipMemcpyAsync(grids.ising_grids, grids.d_ising_grids, (size_t)R * L * L * sizeof(int),
                               hipMemcpyDeviceToHost, e->stream) != hipSuccess) {
                rc = set_error(GASP_ERR_HIP, "GPUassert: grid download");
                break;
            }
        }
        if (do_mag) {
            if (hipMemcpyAsync(magnetisation.data(), e->d_mag, (size_t)R * sizeof(float), hipMemcpyDeviceToHost,
                               e->stream) != hipSuccess ||
                hipMemcpyAsync(counts, e->d_counts, 3 * sizeof(int), hipMemcpyDeviceToHost, e->stream) !=
                    hipSuccess ||
                (use_cluster && hipMemcpyAsync(lclus.data(), e->d_lclus, (size_t)R * sizeof(float),
                                               hipMemcpyDeviceToHost, e->stream) != hipSuccess)) {
                rc = set_error(GASP_ERR_HIP, "GPUassert: collective variable download");
                break;
            }
        }
        if ((rc = engine_sync(e))) break;
        const bool report = do_mag || isweep == tot - 1;
        bool stop = false;
        if (report) stop = itask == 0 ? counts[0] == R : counts[1] + counts[2] == R;
        if (!stop) {
            const int next = isweep + spc;
            const int chk = (!use_cluster && next % mag == 0) ? check_mode : CHECK_NONE;
            if ((rc = engine_advance(e, (uint64_t)spc * (uint64_t)e->N, chk, dn, up))) break;
        }
        if (do_grid || isweep == tot - 1) {
            if (packed_out) {
                // return value ignored, as the reference ignores outfunc's
                append_gridstates(L, R, isweep, h_bits, calc.filename);
                dumped = true;
            } else if (outfunc) {
                outfunc(L, R, grids.ising_grids, isweep, magnetisation.data(), use_cluster ? lclus.data() : nullptr,
                        calc.cv, dn, up, calc.filename);
            }
        }
        if (report) {
            if (itask == 0) {
                const int nnuc = counts[0];
                const int k = isweep / mag;
                if (k < result_size) result[k] = (float)((double)nnuc / (double)R);
                if (py)
                    emit("\r Sweep : %10d, Reached cv = %6.2f : %4d , Unresolved : %4d", isweep, up, nnuc, R - nnuc);
                else
                    emit("%10d  %12.6f\n", isweep, (double)nnuc / (double)R);
            } else {
                const int nA = counts[1], nB = counts[2];
                emit("\r Sweep : %10d, Reached cv = %6.2f : %4d , Reached cv = %6.2f : %4d , Unresolved : %4d", isweep,
                     dn, nA, up, nB, R - nA - nB);
            }
            if (stop) break;
        }
        isweep += spc;
    }
    if (!rc) rc = engine_sync(e);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // final state back into the caller's device grids
    if (!rc) rc = engine_unpack<int>(e, grids.d_ising_grids);
    if (!rc && itask == 1)
        rc = hipMemcpyAsync(grid_fate, e->d_fate, (size_t)R * sizeof(int), hipMemcpyDeviceToHost, e->stream) ==
                     hipSuccess
                 ? GASP_OK
                 : set_error(GASP_ERR_HIP, "GPUassert: fate download");
    if (!rc) rc = engine_store_rng(e);
    t_stats.sweep_kernel_ms = e->kernel_ms - t_ms0;
    t_stats.attempts_per_replica = (double)(e->rng.attempts);
    t_stats.launches = e->launches;
    t_stats.slots_per_replica = e->S;
    t_stats.waves_per_group = e->W;
    t_stats.method = e->method;
    engine_free(e);
    if (!rc && dumped) {
        const size_t n = (size_t)R * L * L;
        for (size_t i = 0; i < n; ++i) grids.ising_grids[i] = ((h_bits[i >> 3] >> (i & 7)) & 1) ? 1 : -1;
    }
    dfree(d_bits);
    if (h_bits) (void)hipHostFree(h_bits);
    if (rc) return;

    if (itask == 0) {
        for (int i = 0; i < result_size; ++i) calc.result[i] = result[i];
    } else {
        // src/mc_gpu.cu:812-826
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
        emit("\n");
    } else {
        emit("\n# Time taken on GPU = %f seconds\n", secs);
        if (itask == 1) emit("pB estimate : %10.6f\n", (double)calc.result[0]);
    }
}
