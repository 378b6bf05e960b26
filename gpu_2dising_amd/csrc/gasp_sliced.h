// gasp_sliced.h -- bit-sliced multispin sweep (engine method 3).
//
// The replicas of a multispin word (64, or 32 with split words) share one
// random site sequence: attempt a of word group g visits one site for all of
// them, reads the five stencil words of the LDS lattice, forms every replica's
// number of up neighbours with bitwise adders and decides every replica's
// acceptance at once by a bit-sliced comparison of its 32-bit acceptance draw
// with its class threshold.  The draw of replica b is assembled from random
// bit planes, most significant first: plane p holds bit 31 - p of every
// replica's draw, so the comparison of draw and threshold is decided, for
// each replica, at the first plane where the two bits differ, and planes are
// generated only while some replica of the wave is undecided.  Each replica
// still follows random-site Metropolis with its own acceptance draws
// (src/mc_gpu.cu:358-434 gives every replica its own site instead); the
// replicas of a word are correlated through their common sites.
//
// Random stream (Philox4x32-10, key = seed): block k of attempt a of group g
// is philox(ctr = (a_lo, a_hi, g, kSlicedDomain + k)); its words, in order,
// continue one word stream w_0, w_1, ... (block k holds w_4k .. w_4k+3).
// Site = floor(w_0 N / 2^32).  With q = WB / 32 dwords per plane, plane p
// (p = 0 .. 31) is the dwords w_q(p+1) .. w_q(p+1)+q-1, bit b of the plane
// (bit b & 31 of its dword b / 32) belonging to replica WB g + b.  Replica b
// flips iff its draw sum_p bit_b(plane p) 2^(31-p) is below the threshold of
// its class (8 s01 + nup as in the other kernels).
//
// Parallel evaluation with exact sequential results.  A workgroup of W waves
// serves one group: its 64 W lanes evaluate 64 W attempts of the chain against
// the lattice as it was at the start of the round ("the window").  Attempt j
// depends on an earlier attempt i iff i's site lies in j's five-site
// stencil; j is committed in the round iff no earlier attempt of the window
// has its site in j's stencil, otherwise it is deferred and evaluated again
// in the next round, before any newer attempt.  Committed attempts read
// exactly what the sequential chain would have shown them: every earlier
// attempt they depend on had committed in an earlier round, and the ones
// they do not depend on neither read nor write what they read or write, so
// the order of independent attempts does not matter.  The window's earliest
// attempt always commits, so every round makes progress.  Dependences are
// found with a per-site table: each attempt writes its launch-local index into
// its site's entry with an atomic minimum, and an attempt whose five stencil
// entries hold a smaller index has an earlier attempt in its stencil.  Two
// tables alternate by round; every lane clears, during the next round, the
// entry it wrote, and the barriers between rounds order those clears before
// the table is used again.
#pragma once

#include "gasp_kernels.h"

namespace gasp {

constexpr uint32_t kSlicedDomain = 0x534C4300u;  // "SLC" + block index

// LDS of the sliced kernel: deferred counts [2][8] (64 B), per-replica up
// counts (64 x 4 B), then 16 bytes per site: the word (dwords 0 .. q - 1) and
// the two tables' entries (dwords 2 and 3).
constexpr int kSlCnt = 0;
constexpr int kSlUp = 64;
constexpr int kSlLat = 320;
constexpr int kSlEntry = 16;

// The classes with a threshold below 2^32 (at most five: for any beta and h,
// the classes whose energy change is positive are the down spins with nup
// below (4 - h) / 2 and the up spins with nup above it).  code = 8 s01 + nup,
// or 7 for an unused slot (no replica has nup = 7, so its mask is empty).
struct SlicedSlots {
    uint32_t code[5];
    uint32_t thr[5];
};

__device__ __forceinline__ SlicedSlots sliced_slots(const uint64_t *thr) {
    SlicedSlots s;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        s.code[i] = 7u;
        s.thr[i] = 0u;
    }
    int n = 0;
#pragma unroll
    for (int j = 0; j < 10; ++j) {
        const uint64_t t = thr[j];
        const uint32_t code = j < 5 ? (uint32_t)j : (uint32_t)(8 + j - 5);
        if (t < (1ull << 32)) {
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i == n) {
                    s.code[i] = code;
                    s.thr[i] = (uint32_t)t;
                }
            ++n;
        }
    }
    return s;
}

// v_bitop3_b32 truth tables (inputs A, B, C = 0xF0, 0xCC, 0xAA)
constexpr int kXor3 = 0x96;        // A ^ B ^ C
constexpr int kAndAndNot = 0x40;   // A & B & ~C
constexpr int kAndXnor = 0x90;     // A & ~(B ^ C)
constexpr int kNor3 = 0x01;        // ~(A | B | C)

// a & m and (a & m) | c for a wave-uniform mask m (all ones or zero) held in
// a scalar register: one v_and_b32 / v_and_or_b32.  Written out because the
// compiler otherwise turns the known all-or-nothing mask into a select.
__device__ __forceinline__ uint32_t and_s(uint32_t a, uint32_t m) {
    uint32_t r;
    asm("v_and_b32 %0, %1, %2" : "=v"(r) : "s"(m), "v"(a));
    return r;
}
__device__ __forceinline__ uint32_t and_or_s(uint32_t a, uint32_t m, uint32_t c) {
    uint32_t r;
    asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(m), "v"(c));
    return r;
}

// Two Philox4x32-10 blocks with the same key, their rounds interleaved.
__device__ __forceinline__ void philox4x32_10x2(uint32_t (&x)[4], uint32_t (&y)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t px0 = (uint64_t)kPhiloxM0 * x[0], px1 = (uint64_t)kPhiloxM1 * x[2];
        const uint64_t py0 = (uint64_t)kPhiloxM0 * y[0], py1 = (uint64_t)kPhiloxM1 * y[2];
        const uint32_t xn0 = __builtin_amdgcn_bitop3_b32((uint32_t)(px1 >> 32), x[1], k0, kXor3);
        const uint32_t xn2 = __builtin_amdgcn_bitop3_b32((uint32_t)(px0 >> 32), x[3], k1, kXor3);
        const uint32_t yn0 = __builtin_amdgcn_bitop3_b32((uint32_t)(py1 >> 32), y[1], k0, kXor3);
        const uint32_t yn2 = __builtin_amdgcn_bitop3_b32((uint32_t)(py0 >> 32), y[3], k1, kXor3);
        x[1] = (uint32_t)px1;
        x[3] = (uint32_t)px0;
        x[0] = xn0;
        x[2] = xn2;
        y[1] = (uint32_t)py1;
        y[3] = (uint32_t)py0;
        y[0] = yn0;
        y[2] = yn2;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
}

template <int WB, int W, bool POW2>
__global__ void __launch_bounds__(64 * W) sweep_sliced(SweepArgs a) {
    static_assert(WB == 64 || WB == 32, "64 or 32 replicas per word");
    static_assert(W == 1 || W == 2 || W == 4 || W == 8, "1, 2, 4 or 8 waves");
    constexpr int Q = WB / 32;  // dwords per word
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int N = a.N;
    const int g = (int)blockIdx.x;  // group of WB replicas
    const int r0 = g * WB;
    const bool own = lane < WB && r0 + lane < a.R;
    if (lds_origin(smem) != 0u || (a.rng_dev && !rng_header_ok(a.rng_dev, a.R))) {
        if (wave == 0 && own) a.mag[r0 + lane] = __builtin_nanf("");
        return;
    }

    // replicas that may flip: present, and not frozen with a decided fate
    bool may = own && !(a.freeze && a.fate[r0 + lane] != -1);
    const uint64_t may_mask = wave_mask(may);
    uint32_t allow[Q];
#pragma unroll
    for (int d = 0; d < Q; ++d) allow[d] = (uint32_t)(may_mask >> (32 * d));

    // ---------------------------------------------------------- prologue
    // HBM: the 64-replica group's two dword planes, a.stride dwords each;
    // workgroup g holds both (WB = 64) or plane g (WB = 32)
    const unsigned *plane0 = (const unsigned *)a.lattice + (size_t)(Q == 2 ? 2 * g : g) * a.stride;
    for (int x = tid; x < N; x += 64 * W) {
        const unsigned lo = plane0[x];
        const unsigned hi = Q == 2 ? plane0[a.stride + x] : 0u;
        lds_u64 *e = (lds_u64 *)(uintptr_t)(kSlLat + kSlEntry * (unsigned)x);
        e[0] = ((uint64_t)hi << 32) | (uint64_t)lo;
        e[1] = ~0ull;  // both tables' entries: empty
    }
    if (tid < 64) lds_ptr32(kSlUp)[tid] = 0u;
    __syncthreads();

    const uint64_t a_begin = a.rng_dev ? a.rng_dev->attempts : a.a_begin;
    const uint64_t seed = a.rng_dev ? a.rng_dev->seed : a.seed;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const SlicedSlots sl = sliced_slots(a.thr);
    const Geometry<POW2, kSlEntry> geo{a.L, N, a.lg};

    // ---------------------------------------------------------- rounds
    const uint32_t n = may_mask ? (uint32_t)a.n_attempts : 0u;
    const int t = wave * 64 + lane;
    bool has = (uint32_t)t < n;
    uint32_t l = (uint32_t)t;                 // launch-local attempt index of this lane
    uint32_t next = min((uint32_t)(64 * W), n);
    unsigned par = 0u, prev_pos = 0u;
    bool had_prev = false;
    for (;;) {
        Stencil st{};
        uint32_t cw[Q], flip[Q];
#pragma unroll
        for (int d = 0; d < Q; ++d) cw[d] = flip[d] = 0u;
        uint32_t und[Q], acc[Q];
        uint32_t blk = 0u;  // Philox block of the planes in w[]
        uint32_t w[4];
        const uint64_t A = a_begin + (uint64_t)l;
        if (has) {
            w[0] = (uint32_t)A;
            w[1] = (uint32_t)(A >> 32);
            w[2] = (uint32_t)g;
            w[3] = kSlicedDomain;
            philox4x32_10(w[0], w[1], w[2], w[3], k0, k1);
            st = geo.philox(w[0]);
            const unsigned base = kSlLat + (unsigned)st.c;
            __hip_atomic_fetch_min(lds_ptr32(base + 8u + 4u * par), l, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            uint64_t c64, u64, d64, l64, r64;
            if constexpr (Q == 2) {
                c64 = lds_load64(base);
                u64 = lds_load64(kSlLat + (unsigned)st.up);
                d64 = lds_load64(kSlLat + (unsigned)st.dn);
                l64 = lds_load64(kSlLat + (unsigned)st.lt);
                r64 = lds_load64(kSlLat + (unsigned)st.rt);
            } else {
                c64 = lds_load32(base);
                u64 = lds_load32(kSlLat + (unsigned)st.up);
                d64 = lds_load32(kSlLat + (unsigned)st.dn);
                l64 = lds_load32(kSlLat + (unsigned)st.lt);
                r64 = lds_load32(kSlLat + (unsigned)st.rt);
            }
            // Blocks 1 and 2 hold planes that nearly every round needs (planes
            // 1-4 of whole words, 3-10 of half words); they do not depend on
            // the lattice, so they are computed while its reads are in flight
            // (two independent chains side by side).
            uint32_t v1[4] = {(uint32_t)A, (uint32_t)(A >> 32), (uint32_t)g, kSlicedDomain + 1u};
            uint32_t v2[4] = {(uint32_t)A, (uint32_t)(A >> 32), (uint32_t)g, kSlicedDomain + 2u};
            philox4x32_10x2(v1, v2, k0, k1);
            uint32_t up[Q], dn[Q], lt[Q], rt[Q];
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                cw[d] = (uint32_t)(c64 >> (32 * d));
                up[d] = (uint32_t)(u64 >> (32 * d));
                dn[d] = (uint32_t)(d64 >> (32 * d));
                lt[d] = (uint32_t)(l64 >> (32 * d));
                rt[d] = (uint32_t)(r64 >> (32 * d));
            }
            // slot masks: the replicas whose class is slot i's
            uint32_t ms[5][Q];
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                // number of up neighbours as bits b2 b1 b0 (bitwise adders)
                const uint32_t t1 = up[d] ^ dn[d], c1 = up[d] & dn[d];
                const uint32_t t2 = lt[d] ^ rt[d], c2 = lt[d] & rt[d];
                const uint32_t b0 = t1 ^ t2;
                const uint32_t b1 = __builtin_amdgcn_bitop3_b32(c1, c2, t1 & t2, kXor3);
                const uint32_t b2 = c1 & c2;
                uint32_t any = 0u;
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    // all-ones where the bit of the code is set (uniform)
                    const uint32_t v0 = 0u - (sl.code[i] & 1u), vb1 = 0u - ((sl.code[i] >> 1) & 1u),
                                   vb2 = 0u - ((sl.code[i] >> 2) & 1u), vs = 0u - ((sl.code[i] >> 3) & 1u);
                    const uint32_t eq = __builtin_amdgcn_bitop3_b32(b0 ^ v0, b1 ^ vb1, b2 ^ vb2, kNor3);
                    ms[i][d] = eq & ~(cw[d] ^ vs);
                    any |= ms[i][d];
                }
                und[d] = any & allow[d];
                acc[d] = ~any & allow[d];  // classes at threshold 2^32 always flip
            }
            // Plane p: Tj = the replicas whose class threshold has bit 31 - p
            // set, the union of the slot masks selected by uniform all-ones /
            // zero masks (one v_and_or_b32 per slot and dword); then a replica
            // still undecided flips if Tj is set where the draw's bit is not,
            // and stays undecided where the two bits agree.
            // The planes are applied in order, so the thresholds are carried
            // shifted left by the plane index: bit 31 holds plane p's bit, and
            // an arithmetic shift by 31 gives the mask (scalar operations).
            uint32_t tsh[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) tsh[i] = sl.thr[i];
            auto plane = [&](int p, const uint32_t *rnd) {
                (void)p;
                uint32_t sel[5];
#pragma unroll
                for (int i = 0; i < 5; ++i) {
                    sel[i] = (uint32_t)((int32_t)tsh[i] >> 31);
                    tsh[i] <<= 1;
                }
#pragma unroll
                for (int d = 0; d < Q; ++d) {
                    uint32_t tj = and_s(ms[0][d], sel[0]);
#pragma unroll
                    for (int i = 1; i < 5; ++i) tj = and_or_s(ms[i][d], sel[i], tj);
                    acc[d] |= __builtin_amdgcn_bitop3_b32(und[d], tj, rnd[d], kAndAndNot);
                    und[d] = __builtin_amdgcn_bitop3_b32(und[d], tj, rnd[d], kAndXnor);
                }
            };
            int p;
            if constexpr (Q == 2) {
                plane(0, w + 2);
                plane(1, v1);
                plane(2, v1 + 2);
                plane(3, v2);
                plane(4, v2 + 2);
                p = 5;
            } else {
                plane(0, w + 1);
                plane(1, w + 2);
                plane(2, w + 3);
#pragma unroll
                for (int k = 0; k < 4; ++k) plane(3 + k, v1 + k);
#pragma unroll
                for (int k = 0; k < 4; ++k) plane(7 + k, v2 + k);
                p = 11;
            }
            // two further blocks at a time while any replica of the wave is
            // undecided (a block past the deciding one changes nothing)
            uint32_t pend = und[0];
#pragma unroll
            for (int d = 1; d < Q; ++d) pend |= und[d];
            blk = 2u;
            while (wave_mask(pend != 0u) != 0ull && p < 32) {
                uint32_t va[4] = {(uint32_t)A, (uint32_t)(A >> 32), (uint32_t)g, kSlicedDomain + blk + 1u};
                uint32_t vb[4] = {(uint32_t)A, (uint32_t)(A >> 32), (uint32_t)g, kSlicedDomain + blk + 2u};
                philox4x32_10x2(va, vb, k0, k1);
                blk += 2u;
                if constexpr (Q == 2) {
                    plane(p, va);  // p <= 29 here
                    plane(p + 1, va + 2);
                    if (p + 2 < 32) plane(p + 2, vb);
                    if (p + 3 < 32) plane(p + 3, vb + 2);
                    p += 4;
                } else {
#pragma unroll
                    for (int k = 0; k < 4; ++k) plane(p + k, va + k);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (p + 4 + k < 32) plane(p + 4 + k, vb + k);
                    p += 8;
                }
                pend = und[0];
#pragma unroll
                for (int d = 1; d < Q; ++d) pend |= und[d];
            }
#pragma unroll
            for (int d = 0; d < Q; ++d) flip[d] = acc[d];
        }
        if (had_prev) lds_ptr32(kSlLat + prev_pos + 8u + 4u * (par ^ 1u))[0] = ~0u;
        __syncthreads();
        bool deferred = false;
        if (has) {
            const unsigned o = 8u + 4u * par;
            const uint32_t m = min(min(min(lds_load32(kSlLat + (unsigned)st.c + o), lds_load32(kSlLat + (unsigned)st.up + o)),
                                       min(lds_load32(kSlLat + (unsigned)st.dn + o), lds_load32(kSlLat + (unsigned)st.lt + o))),
                                   lds_load32(kSlLat + (unsigned)st.rt + o));
            deferred = m < l;
            uint32_t fo = 0u;
#pragma unroll
            for (int d = 0; d < Q; ++d) fo |= flip[d];
            if (!deferred && fo) {
                if constexpr (Q == 2)
                    *(lds_u64 *)(uintptr_t)(kSlLat + (unsigned)st.c) =
                        ((uint64_t)(cw[1] ^ flip[1]) << 32) | (uint64_t)(cw[0] ^ flip[0]);
                else
                    lds_ptr32(kSlLat + (unsigned)st.c)[0] = cw[0] ^ flip[0];
            }
        }
        const uint64_t dmask = wave_mask(deferred);
        if (lane == 0) lds_ptr32(kSlCnt + 4u * (8u * par + (unsigned)wave))[0] = (uint32_t)__popcll(dmask);
        had_prev = has;
        prev_pos = (unsigned)st.c;
        __syncthreads();
        // deferred attempts keep their lanes; the others take the next new
        // attempts in wave order, lane order
        uint32_t D = 0u, before = 0u;
#pragma unroll
        for (int v = 0; v < W; ++v) {
            const uint32_t dv = lds_load32(kSlCnt + 4u * (8u * par + (unsigned)v));
            D += dv;
            if (v < wave) before += 64u - dv;
        }
        const uint32_t rem = n - next;
        if (D == 0u && rem == 0u) break;
        if (!deferred) {
            const uint64_t free_mask = ~dmask;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(free_mask >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)free_mask, 0u));
            const uint32_t idx = before + rank;
            has = idx < rem;
            l = next + idx;
        }
        next += min((uint32_t)(64 * W) - D, rem);
        par ^= 1u;
    }

    // ---------------------------------------------------------- epilogue
    // up spins per replica: wave v counts sites v, v + W, ...; lane b bit b
    __syncthreads();
    if (lane < WB) {
        unsigned cnt = 0u;
        for (int x = wave; x < N; x += W) {
            const unsigned dw = lds_load32(kSlLat + kSlEntry * (unsigned)x + 4u * (unsigned)(lane >> 5));
            cnt += (dw >> (lane & 31)) & 1u;
        }
        __hip_atomic_fetch_add(lds_ptr32(kSlUp + 4u * (unsigned)lane), cnt, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    if (wave == 0) finish_replica(a, own, r0 + lane, own ? lds_load32(kSlUp + 4u * (unsigned)lane) : 0u);
    unsigned *out0 = (unsigned *)a.lattice + (size_t)(Q == 2 ? 2 * g : g) * a.stride;
    for (int x = tid; x < N; x += 64 * W) {
        const unsigned e = kSlLat + kSlEntry * (unsigned)x;
        out0[x] = lds_load32(e);
        if constexpr (Q == 2) out0[a.stride + x] = lds_load32(e + 4u);
    }
}

}  // namespace gasp
