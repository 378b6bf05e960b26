// gasp_kernels.h -- gfx950 device code of the replica-parallel Metropolis engine.
//
// One sweep kernel template covers the three lattice layouts of the reference's
// gpu_method 0/1/2 (src/mc_gpu.cu:88-151, 173-275, 279-434), the two random
// streams (counter-based Philox4x32-10, or a pre-generated per-replica MT19937
// draw stream that reproduces src/mc_cpu.c:41-89), and a speculation width S
// up to one wave; sweep_hbm_waves extends layout 0 to a replica spread over
// two, four or eight waves of a workgroup (S = 128, 256, 512).
//
// Dynamics (identical in every layout).  Attempt a of a replica picks a site,
// reads the spin s and its four periodic neighbours, forms
// idx = 5*(s+1) + nsum + 4 (src/mc_cpu.c:65), and flips s iff the 32-bit
// acceptance draw u satisfies u < T[idx], where T is the exact integer image of
// the reference's acceptance table (xi = (u+0.5)*2^-32 < P  <=>  u < T).
//
// Speculation.  A replica is a sequential Markov chain.  S lanes of one wave
// serve one replica and each lane evaluates two consecutive attempts (an
// attempt pair shares one Philox block), so one round evaluates 2*S attempts
// a0 .. a0+2S-1 against the current lattice.  The attempts are then committed
// in order up to (excluding) the first one whose 5-site stencil contains the
// site of an earlier accepted attempt of the same round: every committed
// attempt saw exactly the lattice the sequential chain would have shown it,
// so the trajectory is the sequential one, bit for bit.  The first
// conflicting attempt is redone in the next round (counter-based draws make
// that free).  S = 1 is the plain one-lane-per-replica chain.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gasp {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;
constexpr uint32_t kPhiloxDomain = 0x47415350u;  // "GASP"

enum { LAT_HBM = 0, LAT_BITS = 1, LAT_MULTISPIN = 2, LAT_MULTIHALF = 3, LAT_MULTI16 = 4, LAT_MULTI8 = 5, LAT_LDS8 = 6,
       LAT_LDS1 = 7 };
enum { RNG_PHILOX = 0, RNG_DRAWS = 1 };
enum { CHECK_NONE = 0, CHECK_COUNT = 1, CHECK_FATES = 2 };

// LDS header shared by all layouts: the 10 thresholds as entries 8 s01 + nup
// of a 16-entry table (128 B; s01 = centre spin bit, nup = up neighbours), 64
// per-replica counters and the 64 row/column neighbourhood masks of the
// conflict filter plus a zero entry.
constexpr int kLdsThr = 0;
constexpr int kLdsCnt = 128;
constexpr int kLdsNbr = 384;
constexpr int kLdsLattice = 912;  // 384 + 65 masks of 8 bytes, rounded up to 16
// Fourth dword of a raw buffer resource on gfx9 parts (no format conversion;
// byte offsets checked against the record count).
constexpr int kBufferWord3 = 0x00020000;

// LDS by absolute address.  The sweep kernels declare no static LDS, so their
// dynamic LDS (smem) starts at LDS address 0 and smem + x is address x.  A
// pointer formed from the integer lets the compiler fold constant offsets
// into the ds instruction; offsets from smem itself cost an add of its base,
// which is only resolved after instruction selection.
typedef __attribute__((address_space(3))) unsigned lds_u32;
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(3))) unsigned char lds_u8;
typedef __attribute__((address_space(3))) unsigned short lds_u16;
__device__ __forceinline__ lds_u32 *lds_ptr32(unsigned addr) { return (lds_u32 *)(uintptr_t)addr; }
__device__ __forceinline__ unsigned lds_load32(unsigned addr) { return *lds_ptr32(addr); }
__device__ __forceinline__ uint64_t lds_load64(unsigned addr) { return *(const lds_u64 *)(uintptr_t)addr; }
__device__ __forceinline__ void lds_xor32(unsigned addr, unsigned v) {
    __hip_atomic_fetch_xor(lds_ptr32(addr), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// LDS address of smem (0 unless the kernel gains static LDS)
__device__ __forceinline__ unsigned lds_origin(unsigned char *smem) {
    return (unsigned)(uintptr_t)(lds_u8 *)smem;
}

// Header of a gasp_rng_state device allocation (its first 64 bytes).
struct RngHeader {
    uint32_t magic;
    uint32_t kind;  // KIND_PHILOX or KIND_MT
    uint64_t seed;
    int32_t ngrids;
    int32_t reserved0;
    uint64_t attempts;  // attempts consumed per replica
    uint64_t reserved[4];
};
static_assert(sizeof(RngHeader) == 64, "header size");
constexpr uint32_t kRngMagic = 0x474E5247u;  // "GRNG"
enum { KIND_PHILOX = 0, KIND_MT = 1 };

// A header the sweep kernels may take their Philox key and attempt counter
// from: made by gpuInitRand, counter-based, with streams for R replicas.
__device__ __forceinline__ bool rng_header_ok(const RngHeader *h, int R) {
    return h->magic == kRngMagic && h->kind == (uint32_t)KIND_PHILOX && h->ngrids >= R;
}

struct SweepArgs {
    void *lattice;           // native layout (see engine)
    int L, N, R;
    int lg;                  // log2(L) when L is a power of two
    int rpw;                 // replicas per workgroup (layouts 0/1)
    int stride;              // layout 0: bytes per replica; 1: dwords per replica; 2: dwords per plane (2 per group)
    uint64_t a_begin;        // per-replica attempt index of the first attempt
    const RngHeader *rng_dev;  // if set, a_begin and seed are read from this device header
                               // (graph replays, stream-ordered sweeps of caller grids)
    uint64_t n_attempts;     // attempts per replica in this launch (< 2^31)
    uint64_t seed;           // Philox key
    const uint32_t *draws;   // RNG_DRAWS: [R][3*n_attempts], draw stream from a_begin
    const uint64_t *thr;     // 10 thresholds, j = 5*s01 + (number of up neighbours)
    float *mag;              // out: magnetisation per replica
    int *fate;               // CHECK_FATES: fates per replica
    int *counts;             // out: [0] #(m>up), [1] #fate 0, [2] #fate 1
    int check;
    int freeze;              // skip replicas whose fate is already decided
    unsigned tbl_mask;       // sweep_hbm_waves: entries of its LDS conflict table - 1 (a power of two - 1)
    double dn_thr, up_thr;
};

__device__ __forceinline__ void philox4x32_10(uint32_t &c0, uint32_t &c1, uint32_t &c2,
                                              uint32_t &c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        // three-input XOR in one v_bitop3_b32 (truth table 0x96 = parity)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
}

// Wave mask of the lanes where p holds.  HIP's __ballot(int) widens p to an
// integer in a vector register and compares it again; the builtin takes the
// compare's lane mask as it is.
__device__ __forceinline__ uint64_t wave_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Exact site maps and periodic stencils.
//  Philox: site = floor(w*N / 2^32), computed as row = floor(w*L / 2^32) and
//  col = floor(frac*L) from the fractional part, which is the same integer;
//  for L = 2^lg this is w >> (32 - 2 lg).
//  MT19937: row = floor(L * (u+0.5)*2^-32) = ((2u+1)*L) >> 33 (src/mc_cpu.c:51),
//  which for L = 2^lg is u >> (32 - lg).
// Stencil positions are site indices times the geometry's unit U: U = 1 for
// site indices, U = 4 for byte offsets of a dword-per-site LDS lattice (the
// split multispin layout), U = 16 for the 16-byte site entries of the
// bit-sliced kernel; the offsets then serve as LDS addresses directly.
// row and col are the site's row and column mod 64, all the conflict filter
// uses.
struct Stencil {
    int c, up, dn, lt, rt;
    int row, col;
};
__device__ __forceinline__ bool in_stencil(int x, const Stencil &s) {
    return (x == s.c) | (x == s.up) | (x == s.dn) | (x == s.lt) | (x == s.rt);
}
// Wave mask of the lanes whose stencil holds the (wave-uniform) position x:
// one compare-to-mask per stencil site, ORed as scalars.
__device__ __forceinline__ uint64_t ballot_in_stencil(int x, const Stencil &s) {
    return wave_mask(x == s.c) | wave_mask(x == s.up) | wave_mask(x == s.dn) | wave_mask(x == s.lt) |
           wave_mask(x == s.rt);
}

template <bool POW2, int U = 1>
struct Geometry {
    static_assert(U == 1 || U == 2 || U == 4 || U == 16,
                  "positions in sites or in byte offsets of 2-, 4- or 16-byte entries");
    static constexpr int kLgU = U == 16 ? 4 : (U == 4 ? 2 : (U == 2 ? 1 : 0));
    int L, N, lg;  // lg = log2(L) when POW2

    __device__ __forceinline__ Stencil from_rowcol(int row, int col) const {
        Stencil s;
        s.row = row & 63;
        s.col = col & 63;
        s.c = (row * L + col) * U;
        s.up = (row == 0) ? s.c + (N - L) * U : s.c - L * U;
        s.dn = (row == L - 1) ? s.c - (N - L) * U : s.c + L * U;
        s.lt = (col == 0) ? s.c + (L - 1) * U : s.c - U;
        s.rt = (col == L - 1) ? s.c - (L - 1) * U : s.c + U;
        return s;
    }
    // cu = site * U for L = 2^lg: wrap by masks
    __device__ __forceinline__ Stencil from_pos_pow2(int cu) const {
        const int m = N * U - 1, cm = (L - 1) * U;
        Stencil s;
        s.row = (int)__builtin_amdgcn_ubfe((uint32_t)cu, (uint32_t)(lg + kLgU), 6u);
        s.col = (int)__builtin_amdgcn_ubfe((uint32_t)cu, (uint32_t)kLgU, (uint32_t)(lg < 6 ? lg : 6));
        s.c = cu;
        s.up = (cu - L * U) & m;
        s.dn = (cu + L * U) & m;
        s.lt = (cu & ~cm) | ((cu - U) & cm);
        s.rt = (cu & ~cm) | ((cu + U) & cm);
        return s;
    }
    __device__ __forceinline__ Stencil philox(uint32_t w) const {
        if constexpr (POW2) {
            // site = w >> (32 - 2 lg); site * 4 keeps two more bits, cleared
            if constexpr (U == 1) return from_pos_pow2((int)(w >> (32 - 2 * lg)));
            else return from_pos_pow2((int)((w >> (32 - 2 * lg - kLgU)) & ~(uint32_t)(U - 1)));
        } else {
            const uint64_t t = (uint64_t)w * (uint32_t)L;
            const int row = (int)(t >> 32);
            const int col = (int)(((uint64_t)(uint32_t)t * (uint32_t)L) >> 32);
            return from_rowcol(row, col);
        }
    }
    __device__ __forceinline__ Stencil mt(uint32_t urow, uint32_t ucol) const {
        if constexpr (POW2) {
            return from_pos_pow2((int)((((urow >> (32 - lg)) << lg) | (ucol >> (32 - lg))) << kLgU));
        } else {
            const int row = (int)((((uint64_t)urow << 1) + 1ull) * (uint32_t)L >> 33);
            const int col = (int)((((uint64_t)ucol << 1) + 1ull) * (uint32_t)L >> 33);
            return from_rowcol(row, col);
        }
    }
};

// ------------------------------------------------------------ lattice access
// Layout 2 (multispin): uint64 word per site, bit b = replica (64*group + b).
//   Lane reads the byte holding its bit; flips with an LDS atomic XOR on the
//   dword holding it (other waves flip other bits of the same word).
// Layout 3 (multispin, split word): the same HBM words, but a workgroup holds
//   one 32-bit half of each (32 replicas): a dword per site in LDS, half the
//   LDS of layout 2, so twice the workgroups for the same replicas.  Positions
//   are the dwords' byte offsets (geometry unit 4); a lane reads the whole
//   dword at LDS address kLdsLattice + position and takes bit rl.
// Layouts 4 and 5 (multispin, quarter and eighth word): 16 or 8 replicas per
//   workgroup, a 16- or 8-bit entry per site in LDS (positions are byte
//   offsets: units 2 and 1), for few replicas spread over more CUs.  A flip
//   XORs bit 8 (position mod 4) + rl of the dword that holds the entry.
// Layout 1 (bit lattice): replica-private region of `stride` dwords in LDS,
//   bit (site & 31) of dword (site >> 5).
// Layout 0 (HBM): one byte (0/1) per site in global memory, replica stride
//   `stride` bytes.  A workgroup's replicas are one buffer resource (scalar
//   registers) and a lane addresses its replica's bytes by a 32-bit offset:
//   no 64-bit address arithmetic per read, and reads past the workgroup's
//   replicas return 0.  The sweeps order a round's flips before the next
//   round's reads with fences (one wave) or barriers (several).
// Layout 6 (layout 0 staged in LDS): one replica's bytes at LDS address
//   kLdsLattice + site, for the multi-wave sweep of one replica per workgroup;
//   its barriers order the flips (plain byte stores) before the next reads.
// Layout 7 (layout 0 staged in LDS as bits): one replica's spins as bit
//   (site & 31) of the dword at LDS address kLdsLattice + 4 (site >> 5), an
//   eighth of layout 6's LDS; a flip is an atomic XOR (a round's flips are at
//   distinct sites, but several may share a dword).
template <int LAT>
struct Lattice {
    unsigned char *base;  // LDS (1/2) or global (0) base of this lane's replica
    unsigned *dw;         // layout 2: LDS lattice as dwords
    int rl;               // layout 2: replica bit within the 64-bit word
    __amdgpu_buffer_rsrc_t rsrc;  // layout 0: the workgroup's replicas
    unsigned off;                 // layout 0: byte offset of this lane's replica in rsrc

    __device__ __forceinline__ int spin(int site) const {
        if constexpr (LAT == LAT_MULTISPIN) {
            return (base[site * 8] >> (rl & 7)) & 1;
        } else if constexpr (LAT == LAT_MULTIHALF) {
            return (int)__builtin_amdgcn_ubfe(lds_load32(kLdsLattice + (unsigned)site), (uint32_t)rl, 1u);
        } else if constexpr (LAT == LAT_MULTI16) {
            return (int)__builtin_amdgcn_ubfe(*(const lds_u16 *)(uintptr_t)(kLdsLattice + (unsigned)site),
                                              (uint32_t)rl, 1u);
        } else if constexpr (LAT == LAT_MULTI8) {
            return (int)__builtin_amdgcn_ubfe(*(const lds_u8 *)(uintptr_t)(kLdsLattice + (unsigned)site),
                                              (uint32_t)rl, 1u);
        } else if constexpr (LAT == LAT_BITS) {
            return (base[site >> 3] >> (site & 7)) & 1;
        } else if constexpr (LAT == LAT_LDS8) {
            return (int)*(const lds_u8 *)(uintptr_t)(kLdsLattice + (unsigned)site);
        } else if constexpr (LAT == LAT_LDS1) {
            return (int)__builtin_amdgcn_ubfe(lds_load32(kLdsLattice + (((unsigned)site >> 5) << 2)),
                                              (uint32_t)site & 31u, 1u);
        } else {
            return (int)__builtin_amdgcn_raw_buffer_load_b8(rsrc, off + (unsigned)site, 0, 0);
        }
    }
    __device__ __forceinline__ void flip(int site, int old) const {
        if constexpr (LAT == LAT_MULTISPIN) {
            // word `site` is dwords 2*site (bits 0..31) and 2*site+1 (bits 32..63)
            atomicXor(dw + 2 * site + (rl >> 5), 1u << (rl & 31));
        } else if constexpr (LAT == LAT_MULTIHALF) {
            lds_xor32(kLdsLattice + (unsigned)site, 1u << rl);
        } else if constexpr (LAT == LAT_MULTI16 || LAT == LAT_MULTI8) {
            lds_xor32(kLdsLattice + ((unsigned)site & ~3u), 1u << ((((unsigned)site & 3u) << 3) + (unsigned)rl));
        } else if constexpr (LAT == LAT_BITS) {
            atomicXor((unsigned *)base + (site >> 5), 1u << (site & 31));
        } else if constexpr (LAT == LAT_LDS8) {
            *(lds_u8 *)(uintptr_t)(kLdsLattice + (unsigned)site) = (unsigned char)(old ^ 1);
        } else if constexpr (LAT == LAT_LDS1) {
            lds_xor32(kLdsLattice + (((unsigned)site >> 5) << 2), 1u << ((unsigned)site & 31u));
        } else {
            __builtin_amdgcn_raw_buffer_store_b8((unsigned char)(old ^ 1), rsrc, off + (unsigned)site, 0, 0);
        }
    }
};

struct Attempt {
    Stencil st;
    int s01;
    bool acc;
};

// The five spins of a stencil, read before any is used: a kernel reads both
// attempts' spins first, so that on the HBM layout (atomic loads, which the
// compiler keeps in order with the threshold read) the ten reads share one
// round trip to L2 instead of two.
struct Spins {
    int c, up, dn, lt, rt;
};
template <int LAT>
__device__ __forceinline__ Spins read_spins(const Lattice<LAT> &lat, const Stencil &st) {
    if constexpr (LAT == LAT_HBM) {
        // The centre row's bytes lo .. hi (c - 1 .. c + 1 where the row does
        // not wrap) come from the aligned dword holding lo and, for the lanes
        // whose hi lies in the next one, that dword too (hi < N, so it is in
        // the lattice): fewer scattered loads per attempt, which is what the
        // byte lattice's rounds wait on.  A wrapped neighbour is read alone.
        const int c = st.c;
        const bool wl = st.lt != c - 1, wr = st.rt != c + 1;  // column 0 / L - 1
        const int lo = wl ? c : c - 1, hi = wr ? c : c + 1;
        const int d0 = lo & ~3;
        const unsigned v0 = __builtin_amdgcn_raw_buffer_load_b32(lat.rsrc, lat.off + (unsigned)d0, 0, 0);
        unsigned v1 = 0u;
        if ((hi & ~3) != d0) v1 = __builtin_amdgcn_raw_buffer_load_b32(lat.rsrc, lat.off + (unsigned)d0 + 4u, 0, 0);
        const int up = lat.spin(st.up), dn = lat.spin(st.dn);
        const int lt_far = wl ? lat.spin(st.lt) : 0, rt_far = wr ? lat.spin(st.rt) : 0;
        // the window lo .. hi lies within the 8 bytes from d0: one 64-bit shift per byte
        const uint64_t v = ((uint64_t)v1 << 32) | v0;
        auto byte_at = [&](int x) { return (int)((v >> (8 * (x - d0))) & 1ull); };
        return Spins{byte_at(c), up, dn, wl ? lt_far : byte_at(c - 1), wr ? rt_far : byte_at(c + 1)};
    } else {
        return Spins{lat.spin(st.c), lat.spin(st.up), lat.spin(st.dn), lat.spin(st.lt), lat.spin(st.rt)};
    }
}
// acc: the draw accepts the flip (whether the attempt is enabled is the caller's)
__device__ __forceinline__ Attempt decide(const Stencil &st, const Spins &x, uint32_t ua) {
    Attempt t;
    t.st = st;
    t.s01 = x.c;
    const unsigned n3 = (unsigned)(x.up + x.dn + x.lt);
    // byte offset of table entry 8 s01 + nup: shifts and adds, no multiply or select
    const unsigned j3 = ((unsigned)x.c << 3) + n3;
    const unsigned off = (j3 + (unsigned)x.rt) << 3;
    const uint64_t T = lds_load64(kLdsThr + off);
    t.acc = (uint64_t)ua < T;
    return t;
}

// Host threshold j = 5 s01 + nup (j < 10) goes to LDS table entry 8 s01 + nup.
__device__ __forceinline__ void stage_thresholds(const SweepArgs &a, uint64_t *thr_lds, int tid) {
    if (tid < 10) thr_lds[tid < 5 ? tid : tid + 3] = a.thr[tid];
}

// ------------------------------------------------------------ conflict filter
// An attempt at (r', c') conflicts with an accepted attempt at (r, c) iff
// (dr, dc) = (r' - r, c' - c) mod L is one of the five stencil offsets, so
// only if its row and its column each differ by at most 1 (periodically).
// The filter tests rows and columns as 64-bit masks over the values mod 64.
// nbr_mask(k): the union, over every value t < L with t mod 64 = k, of the
// values within periodic distance 1 of t, folded mod 64.  Exact for L <= 64;
// for larger L the fold only adds bits (for powers of two each t gives the
// same bits), so the test stays conservative: a true conflict always passes.
__device__ __forceinline__ uint64_t nbr_mask(int k, int L) {
    uint64_t m = 0ull;
    for (int t = k; t < L; t += 64) {
        const int rm1 = (t == 0) ? L - 1 : t - 1;
        const int rp1 = (t == L - 1) ? 0 : t + 1;
        m |= (1ull << (rm1 & 63)) | (1ull << (t & 63)) | (1ull << (rp1 & 63));
    }
    return m;
}

// DPP row_shr:K (lane l reads lane l - K of its 16-lane row; 0 outside it).
template <int K>
__device__ __forceinline__ uint32_t row_shr(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + K, 0xf, 0xf, true);
}

// OR of w[i] over the lanes of the group (S consecutive, aligned lanes)
// that precede this lane, for each of the NW dwords: a shift by one lane, then
// an inclusive Hillis-Steele OR-scan (steps k < S) inside the 16-lane DPP row.
template <int S, int K, int NW>
__device__ __forceinline__ void scan_step(uint32_t (&w)[NW], int slot) {
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        uint32_t t = row_shr<K>(w[i]);
        if constexpr (S < 16) {
            if (slot < K) t = 0u;  // the source lane belongs to the previous group
        }
        w[i] |= t;
    }
}

// Groups of 32 or 64 lanes span DPP rows: after the scan inside each row,
// row_bcast:15 ORs the total of row 0 (2) into every lane of row 1 (3), and
// for 64-lane groups row_bcast:31 then ORs the total of rows 0-1 into rows 2
// and 3.  `x` holds each lane's own values (the row totals are inclusive).
template <int S, int NW>
__device__ __forceinline__ void group_excl_or(uint32_t (&w)[NW], int slot) {
    static_assert(S >= 2 && S <= 64, "2 to 64 lanes per group");
    constexpr int SR = S < 16 ? S : 16;  // lanes of the group inside one row
    uint32_t x[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        x[i] = w[i];
        w[i] = row_shr<1>(w[i]);
        if constexpr (S < 16) {
            if (slot == 0) w[i] = 0u;
        }
    }
    scan_step<SR, 1>(w, slot);
    if constexpr (SR > 2) scan_step<SR, 2>(w, slot);
    if constexpr (SR > 4) scan_step<SR, 4>(w, slot);
    if constexpr (SR > 8) scan_step<SR, 8>(w, slot);
    if constexpr (S >= 32) {
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const uint32_t tot = w[i] | x[i];  // inclusive within the row
            // rows 1 and 3 receive lane 15 / lane 47; rows 0 and 2 keep 0
            const uint32_t p = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)tot, 0x142, 0xa, 0xf, false);
            uint32_t e = w[i] | p;
            if constexpr (S == 64) {
                // rows 2 and 3 receive lane 31, the total of rows 0 and 1
                const uint32_t q = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(tot | p), 0x143, 0xc, 0xf, false);
                e |= q;
            }
            w[i] = e;
        }
    }
}

// The filter tests rows and columns only: once the exact loop visits only the
// groups that hold a candidate, the diagonal projections (two more masks per
// attempt and two more scanned dwords) cost more than the extra candidates
// they remove.  The test admits the stencil offsets and the four diagonal
// neighbours of each earlier accepted attempt.

// Bits strictly above the lowest set bit of x inside each aligned S-bit group
// (none for a group without a set bit): per group, ~(x ^ (x - 1)).  The even
// and the odd groups are done as one 64-bit subtraction each; the lowest bit
// of each group of the other parity is set as a guard, so that the borrow out
// of an empty group stops there (out of the top group it leaves the word).
template <int S>
__device__ __forceinline__ uint64_t group_after(uint64_t x) {
    if constexpr (S == 64) {
        return ~(x ^ (x - 1ull));  // x = 0: ~(all ones) = 0
    } else {
        constexpr uint64_t kOnes = (S == 32) ? 0xffffffffull : ((1ull << S) - 1ull);
        uint64_t lo = 0ull, ev = 0ull;
#pragma unroll
        for (int p = 0; p < 64; p += S) lo |= 1ull << p;
#pragma unroll
        for (int p = 0; p < 64; p += 2 * S) ev |= kOnes << p;
        const uint64_t od = ~ev, lo_ev = lo & ev, lo_od = lo & od;
        const uint64_t e = (x & ev) | lo_od;
        const uint64_t o = (x & od) | lo_ev;
        return (ev & ~(e ^ (e - lo_ev))) | (od & ~(o ^ (o - lo_od)));
    }
}

// Lanes of the S-lane groups (aligned, S < 64) that hold at least one set bit
// of the wave mask x: OR each group into its first lane, keep those bits,
// and spread each back over its group ((b << S) - b sets bits p .. p+S-1 for
// a start bit p; the top group wraps to -2^p, also bits p .. 63).
template <int S>
__device__ __forceinline__ uint64_t group_span(uint64_t x) {
    if constexpr (S == 64) return x ? ~0ull : 0ull;  // one group: the whole wave
#pragma unroll
    for (int k = S / 2; k >= 1; k >>= 1) x |= x >> k;
    uint64_t starts = 0ull;
#pragma unroll
    for (int p = 0; p < 64; p += S) starts |= 1ull << p;
    const uint64_t b = x & starts;
    return (b << S) - b;
}

// Per-replica epilogue: magnetisation and the threshold check.  Must be
// called by a whole wave (ballots); `own` marks the lane reporting replica r.
__device__ __forceinline__ void finish_replica(const SweepArgs &a, bool own, int r,
                                               unsigned nup_count) {
    float m = 0.0f;
    bool gt_up = false, is_a = false, is_b = false;
    if (own) {
        const int sum = 2 * (int)nup_count - a.N;
        // src/mc_gpu.cu:445-449: float sum of +-1 spins (exact) over float N
        m = (float)sum / (float)a.N;
        a.mag[r] = m;
        gt_up = (double)m > a.up_thr;
        if (a.check == CHECK_FATES) {
            int f = a.fate[r];
            if (f == -1) {
                if ((double)m > a.up_thr) f = 1;
                else if ((double)m < a.dn_thr) f = 0;
                if (f != -1) a.fate[r] = f;
            }
            is_a = (f == 0);
            is_b = (f == 1);
        }
    }
    if (a.check == CHECK_COUNT) {
        const uint64_t b = wave_mask(gt_up);
        if ((threadIdx.x & 63) == 0 && b) atomicAdd(&a.counts[0], (int)__popcll(b));
    } else if (a.check == CHECK_FATES) {
        const uint64_t ba = wave_mask(is_a), bb = wave_mask(is_b);
        if ((threadIdx.x & 63) == 0) {
            if (ba) atomicAdd(&a.counts[1], (int)__popcll(ba));
            if (bb) atomicAdd(&a.counts[2], (int)__popcll(bb));
        }
    }
}

// ------------------------------------------------------------ mark plane
// An exact candidate test for the speculation groups of the split-word layouts
// (an entry of 32, 16 or 8 bits per site, bit rl = replica): a second plane of
// the same shape, MOFF bytes above the lattice, all zero between rounds.  Each
// accepted attempt ORs its replica's bit into its site's mark (returning the
// old mark), then every attempt after an accepted attempt of its group reads
// the marks of its five stencil sites (the earlier ones cannot follow an
// accepted attempt), at the lattice addresses plus MOFF.  An attempt whose
// stencil holds the site of an accepted attempt of its replica from the same
// round sees that bit: at a neighbour through the neighbour's mark; at its own
// site through the site's mark if it did not accept, and if it did through the
// old mark returned to one of the two ORs (a wave's lanes and its successive
// instructions reach LDS in order; an accepted attempt that skips the reads
// still has that old mark).  Those attempts are the candidates; the exact loop
// sorts out which accepted attempt came first.  Unlike the row and column
// masks, whose unions admit a growing share of attempts as a group's accepted
// attempts multiply, it admits none without an accepted attempt in the
// stencil.
// The accepted attempts clear their bits at the end of the round.
__device__ __forceinline__ unsigned lds_or_rtn32(unsigned addr, unsigned v) {
    return __hip_atomic_fetch_or(lds_ptr32(addr), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_and32(unsigned addr, unsigned v) {
    __hip_atomic_fetch_and(lds_ptr32(addr), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// The entry at byte position pos of a split-word plane: half words fill a
// dword; quarter and eighth words share one, at bits 8 (pos mod 4) + rl.
template <int LAT>
__device__ __forceinline__ unsigned entry_shift(unsigned pos, int rl) {
    if constexpr (LAT == LAT_MULTIHALF) return (unsigned)rl;
    else return ((pos & 3u) << 3) + (unsigned)rl;
}
template <int LAT>
__device__ __forceinline__ unsigned entry_dword(unsigned pos) {
    if constexpr (LAT == LAT_MULTIHALF) return pos;
    else return pos & ~3u;
}
template <int LAT>
__device__ __forceinline__ unsigned read_entry(unsigned addr) {
    if constexpr (LAT == LAT_MULTIHALF) return lds_load32(addr);
    else if constexpr (LAT == LAT_MULTI16) return *(const lds_u16 *)(uintptr_t)addr;
    else return *(const lds_u8 *)(uintptr_t)addr;
}
// Marks the site; returns whether its old mark held the replica's bit.
template <int LAT, int MOFF>
__device__ __forceinline__ bool mark_site(int pos, int rl) {
    constexpr unsigned M = (unsigned)(kLdsLattice + MOFF);
    const unsigned sh = entry_shift<LAT>((unsigned)pos, rl);
    return ((lds_or_rtn32(M + entry_dword<LAT>((unsigned)pos), 1u << sh) >> sh) & 1u) != 0u;
}
template <int LAT, int MOFF>
__device__ __forceinline__ void unmark_site(int pos, int rl) {
    constexpr unsigned M = (unsigned)(kLdsLattice + MOFF);
    const unsigned sh = entry_shift<LAT>((unsigned)pos, rl);
    lds_and32(M + entry_dword<LAT>((unsigned)pos), ~(1u << sh));
}
// Half words: the OR of the stencil's marks, `own` (the old mark of an
// accepted attempt) in place of the centre's.
template <int MOFF>
__device__ __forceinline__ unsigned half_marks(const Stencil &s, bool acc, unsigned own) {
    constexpr unsigned M = (unsigned)(kLdsLattice + MOFF);
    const unsigned c = lds_load32(M + (unsigned)s.c);
    const unsigned n = lds_load32(M + (unsigned)s.up) | lds_load32(M + (unsigned)s.dn) |
                       lds_load32(M + (unsigned)s.lt) | lds_load32(M + (unsigned)s.rt);
    return n | (acc ? own : c);
}
// Quarter and eighth words: whether the stencil's marks hold the replica's
// bit; `own` stands for the centre's mark of an accepted attempt (its old
// mark).
template <int LAT, int MOFF>
__device__ __forceinline__ bool stencil_hit(const Stencil &s, int rl, bool acc, bool own) {
    constexpr unsigned M = (unsigned)(kLdsLattice + MOFF);
    const unsigned c = read_entry<LAT>(M + (unsigned)s.c);
    const unsigned n = read_entry<LAT>(M + (unsigned)s.up) | read_entry<LAT>(M + (unsigned)s.dn) |
                       read_entry<LAT>(M + (unsigned)s.lt) | read_entry<LAT>(M + (unsigned)s.rt);
    return ((n >> rl) & 1u) != 0u || (acc ? own : ((c >> rl) & 1u) != 0u);
}

// MOFF: 0 for the row/column conflict filter, else the byte offset of the
// mark plane above the lattice (split words).
template <int LAT, int RNG, int S, bool POW2, int MOFF = 0>
__global__ void __launch_bounds__(1024) sweep_kernel(SweepArgs a) {
    static_assert(S >= 1 && S <= 64 && (S & (S - 1)) == 0, "S must be a power of two <= 64");
    static_assert(MOFF == 0 || ((LAT == LAT_MULTIHALF || LAT == LAT_MULTI16 || LAT == LAT_MULTI8) && S >= 2 && S <= 32),
                  "mark plane: split words, 2 to 32 slots");
    constexpr bool SPLIT = (LAT == LAT_MULTIHALF) || (LAT == LAT_MULTI16) || (LAT == LAT_MULTI8);
    constexpr bool MULTI = (LAT == LAT_MULTISPIN) || SPLIT;
    // replicas per workgroup in the multispin layouts (whole, half, quarter, eighth words)
    constexpr int RPG = (LAT == LAT_MULTIHALF) ? 32 : (LAT == LAT_MULTI16) ? 16 : (LAT == LAT_MULTI8) ? 8 : 64;
    static_assert(!MULTI || RPG * S <= 1024, "workgroup too large");
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *thr_lds = (uint64_t *)(smem + kLdsThr);
    unsigned *cnt_lds = (unsigned *)(smem + kLdsCnt);
    uint64_t *nbr_lds = (uint64_t *)(smem + kLdsNbr);
    unsigned char *lat_lds = smem + kLdsLattice;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int L = a.L, N = a.N;
    const int slot = lane & (S - 1);
    const int gbase = lane & ~(S - 1);
    // the lanes of this lane's group below it (mark-plane filter: does an
    // earlier lane of the group hold an accepted attempt)
    const uint64_t below = ((1ull << lane) - 1ull) & ~((1ull << gbase) - 1ull);

    int rl, r0;
    if constexpr (MULTI) {
        rl = wave * (64 / S) + (lane / S);
        r0 = (int)blockIdx.x * RPG;  // split: the groups' 64 replicas over 64 / RPG workgroups
    } else {
        rl = tid / S;
        r0 = blockIdx.x * a.rpw;
    }
    const int r = r0 + rl;
    // live: the replica exists; active: it makes attempts in this launch
    // (replicas past R, and when freezing those with a decided fate, do not)
    const bool live = r < a.R;
    const bool active = live && !(a.freeze && a.fate[r] != -1);
    // absolute LDS addresses assume smem at address 0, and a device header
    // must be a Philox one for these replicas: otherwise leave the lattice
    // alone and report NaN magnetisations (uniform over the workgroup)
    if (lds_origin(smem) != 0u || (a.rng_dev && !rng_header_ok(a.rng_dev, a.R))) {
        if (live && slot == 0) a.mag[r] = __builtin_nanf("");
        return;
    }

    // ---------------------------------------------------------- prologue
    stage_thresholds(a, thr_lds, tid);
    if (tid < 64) cnt_lds[tid] = 0u;
    // entry 64 is zero: the mask of an attempt that was not accepted
    for (int t = tid; t <= 64; t += blockDim.x) nbr_lds[t] = t < 64 ? nbr_mask(t, a.L) : 0ull;
    if constexpr (LAT == LAT_MULTISPIN) {
        // HBM: the group's two dword planes (bits 0..31, then 32..63 of each
        // site's word), a.stride dwords each (N rounded up to 4); LDS: one
        // 64-bit word per site (dwords 2 site and 2 site + 1)
        const uint4 *lo = (const uint4 *)((const unsigned *)a.lattice + (size_t)(2 * blockIdx.x) * a.stride);
        const uint4 *hi = lo + a.stride / 4;
        uint4 *dst = (uint4 *)lat_lds;
        for (int i = tid; i < a.stride / 4; i += blockDim.x) {
            const uint4 l = lo[i], h = hi[i];
            dst[2 * i] = make_uint4(l.x, h.x, l.y, h.y);
            dst[2 * i + 1] = make_uint4(l.z, h.z, l.w, h.w);
        }
    } else if constexpr (LAT == LAT_MULTIHALF) {
        // workgroup 2g + h stages plane h of group g, which is plane blockIdx.x
        const uint4 *src = (const uint4 *)((const unsigned *)a.lattice + (size_t)blockIdx.x * a.stride);
        uint4 *dst = (uint4 *)lat_lds;
        for (int i = tid; i < a.stride / 4; i += blockDim.x) dst[i] = src[i];
        if constexpr (MOFF > 0) {
            uint4 *mk = (uint4 *)(lat_lds + MOFF);
            for (int i = tid; i < a.stride / 4; i += blockDim.x) mk[i] = make_uint4(0u, 0u, 0u, 0u);
        }
    } else if constexpr (LAT == LAT_MULTI16 || LAT == LAT_MULTI8) {
        // bits o .. o + RPG - 1 of plane (64 g + o) / 32 of each site's dword,
        // where replica r0 = blockIdx.x RPG is bit o of that plane
        const unsigned *src = (const unsigned *)a.lattice + (size_t)(r0 >> 5) * a.stride;
        const int o = r0 & 31;
        for (int i = tid; i < N; i += blockDim.x) {
            const unsigned v = src[i] >> o;
            if constexpr (LAT == LAT_MULTI16) ((unsigned short *)lat_lds)[i] = (unsigned short)v;
            else lat_lds[i] = (unsigned char)v;
        }
        if constexpr (MOFF > 0) {
            // the plane's 2 or 1 bytes per site, a.stride (a multiple of 4) sites
            unsigned *mk = (unsigned *)(lat_lds + MOFF);
            const int ndw = (LAT == LAT_MULTI16 ? 2 : 1) * a.stride / 4;
            for (int i = tid; i < ndw; i += blockDim.x) mk[i] = 0u;
        }
    } else if constexpr (LAT == LAT_BITS) {
        const int nrep = min(a.rpw, a.R - r0);
        const int ndw = nrep * a.stride;
        const unsigned *src = (const unsigned *)a.lattice + (size_t)r0 * a.stride;
        unsigned *dst = (unsigned *)lat_lds;
        for (int i = tid; i < ndw; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();

    Lattice<LAT> lat;
    lat.dw = (unsigned *)lat_lds;
    lat.rl = rl;
    if constexpr (MULTI) {
        lat.base = lat_lds + (rl >> 3);
    } else if constexpr (LAT == LAT_BITS) {
        lat.base = lat_lds + (size_t)rl * a.stride * 4;
    } else {
        lat.base = (unsigned char *)a.lattice + (size_t)(live ? r : 0) * a.stride;
        // the workgroup's live replicas r0 .. r0 + nrep - 1 as one buffer
        const int nrep = max(0, min(a.rpw, a.R - r0));
        lat.rsrc = __builtin_amdgcn_make_buffer_rsrc((unsigned char *)a.lattice + (size_t)r0 * a.stride, (short)0,
                                                     nrep * a.stride, kBufferWord3);
        lat.off = (unsigned)(rl * a.stride);
    }

    // Launch-local attempt coordinates shifted by b = a_begin & 1 so that
    // attempt pairs (2q, 2q+1) coincide with global Philox pairs P0 + q.
    const uint64_t a_begin = a.rng_dev ? a.rng_dev->attempts : a.a_begin;
    const uint64_t seed = a.rng_dev ? a.rng_dev->seed : a.seed;
    const uint32_t b = (uint32_t)(a_begin & 1ull);
    const uint64_t P0 = a_begin >> 1;
    const uint32_t e = (uint32_t)a.n_attempts + b;
    uint32_t l0 = active ? b : e;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    // stencil positions: byte offsets in the split layouts (entries of 4, 2, 1 bytes)
    constexpr int U = (LAT == LAT_MULTIHALF) ? 4 : (LAT == LAT_MULTI16) ? 2 : 1;
    const Geometry<POW2, U> geo{L, N, a.lg};
    const uint32_t *drw = nullptr;
    if constexpr (RNG == RNG_DRAWS)
        drw = a.draws + (size_t)(live ? r : 0) * (3 * a.n_attempts);

    // ---------------------------------------------------------- rounds
    while (wave_mask(l0 < e) != 0ull) {
        const uint32_t q = (l0 >> 1) + (uint32_t)slot;
        const uint32_t lA = 2 * q, lB = lA + 1;
        const bool enA = (lA >= l0) && (lA < e);
        const bool enB = (lB >= l0) && (lB < e);

        Stencil sA, sB;
        uint32_t uaA, uaB;
        if constexpr (RNG == RNG_PHILOX) {
            const uint64_t pair = P0 + q;
            uint32_t c0 = (uint32_t)pair, c1 = (uint32_t)(pair >> 32), c2 = (uint32_t)r, c3 = kPhiloxDomain;
            philox4x32_10(c0, c1, c2, c3, k0, k1);
            sA = geo.philox(c0);
            uaA = c1;
            sB = geo.philox(c2);
            uaB = c3;
        } else {
            uint32_t u0 = 0, u1 = 0, u2 = 0, v0 = 0, v1 = 0, v2 = 0;
            if (enA) {
                const uint32_t *d = drw + 3 * (size_t)(lA - b);
                u0 = d[0]; u1 = d[1]; u2 = d[2];
            }
            if (enB) {
                const uint32_t *d = drw + 3 * (size_t)(lB - b);
                v0 = d[0]; v1 = d[1]; v2 = d[2];
            }
            sA = geo.mt(u0, u1);
            uaA = u2;
            sB = geo.mt(v0, v1);
            uaB = v2;
        }

        const Spins xA = read_spins<LAT>(lat, sA), xB = read_spins<LAT>(lat, sB);
        const Attempt A = decide(sA, xA, uaA);
        const Attempt B = decide(sB, xB, uaB);

        // Accepted attempts as wave masks, used from here on as lane
        // predicates through inverse ballots (kept in scalar registers, where
        // per-lane booleans across the branches below cost selects).
        // (each ballot of a single compare: a ballot of a combined boolean
        // costs a select and a compare to clear the inactive lanes)
        const uint64_t eA = wave_mask(lA >= l0) & wave_mask(lA < e);
        const uint64_t eB = wave_mask(lB >= l0) & wave_mask(lB < e);
        uint64_t mA = eA & wave_mask(A.acc), mB = eB & wave_mask(B.acc);
        // lanes whose B holds the same lane's accepted A in its stencil
        uint64_t oB = mA & ballot_in_stencil(A.st.c, B.st);
        // opaque to the optimiser, which would otherwise fold the inverse
        // ballots below back into per-lane booleans
        asm volatile("" : "+s"(mA), "+s"(mB), "+s"(oB));
        if constexpr (S > 1) {
            // Wave masks: cA / cB = lanes whose attempt A / B has an earlier
            // accepted attempt of another lane of its group in its stencil.
            // The loop runs over the accepted attempts only (wave-uniform);
            // `later` (the lanes after i in i's group) is a scalar mask.
            uint64_t cA = 0ull, cB = oB;
            uint64_t m = mA | mB;
            // The filter pays up to 32 lanes per group (a 32-slot split launch
            // at L=64, 16384 replicas: 42.5 -> 37.9 ms); over a whole wave
            // (method 0 at 64 slots) it cost more than it saved (39.1 -> 40.6).
            if constexpr (MOFF > 0) {
                // the mark plane (see mark_site)
                if (m != 0ull) {
                    const bool accA = __builtin_amdgcn_inverse_ballot_w64(mA);
                    const bool accB = __builtin_amdgcn_inverse_ballot_w64(mB);
                    // Only the lanes after their group's first lane with an
                    // accepted attempt can follow an accepted attempt of
                    // another lane, so only they read their stencils' marks
                    // (at L = 64 with 16384 replicas on 16 slots about half
                    // the lanes skip the ten reads).
                    // An accepted attempt among the others still reports a
                    // mark at its own site through the old mark its OR
                    // returned: of two ORs at one site in one instruction only
                    // the second sees the first, and that may be either lane.
                    const bool reads = (m & below) != 0ull;
                    bool hA, hB;
                    if constexpr (LAT == LAT_MULTIHALF) {
                        // whole dwords: one test of the replica's bit at the end
                        constexpr unsigned M = (unsigned)(kLdsLattice + MOFF);
                        const unsigned bit = 1u << rl;
                        unsigned ownA = 0u, ownB = 0u;
                        if (accA) ownA = lds_or_rtn32(M + (unsigned)A.st.c, bit);
                        if (accB) ownB = lds_or_rtn32(M + (unsigned)B.st.c, bit);
                        asm volatile("" ::: "memory");
                        unsigned gA = ownA, gB = ownB;  // 0 unless accepted
                        if (reads) {
                            gA = half_marks<MOFF>(A.st, accA, ownA);
                            gB = half_marks<MOFF>(B.st, accB, ownB);
                        }
                        hA = (gA & bit) != 0u;
                        hB = (gB & bit) != 0u;
                    } else {
                        bool ownA = false, ownB = false;
                        if (accA) ownA = mark_site<LAT, MOFF>(A.st.c, rl);
                        if (accB) ownB = mark_site<LAT, MOFF>(B.st.c, rl);
                        asm volatile("" ::: "memory");
                        hA = ownA;
                        hB = ownB;
                        if (reads) {
                            hA = stencil_hit<LAT, MOFF>(A.st, rl, accA, ownA);
                            hB = stencil_hit<LAT, MOFF>(B.st, rl, accB, ownB);
                        }
                    }
                    const uint64_t cand = (eA & wave_mask(hA)) | (eB & wave_mask(hB));
                    m = cand ? (m & group_span<S>(cand)) : 0ull;
                }
            } else if constexpr (S <= 32) {
                // Filter (see nbr_mask): union the row and column masks of
                // the accepted attempts of the earlier lanes of the group; an
                // attempt within 1 of that union on both is a candidate.
                // Without a candidate in the wave the exact loop is skipped.
                if (m != 0ull) {
                    // not accepted: the zero entry (a select, not a masked read)
                    const bool accA = __builtin_amdgcn_inverse_ballot_w64(mA);
                    const bool accB = __builtin_amdgcn_inverse_ballot_w64(mB);
                    const uint64_t rm = lds_load64(kLdsNbr + 8u * (unsigned)(accA ? A.st.row : 64)) |
                                        lds_load64(kLdsNbr + 8u * (unsigned)(accB ? B.st.row : 64));
                    const uint64_t km = lds_load64(kLdsNbr + 8u * (unsigned)(accA ? A.st.col : 64)) |
                                        lds_load64(kLdsNbr + 8u * (unsigned)(accB ? B.st.col : 64));
                    uint32_t w[4] = {(uint32_t)rm, (uint32_t)(rm >> 32), (uint32_t)km, (uint32_t)(km >> 32)};
                    group_excl_or<S>(w, slot);
                    const uint64_t er = ((uint64_t)w[1] << 32) | w[0], ek = ((uint64_t)w[3] << 32) | w[2];
                    const bool sA = ((er >> A.st.row) & (ek >> A.st.col) & 1ull) != 0ull;
                    const bool sB = ((er >> B.st.row) & (ek >> B.st.col) & 1ull) != 0ull;
                    // the exact loop then visits only the accepted attempts
                    // of groups that hold a candidate (a group without one
                    // has no conflict)
                    const uint64_t cand = wave_mask(sA || sB);
                    m = cand ? (m & group_span<S>(cand)) : 0ull;
                }
            }
            while (m) {
                const int i = __builtin_ctzll(m);
                m &= m - 1;
                const uint64_t grp = (S == 64) ? ~0ull : (((1ull << S) - 1ull) << (i & ~(S - 1)));
                const uint64_t later = grp & ~((2ull << i) - 1ull);  // i = 63: 2 << 63 = 0, later = 0
                if ((mA >> i) & 1ull) {
                    const int x = __builtin_amdgcn_readlane(A.st.c, i);
                    cA |= ballot_in_stencil(x, A.st) & later;
                    cB |= ballot_in_stencil(x, B.st) & later;
                }
                if ((mB >> i) & 1ull) {
                    const int x = __builtin_amdgcn_readlane(B.st.c, i);
                    cA |= ballot_in_stencil(x, A.st) & later;
                    cB |= ballot_in_stencil(x, B.st) & later;
                }
            }
            // Invalid attempts, as wave masks: every lane after the first
            // conflicting lane of its group (both attempts), and at a lane
            // its own conflicts (an invalid A also invalidates its B).
            // (a round without a conflict, the common case, skips the masks
            // and the prefix lengths: nothing is invalid, all 2 S are valid)
            const uint64_t cc = cA | cB;
            uint64_t invA = 0ull, invB = 0ull;
            uint32_t P = 2u * (uint32_t)S;
            if (cc != 0ull) {
                invA = group_after<S>(cc) | cA;
                invB = invA | cB;
                // The valid attempts of a group are a prefix of its order A,
                // B, A, B, ... (an attempt is invalid once an earlier one
                // conflicts); the next round starts after that prefix.
                if constexpr (S == 64) {
                    P = (uint32_t)(__popcll(~invA) + __popcll(~invB));
                } else {
                    constexpr uint32_t kOnes = (S == 32) ? 0xffffffffu : ((1u << S) - 1u);
                    P = (uint32_t)__popc((uint32_t)(~invA >> gbase) & kOnes) +
                        (uint32_t)__popc((uint32_t)(~invB >> gbase) & kOnes);
                }
            }
            if (__builtin_amdgcn_inverse_ballot_w64(mA & ~invA)) lat.flip(A.st.c, A.s01);
            if (__builtin_amdgcn_inverse_ballot_w64(mB & ~invB)) lat.flip(B.st.c, B.s01);
            if constexpr (MOFF > 0) {
                // clear the round's marks (every accepted attempt set one)
                asm volatile("" ::: "memory");
                if (__builtin_amdgcn_inverse_ballot_w64(mA)) unmark_site<LAT, MOFF>(A.st.c, rl);
                if (__builtin_amdgcn_inverse_ballot_w64(mB)) unmark_site<LAT, MOFF>(B.st.c, rl);
            }
            // The next round starts after the valid prefix.  Disabled attempts
            // before l0 are valid (nothing precedes them) and those at or past
            // e end past e, so the clamp to e covers them.
            const uint32_t nxt = (l0 & ~1u) + P;
            l0 = nxt < e ? nxt : e;
            if constexpr (LAT == LAT_HBM) {
                // lanes of a replica must see each other's byte stores next round
                if (((mA & ~invA) | (mB & ~invB)) != 0ull) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
            }
        } else {
            if (__builtin_amdgcn_inverse_ballot_w64(mA)) lat.flip(A.st.c, A.s01);
            if (__builtin_amdgcn_inverse_ballot_w64(mB & ~oB)) lat.flip(B.st.c, B.s01);
            if (__builtin_amdgcn_inverse_ballot_w64(eB & oB)) {
                l0 = lB;
            } else {
                const uint32_t nxt = lA + 2;
                l0 = nxt < e ? nxt : e;
            }
        }
    }

    // ---------------------------------------------------------- epilogue
    if constexpr (LAT == LAT_MULTISPIN) {
        __syncthreads();
        // thread t counts bit (t & 63) over sites t>>6, t>>6 + S, ...
        const int b = tid & 63;
        const int q = tid >> 6;
        const uint64_t *w = (const uint64_t *)lat_lds;
        unsigned c = 0;
        for (int site = q; site < N; site += S) c += (unsigned)((w[site] >> b) & 1ull);
        atomicAdd(&cnt_lds[b], c);
        __syncthreads();
        if (tid < 64) finish_replica(a, (r0 + tid) < a.R, r0 + tid, cnt_lds[tid]);
        uint4 *lo = (uint4 *)((unsigned *)a.lattice + (size_t)(2 * blockIdx.x) * a.stride);
        uint4 *hi = lo + a.stride / 4;
        const uint4 *src = (const uint4 *)lat_lds;
        for (int i = tid; i < a.stride / 4; i += blockDim.x) {
            const uint4 p = src[2 * i], q = src[2 * i + 1];
            lo[i] = make_uint4(p.x, p.z, q.x, q.z);
            hi[i] = make_uint4(p.y, p.w, q.y, q.w);
        }
    } else if constexpr (LAT == LAT_MULTIHALF) {
        __syncthreads();
        // thread t counts bit (t & 31) over sites t>>5, t>>5 + S, ...
        const int b = tid & 31;
        const int q = tid >> 5;
        const unsigned *w = (const unsigned *)lat_lds;
        unsigned c = 0;
        for (int site = q; site < N; site += S) c += (w[site] >> b) & 1u;
        atomicAdd(&cnt_lds[b], c);
        __syncthreads();
        if (tid < 32) finish_replica(a, (r0 + tid) < a.R, r0 + tid, cnt_lds[tid]);
        uint4 *dst = (uint4 *)((unsigned *)a.lattice + (size_t)blockIdx.x * a.stride);
        const uint4 *src = (const uint4 *)lat_lds;
        for (int i = tid; i < a.stride / 4; i += blockDim.x) dst[i] = src[i];
    } else if constexpr (LAT == LAT_MULTI16 || LAT == LAT_MULTI8) {
        __syncthreads();
        // thread t counts bit (t mod RPG) over sites t / RPG, t / RPG + S, ...
        const int b = tid & (RPG - 1);
        const int q = tid / RPG;
        unsigned c = 0;
        for (int site = q; site < N; site += S) {
            const unsigned v = (LAT == LAT_MULTI16) ? (unsigned)((const unsigned short *)lat_lds)[site]
                                                    : (unsigned)lat_lds[site];
            c += (v >> b) & 1u;
        }
        atomicAdd(&cnt_lds[b], c);
        __syncthreads();
        if (tid < RPG) finish_replica(a, (r0 + tid) < a.R, r0 + tid, cnt_lds[tid]);
        // the workgroup's bytes of each dword of the plane (other workgroups
        // write the other bytes)
        unsigned char *dst = (unsigned char *)((unsigned *)a.lattice + (size_t)(r0 >> 5) * a.stride) + ((r0 & 31) >> 3);
        for (int i = tid; i < N; i += blockDim.x) {
            if constexpr (LAT == LAT_MULTI16) *(unsigned short *)(dst + 4 * (size_t)i) = ((const unsigned short *)lat_lds)[i];
            else dst[4 * (size_t)i] = lat_lds[i];
        }
    } else {
        if constexpr (LAT == LAT_BITS) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        unsigned c = 0;
        if (live) {
            if constexpr (LAT == LAT_BITS) {
                const unsigned *w = (const unsigned *)lat.base;
                for (int i = slot; i < a.stride; i += S) c += __popc(w[i]);
            } else {
                const unsigned *w = (const unsigned *)lat.base;
                for (int i = slot; i < a.stride / 4; i += S)
                    c += __popc(__hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) &
                                0x01010101u);
            }
        }
#pragma unroll
        for (int off = S / 2; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
        finish_replica(a, live && slot == 0, r, c);
        if constexpr (LAT == LAT_BITS) {
            const int nrep = min(a.rpw, a.R - r0);
            const int ndw = nrep * a.stride;
            unsigned *dst = (unsigned *)a.lattice + (size_t)r0 * a.stride;
            const unsigned *src = (const unsigned *)lat_lds;
            for (int i = tid; i < ndw; i += blockDim.x) dst[i] = src[i];
        }
    }
}

// ------------------------------------------------------------ multi-wave HBM sweep
// Layout 0 with 64 W lanes per replica spread over the W waves (2, 4, 8 or 16)
// of one workgroup, one replica per workgroup.  For large lattices with few
// replicas (the L = 256 group of a batched run has 256), one wave per replica
// leaves three of each CU's four SIMDs idle and the round latency exposed.
//
// Windowed evaluation, one attempt per lane: a round evaluates a window of
// 64 W attempts (launch-local indices) against the lattice as it was at the
// start of the round.  Attempt j commits iff no earlier attempt of the window
// has its site in j's stencil; otherwise it stays on its lane and is evaluated
// again next round, before any newer attempt.  A committed attempt has read
// what the sequential chain would have shown it: the earlier attempts it
// depends on committed in earlier rounds, and the others neither read nor
// wrote what it reads or writes; the window's earliest attempt always
// commits.  The other lanes take the next attempts in wave and lane order.
// Dependences come from an LDS table of attempt indices written with atomic
// minima at entry site & tbl_mask (sites that share an entry only add
// deferrals): an attempt whose five stencil entries hold an index below its
// own has an earlier attempt of the window in its stencil.  Two tables
// alternate by round, and each lane clears its entry of the previous round's
// table during the next round.  Two barriers per round: after the marks,
// and after the flips (which the next round's reads must see) and the counts
// of deferred attempts per wave.
// STAGE: 0, the replica's bytes stay in HBM; 1, they are staged in LDS for
// the launch as bytes (layout 6), 2, as bits (layout 7), at kLdsLattice, and
// the tables follow them.
template <int RNG, int W, bool POW2, int STAGE>
__global__ void __launch_bounds__(64 * W) sweep_hbm_waves(SweepArgs a) {
    static_assert(W == 2 || W == 4 || W == 8 || W == 16, "2, 4, 8 or 16 waves per replica");
    static_assert(STAGE >= 0 && STAGE <= 2, "HBM, LDS bytes or LDS bits");
    constexpr int S = 64 * W;
    constexpr int LAT = STAGE == 1 ? LAT_LDS8 : STAGE == 2 ? LAT_LDS1 : LAT_HBM;
    constexpr unsigned kNone = 0xffffffffu;
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *thr_lds = (uint64_t *)(smem + kLdsThr);
    unsigned *cnt_lds = (unsigned *)(smem + kLdsCnt);  // [0]: up spins; [16 + 16 par + wave]: deferred attempts
    const unsigned hm = a.tbl_mask;
    const unsigned tsz = 4u * (hm + 1u);  // bytes per table; table p at tbase + p tsz
    // a.stride (bytes per replica) is a multiple of 16; layout 7 takes N bits
    // rounded up to 16 bytes
    const unsigned nbw = ((unsigned)a.N + 31u) >> 5;  // layout 7: dwords
    const unsigned tbase = (unsigned)kLdsLattice +
                           (STAGE == 1 ? (unsigned)a.stride : STAGE == 2 ? ((nbw + 3u) >> 2) << 4 : 0u);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int r = blockIdx.x;  // grid = R workgroups
    // uniform over the workgroup
    const bool active = !(a.freeze && a.fate[r] != -1);
    if (lds_origin(smem) != 0u || (a.rng_dev && !rng_header_ok(a.rng_dev, a.R))) {  // as in sweep_kernel
        if (tid == 0) a.mag[r] = __builtin_nanf("");
        return;
    }

    stage_thresholds(a, thr_lds, tid);
    if (tid == 0) cnt_lds[0] = 0u;
    for (unsigned i = (unsigned)tid; i < 2u * (hm + 1u); i += S) *lds_ptr32(tbase + 4u * i) = kNone;
    Lattice<LAT> lat;
    lat.base = (unsigned char *)a.lattice + (size_t)r * a.stride;
    lat.dw = nullptr;
    lat.rl = 0;
    lat.rsrc = __builtin_amdgcn_make_buffer_rsrc(lat.base, (short)0, a.stride, kBufferWord3);
    lat.off = 0u;
    if constexpr (STAGE == 1) {
        const uint4 *src = (const uint4 *)lat.base;
        uint4 *dst = (uint4 *)(smem + kLdsLattice);
        for (int i = tid; i < a.stride / 16; i += S) dst[i] = src[i];
    } else if constexpr (STAGE == 2) {
        // bit k of dword i from byte 32 i + k (0 or 1); loads past the
        // replica's stride read 0, and the bits of sites from N on are cleared
        for (unsigned i = (unsigned)tid; i < nbw; i += S) {
            unsigned v = 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const unsigned w = __builtin_amdgcn_raw_buffer_load_b32(lat.rsrc, 32u * i + 4u * (unsigned)k, 0, 0);
                v |= ((w & 1u) | ((w >> 7) & 2u) | ((w >> 14) & 4u) | ((w >> 21) & 8u)) << (4 * k);
            }
            const unsigned rest = (unsigned)a.N - 32u * i;
            if (rest < 32u) v &= (1u << rest) - 1u;
            *lds_ptr32(kLdsLattice + 4u * i) = v;
        }
    }
    __syncthreads();

    // attempt l of the launch is global attempt a_begin + l: Philox pair
    // (a_begin + l) >> 1, first or second half by its low bit
    const uint64_t a_begin = a.rng_dev ? a.rng_dev->attempts : a.a_begin;
    const uint64_t seed = a.rng_dev ? a.rng_dev->seed : a.seed;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const Geometry<POW2> geo{a.L, a.N, a.lg};
    const uint32_t *drw = nullptr;
    if constexpr (RNG == RNG_DRAWS) drw = a.draws + (size_t)r * (3 * a.n_attempts);

    const uint32_t n = active ? (uint32_t)a.n_attempts : 0u;
    bool has = (uint32_t)tid < n;
    uint32_t l = (uint32_t)tid;  // launch-local attempt index of this lane
    uint32_t next = min((uint32_t)S, n);
    unsigned par = 0u, prev_ent = 0u;
    bool had_prev = false;
    for (;;) {
        Stencil st{};
        bool acc = false;
        int s01 = 0;
        if (has) {
            uint32_t ua;
            if constexpr (RNG == RNG_PHILOX) {
                const uint64_t A = a_begin + (uint64_t)l;
                const uint64_t pair = A >> 1;
                uint32_t c0 = (uint32_t)pair, c1 = (uint32_t)(pair >> 32), c2 = (uint32_t)r, c3 = kPhiloxDomain;
                philox4x32_10(c0, c1, c2, c3, k0, k1);
                const bool second = (A & 1ull) != 0ull;
                st = geo.philox(second ? c2 : c0);
                ua = second ? c3 : c1;
            } else {
                const uint32_t *d = drw + 3 * (size_t)l;
                st = geo.mt(d[0], d[1]);
                ua = d[2];
            }
            __hip_atomic_fetch_min(lds_ptr32(tbase + par * tsz + 4u * ((unsigned)st.c & hm)), l,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const Spins x = read_spins<LAT>(lat, st);
            const Attempt t = decide(st, x, ua);
            acc = t.acc;
            s01 = t.s01;
        }
        // the previous round's mark, in the other table (read by nobody now)
        if (had_prev) *lds_ptr32(tbase + (par ^ 1u) * tsz + 4u * prev_ent) = kNone;
        __syncthreads();
        bool deferred = false;
        if (has) {
            const unsigned tb = tbase + par * tsz;
            const uint32_t m = min(min(min(lds_load32(tb + 4u * ((unsigned)st.c & hm)),
                                           lds_load32(tb + 4u * ((unsigned)st.up & hm))),
                                       min(lds_load32(tb + 4u * ((unsigned)st.dn & hm)),
                                           lds_load32(tb + 4u * ((unsigned)st.lt & hm)))),
                                   lds_load32(tb + 4u * ((unsigned)st.rt & hm)));
            deferred = m < l;
            if (!deferred && acc) lat.flip(st.c, s01);
        }
        const uint64_t dmask = wave_mask(deferred);
        if (lane == 0) cnt_lds[16u + 16u * par + (unsigned)wave] = (uint32_t)__popcll(dmask);
        had_prev = has;
        prev_ent = (unsigned)st.c & hm;
        // flips visible to every wave before the next round's reads
        __syncthreads();
        // deferred attempts keep their lanes; the others take the next new
        // attempts in wave order, lane order
        uint32_t D = 0u, before = 0u;
#pragma unroll
        for (int v = 0; v < W; ++v) {
            const uint32_t dv = cnt_lds[16u + 16u * par + (unsigned)v];
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
        next += min((uint32_t)S - D, rem);
        par ^= 1u;
    }

    // ---------------------------------------------------------- epilogue
    __syncthreads();
    unsigned c = 0;
    if constexpr (STAGE == 2) {
        // the bits back to bytes (sites from N on, within the stride, get 0)
        for (unsigned i = (unsigned)tid; i < nbw; i += S) {
            const unsigned v = lds_load32(kLdsLattice + 4u * i);
            c += __popc(v);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const unsigned q = v >> (4 * k);
                const unsigned w = (q & 1u) | ((q & 2u) << 7) | ((q & 4u) << 14) | ((q & 8u) << 21);
                __builtin_amdgcn_raw_buffer_store_b32(w, lat.rsrc, 32u * i + 4u * (unsigned)k, 0, 0);
            }
        }
    } else if constexpr (STAGE == 1) {
        const uint4 *src = (const uint4 *)(smem + kLdsLattice);
        uint4 *dst = (uint4 *)lat.base;
        for (int i = tid; i < a.stride / 16; i += S) {
            const uint4 v = src[i];
            c += __popc(v.x & 0x01010101u) + __popc(v.y & 0x01010101u) + __popc(v.z & 0x01010101u) +
                 __popc(v.w & 0x01010101u);
            dst[i] = v;
        }
    } else {
        const unsigned *w = (const unsigned *)lat.base;
        for (int i = tid; i < a.stride / 4; i += S)
            c += __popc(__hip_atomic_load(w + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & 0x01010101u);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
    if (lane == 0) atomicAdd(&cnt_lds[0], c);
    __syncthreads();
    if (wave == 0) finish_replica(a, tid == 0, r, cnt_lds[0]);
}

// ------------------------------------------------------------ layout kernels
// int32 +-1 grids [R][N] <-> native layouts.  1-D grid-stride loops over
// (replica or group, element) so that any R fits the launch limits.
// `per` consecutive replicas share one source grid (1: one grid per replica;
// R / ninputs: committor inputs replicated on the device).
// Layout 2 in HBM: for group g (replicas 64g .. 64g+63) two dword planes of
// `stride` dwords: plane 2g holds bit b < 32 of site s as bit b of dword s,
// plane 2g + 1 the bits 32..63.  A split workgroup stages one plane.
__global__ void pack_multispin(const int *__restrict__ src, unsigned *__restrict__ dst, int R, int N, int stride,
                               int per) {
    const int G = (R + 63) / 64;
    const size_t total = (size_t)G * N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(i / N), site = (int)(i - (size_t)g * N);
        uint64_t w = 0;
        const int nb = min(64, R - 64 * g);
        for (int b = 0; b < nb; ++b)
            if (src[(size_t)((64 * g + b) / per) * N + site] == 1) w |= 1ull << b;
        dst[(size_t)(2 * g) * stride + site] = (unsigned)w;
        dst[(size_t)(2 * g + 1) * stride + site] = (unsigned)(w >> 32);
    }
}

template <typename T>
__global__ void unpack_multispin(const unsigned *__restrict__ src, T *__restrict__ dst, int R, int N, int stride) {
    const int G = (R + 63) / 64;
    const size_t total = (size_t)G * N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int g = (int)(i / N), site = (int)(i - (size_t)g * N);
        const uint64_t w = (uint64_t)src[(size_t)(2 * g) * stride + site] |
                           ((uint64_t)src[(size_t)(2 * g + 1) * stride + site] << 32);
        const int nb = min(64, R - 64 * g);
        for (int b = 0; b < nb; ++b) dst[(size_t)(64 * g + b) * N + site] = ((w >> b) & 1ull) ? T(1) : T(-1);
    }
}

__global__ void pack_bits(const int *__restrict__ src, unsigned *__restrict__ dst, int R, int N, int stride,
                          int per) {
    const size_t total = (size_t)R * stride;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / stride), w = (int)(i - (size_t)r * stride);
        unsigned v = 0;
        for (int k = 0; k < 32; ++k) {
            const int site = 32 * w + k;
            if (site < N && src[(size_t)(r / per) * N + site] == 1) v |= 1u << k;
        }
        dst[i] = v;
    }
}

template <typename T>
__global__ void unpack_bits(const unsigned *__restrict__ src, T *__restrict__ dst, int R, int N, int stride) {
    const size_t total = (size_t)R * N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / N), site = (int)(i - (size_t)r * N);
        const unsigned v = src[(size_t)r * stride + (site >> 5)];
        dst[i] = ((v >> (site & 31)) & 1u) ? T(1) : T(-1);
    }
}

__global__ void pack_hbm(const int *__restrict__ src, unsigned char *__restrict__ dst, int R, int N, int stride,
                         int per) {
    const size_t total = (size_t)R * stride;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / stride), site = (int)(i - (size_t)r * stride);
        dst[i] = (site < N && src[(size_t)(r / per) * N + site] == 1) ? 1 : 0;
    }
}

template <typename T>
__global__ void unpack_hbm(const unsigned char *__restrict__ src, T *__restrict__ dst, int R, int N, int stride) {
    const size_t total = (size_t)R * N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / N), site = (int)(i - (size_t)r * N);
        dst[i] = src[(size_t)r * stride + site] ? T(1) : T(-1);
    }
}

// ------------------------------------------------------------ MT19937 draws
// Per-replica states interleaved [624][R] (word k of replica r at k*R + r).
__global__ void mt_seed_kernel(uint32_t *st, int R, uint64_t seed_base) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    uint32_t x = (uint32_t)(seed_base + (uint64_t)r);
    st[r] = x;
    for (int k = 1; k < 624; ++k) {
        x = 1812433253u * (x ^ (x >> 30)) + (uint32_t)k;
        st[(size_t)k * R + r] = x;
    }
}

// One wave per replica: the 624-word state lives in LDS; each twist is done in
// four dependent phases (words 0..226 use only old words; 227..453 use new
// words 0..226; 454..622 use new words 227..395; 623 uses new words 0 and 396),
// then the block is tempered and written to the replica's draw row.
// Draws [q0, q0 + ndraws) of the replica's stream go to out[r*ndraws + i].
__global__ void __launch_bounds__(64) mt_draws_kernel(uint32_t *st, int R, uint64_t q0, uint64_t ndraws,
                                                      uint32_t *out) {
    __shared__ uint32_t mt[624];
    const int r = blockIdx.x;
    const int t = threadIdx.x;
    for (int k = t; k < 624; k += 64) mt[k] = st[(size_t)k * R + r];
    __syncthreads();
    // position within the current block: after q draws from a fresh seed the
    // reference's mti is 624 when q == 0, else ((q-1) % 624) + 1
    int mti = (q0 == 0) ? 624 : (int)((q0 - 1) % 624) + 1;
    uint64_t done = 0;
    uint32_t *o = out + (size_t)r * ndraws;
    auto twist_word = [&](int k, uint32_t next, uint32_t far) {
        const uint32_t y = (mt[k] & 0x80000000u) | (next & 0x7fffffffu);
        return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    };
    while (done < ndraws) {
        if (mti >= 624) {
            // phase 1: k in [0, 227)
            for (int base = 0; base < 227; base += 64) {
                const int k = base + t;
                uint32_t nv = 0;
                if (k < 227) nv = twist_word(k, mt[k + 1], mt[k + 397]);
                __syncthreads();
                if (k < 227) mt[k] = nv;
                __syncthreads();
            }
            // phase 2: k in [227, 454)
            for (int base = 227; base < 454; base += 64) {
                const int k = base + t;
                uint32_t nv = 0;
                if (k < 454) nv = twist_word(k, mt[k + 1], mt[k - 227]);
                __syncthreads();
                if (k < 454) mt[k] = nv;
                __syncthreads();
            }
            // phase 3: k in [454, 623)
            for (int base = 454; base < 623; base += 64) {
                const int k = base + t;
                uint32_t nv = 0;
                if (k < 623) nv = twist_word(k, mt[k + 1], mt[k - 227]);
                __syncthreads();
                if (k < 623) mt[k] = nv;
                __syncthreads();
            }
            if (t == 0) mt[623] = twist_word(623, mt[0], mt[396]);
            __syncthreads();
            mti = 0;
        }
        // temper and emit min(624 - mti, remaining) draws
        const uint64_t take64 = (uint64_t)(624 - mti) < (ndraws - done) ? (uint64_t)(624 - mti) : (ndraws - done);
        const int take = (int)take64;
        for (int i = t; i < take; i += 64) {
            uint32_t y = mt[mti + i];
            y ^= (y >> 11);
            y ^= (y << 7) & 0x9d2c5680u;
            y ^= (y << 15) & 0xefc60000u;
            y ^= (y >> 18);
            o[done + i] = y;
        }
        mti += take;
        done += take64;
        __syncthreads();
    }
    for (int k = t; k < 624; k += 64) st[(size_t)k * R + r] = mt[k];
}

// ------------------------------------------------------------ small kernels
// (kernels rather than memset / host updates, so that they can be captured in
// a graph and replayed)
// The body of a gridstates.bin record (src/io.c:67-130): the R*N spins as one
// bit stream, replica-major then site-major, bit k of byte b = spin 8b + k
// (1 = up), padded with zero bits to whole bytes.  One thread per output byte.
__global__ void pack_gridstates(const void *__restrict__ lat, int method, int stride, int R, int N,
                                uint8_t *__restrict__ out, size_t nbytes) {
    const size_t total = (size_t)R * N;
    for (size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x; b < nbytes; b += (size_t)gridDim.x * blockDim.x) {
        unsigned v = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const size_t i = 8 * b + k;
            if (i >= total) break;
            const int r = (int)(i / N), site = (int)(i - (size_t)r * N);
            unsigned up;
            if (method == 2) {
                up = (((const unsigned *)lat)[(size_t)(r >> 5) * stride + site] >> (r & 31)) & 1u;
            } else if (method == 1) {
                up = (((const unsigned *)lat)[(size_t)r * stride + (site >> 5)] >> (site & 31)) & 1u;
            } else {
                up = ((const unsigned char *)lat)[(size_t)r * stride + site];
            }
            v |= up << k;
        }
        out[b] = (uint8_t)v;
    }
}

// Selected replicas idx[0..n) as int8 +-1 grids [n][N], from any native layout.
__global__ void gather_spins8(const void *__restrict__ lat, int method, int stride, const int *__restrict__ idx,
                              int n, int N, int8_t *__restrict__ out) {
    const size_t total = (size_t)n * N;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i / N), site = (int)(i - (size_t)k * N);
        const int r = idx[k];
        int up;
        if (method == 2) {
            up = (int)((((const unsigned *)lat)[(size_t)(r >> 5) * stride + site] >> (r & 31)) & 1u);
        } else if (method == 1) {
            up = (int)((((const unsigned *)lat)[(size_t)r * stride + (site >> 5)] >> (site & 31)) & 1u);
        } else {
            up = ((const unsigned char *)lat)[(size_t)r * stride + site];
        }
        out[i] = up ? (int8_t)1 : (int8_t)-1;
    }
}

__global__ void zero_ints(int *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}
// Advances the attempt counter of a device RNG header by n, if the header
// is one the sweep kernels accept for R replicas (they did not sweep otherwise).
__global__ void advance_counter(RngHeader *hdr, int R, uint64_t n) {
    if (threadIdx.x == 0 && rng_header_ok(hdr, R)) hdr->attempts += n;
}

// ------------------------------------------------------------ checks
// Threshold check on an arbitrary per-replica CV array (largest-cluster CV).
__global__ void check_cv_kernel(const float *cv, int R, int check, double dn_thr, double up_thr,
                                int *fate, int *counts) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool own = r < R;
    bool gt_up = false, is_a = false, is_b = false;
    if (own) {
        const double v = (double)cv[r];
        gt_up = v > up_thr;
        if (check == CHECK_FATES) {
            int f = fate[r];
            if (f == -1) {
                if (v > up_thr) f = 1;
                else if (v < dn_thr) f = 0;
                if (f != -1) fate[r] = f;
            }
            is_a = f == 0;
            is_b = f == 1;
        }
    }
    const uint64_t b0 = wave_mask(gt_up), ba = wave_mask(is_a), bb = wave_mask(is_b);
    if ((threadIdx.x & 63) == 0) {
        if (check == CHECK_COUNT && b0) atomicAdd(&counts[0], (int)__popcll(b0));
        if (check == CHECK_FATES) {
            if (ba) atomicAdd(&counts[1], (int)__popcll(ba));
            if (bb) atomicAdd(&counts[2], (int)__popcll(bb));
        }
    }
}

}  // namespace gasp
