// gasp_cnn.hip -- committor-surrogate CNN: training and inference kernels for
// gfx950 (BF16 MFMA, FP32 accumulation, FP32 master weights, Adam).
//
// No counterpart in the reference: this is the learned addition of the engine.
// Network (per grid of L x L spins, L a power of two >= 16):
//   conv1 1->32, conv2..conv4 32->32: 3x3 circular cross-correlation + bias + ReLU
//   global average pool -> fc1 32->64 + ReLU -> fc2 64->1 -> sigmoid = pB
// trained with binary cross-entropy against shooting labels.
//
// Storage: activations NHWC bfloat16 [n][L][L][32]; conv2..4 weights as MFMA
// fragments (bfloat16); conv1 weights and the MLP in float32.  Parameters are
// one float32 vector (30241 entries) in the order
//   w1[9][32] b1[32] w2[9][32][32] b2[32] w3 b3 w4 b4 fw1[32][64] fb1[64] fw2[64] fb2[1]
// with taps t = 3 (dy + 1) + (dx + 1) and w[t][cin][cout].
//
// MFMA use (v_mfma_f32_16x16x32_bf16, D[16x16] += A[16x32] B[32x16]):
//   lane l holds A[l % 16][8 (l / 16) + i], B[8 (l / 16) + i][l % 16] (i < 8) and
//   D[4 (l / 16) + j][l % 16] (j < 4).
//   forward / backward-data:  D[cout][pos] += W^T[cout][cin] X[cin][pos], one
//     MFMA per tap and 16-cout half: 18 per 16-position tile.
//   weight gradient:          D[cout][cin] += dZ^T[cout][pos] X_t[pos][cin] over
//     32-position chunks staged transposed in LDS, all nine taps per chunk.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gasp_internal.h"

using gasp::set_error;

namespace cnn {

constexpr int C = 32;
constexpr int HID = 64;
constexpr int NPARAM = 30241;
constexpr int OFF_W1 = 0, OFF_B1 = 288;
constexpr int OFF_W[5] = {0, 0, 320, 9568, 18816};     // conv k = 2..4
constexpr int OFF_B[5] = {0, 288, 9536, 18784, 28032};
constexpr int OFF_FW1 = 28064, OFF_FB1 = 30112, OFF_FW2 = 30176, OFF_FB2 = 30240;
constexpr int HEAD_N = NPARAM - OFF_FW1;               // 2177
constexpr int kRedChunk = 32;                          // partials per first-pass chunk

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest, ties to even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ f32x4 mfma(uint4 a, uint4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                   c, 0, 0, 0);
}

// ------------------------------------------------------------------ weights
// Fragment layout of conv k (k = 2..4): frag[(t * 2 + half) * 64 + lane] = 8
// bfloat16 of W^T rows: lane l -> cout = 16 half + l % 16, cin = 8 (l / 16) + i.
// Backward-data weights are those of the transposed, tap-flipped convolution:
// W'[t'][co][ci] = W[8 - t'][ci][co].
__global__ void prep_weights(const float *__restrict__ p, uint16_t *__restrict__ wf, uint16_t *__restrict__ wb,
                             float *__restrict__ w1r) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    // 3 layers x 9 taps x 2 halves x 64 lanes x 8 = 27648 entries per layout
    if (tid < 3 * 9 * 2 * 64 * 8) {
        const int i = tid & 7, l = (tid >> 3) & 63, half = (tid >> 9) & 1, t = (tid >> 10) % 9, k = tid / (9 * 1024);
        const float *w = p + OFF_W[k + 2];
        const int co = 16 * half + (l & 15), ci = 8 * (l >> 4) + i;
        wf[tid] = f2bf(w[(t * C + ci) * C + co]);
        // transposed conv: output channel = ci of the forward, input channel = co
        const int tf = 8 - t;
        wb[tid] = f2bf(w[(tf * C + co) * C + ci]);
    }
    if (tid < 288) w1r[tid] = bf2f(f2bf(p[OFF_W1 + tid]));
}

// ------------------------------------------------------------------ conv1
// a1[n][p][c] = bf16(relu(b1[c] + sum_t w1[t][c] s[p + off_t]))
__global__ void conv1_fwd(const int8_t *__restrict__ s, const float *__restrict__ w1r, const float *__restrict__ p,
                          uint16_t *__restrict__ a1, int n, int L) {
    const int N = L * L;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * N) return;
    const int img = (int)(idx / N), pos = (int)(idx % N);
    const int y = pos / L, x = pos % L, m = L - 1;
    float sv[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int yy = (y + t / 3 - 1) & m, xx = (x + t % 3 - 1) & m;
        sv[t] = (float)s[(size_t)img * N + yy * L + xx];
    }
    uint32_t out[16];
#pragma unroll
    for (int c2 = 0; c2 < 16; ++c2) {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = 2 * c2 + h;
            float z = p[OFF_B1 + c];
#pragma unroll
            for (int t = 0; t < 9; ++t) z += w1r[t * C + c] * sv[t];
            v[h] = fmaxf(z, 0.0f);
        }
        out[c2] = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
    }
    uint4 *dst = (uint4 *)(a1 + idx * C);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
}

// ------------------------------------------------------------------ conv 32->32
// A workgroup of 4 waves computes 4 consecutive rows of one image (L is a
// multiple of 4); wave w owns row y0 + w, in 16-position tiles x 32 output
// channels.  The input band (rows y0 - 1 .. y0 + 4, periodic) is staged in LDS
// one x-chunk of up to 64 positions (+ 1 halo position each side) at a time,
// so each input element is read from memory about 1.5 times instead of once
// per tap; the 9 taps then take their B fragments from LDS.  LDS positions
// are 80 bytes apart (32 channels + 8 pad), which makes the 16 lanes of a
// 16-byte fragment read start on 16 different bank quads.
// BWD = false: out = bf16(relu(acc + bias)); BWD = true (backward data with
// the flipped/transposed weights): out = bf16(acc * (mask > 0)).
constexpr int kConvRows = 4, kConvXT = 64, kConvPS = 40;  // rows per block, x-chunk, position stride (bf16)

template <bool BWD>
__global__ void __launch_bounds__(256) conv_mfma(const uint16_t *__restrict__ in, const uint16_t *__restrict__ wfrag,
                                                 const float *__restrict__ bias, const uint16_t *__restrict__ mask,
                                                 uint16_t *__restrict__ out, int n, int L) {
    __shared__ __align__(16) uint16_t tile[(kConvRows + 2) * (kConvXT + 2) * kConvPS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = blockIdx.x * kConvRows;  // grid = n * L / kConvRows exactly
    const int img = row0 / L, y0 = row0 % L, m = L - 1, N = L * L;
    uint4 wa[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h) wa[t][h] = ((const uint4 *)wfrag)[(t * 2 + h) * 64 + lane];
    const int pl = lane & 15, kq = lane >> 4;
    const uint16_t *base = in + (size_t)img * N * C;
    const int xw = L < kConvXT ? L : kConvXT;
    const int y = y0 + w;
    for (int xc = 0; xc < L; xc += xw) {
        __syncthreads();  // the previous chunk's fragments have been read
        const int items = (kConvRows + 2) * (xw + 2) * 4;  // 16-byte pieces: 4 per position
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            const int q = it & 3, p = (it >> 2) % (xw + 2), r = (it >> 2) / (xw + 2);
            const int yy = (y0 + r - 1) & m, xx = (xc + p - 1) & m;
            *(uint4 *)&tile[(r * (kConvXT + 2) + p) * kConvPS + 8 * q] =
                *(const uint4 *)(base + ((size_t)(yy * L + xx) * C + 8 * q));
        }
        __syncthreads();
        for (int x0 = 0; x0 < xw; x0 += 16) {
            f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const int r = w + t / 3, p = x0 + pl + t % 3;
                const uint4 b = *(const uint4 *)&tile[(r * (kConvXT + 2) + p) * kConvPS + 8 * kq];
                acc[0] = mfma(wa[t][0], b, acc[0]);
                acc[1] = mfma(wa[t][1], b, acc[1]);
            }
            const size_t opos = (size_t)img * N + (size_t)y * L + xc + x0 + pl;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c0 = 16 * h + 4 * kq;
                float v[4];
                if constexpr (BWD) {
                    const uint2 mk = *(const uint2 *)(mask + opos * C + c0);
                    const uint16_t mb[4] = {(uint16_t)mk.x, (uint16_t)(mk.x >> 16), (uint16_t)mk.y,
                                            (uint16_t)(mk.y >> 16)};
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = bf2f(mb[j]) > 0.0f ? acc[h][j] : 0.0f;
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = fmaxf(acc[h][j] + bias[c0 + j], 0.0f);
                }
                uint2 o;
                o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *(uint2 *)(out + opos * C + c0) = o;
            }
        }
    }
}

// ------------------------------------------------------------------ weight gradient
// dW[t][ci][co] = sum over positions p of a[p + off_t][ci] dz[p][co]: a GEMM
// whose K is the positions.  A workgroup takes bands of 4 rows x XW = min(L,
// 64) columns of one image in turn.  Per band it stages, through registers,
// dz transposed (zT[co][row][x]) and the input rows y0 - 1 .. y0 + 4
// transposed with one periodic halo column each side (aT[ci][row][8 + x],
// x = -1 .. XW): a thread loads four positions x eight channels (four 16-byte
// loads) and writes eight 8-byte channel runs.  The band's positions then go
// to the waves as K-chunks of 32 (four runs of 8): the A fragment (dz^T, 8
// positions of one co per lane) is one 16-byte LDS read; for each kernel row
// one 16-byte read plus the dword either side give the B fragments of the
// taps dx = -1, 0, +1 (positions shifted by one bf16 with v_alignbit).  Waves
// 2 j and 2 j + 1 take the same K-chunks for ci half 0 and 1, so a K-chunk
// costs a wave 2 + 3 x 3 LDS reads for 18 MFMAs (nine taps, two co halves)
// and the nine taps' accumulators (72 registers) leave room for three
// workgroups per CU (160 registers a lane, 49 KB of LDS each).  They stay in registers over all of the workgroup's
// bands; at the end the four waves add theirs in a fixed order through LDS
// and the workgroup writes one partial row part[blk] = [dW[t][ci][co] |
// db[co] (sum of dz)], the order of the parameter vector, so one reduction
// covers both and the result does not depend on scheduling.  Channel runs are 260 (aT) and 132 (zT) dwords apart,
// 4 mod 64: the 16 lanes of a fragment read start on different bank quads.
constexpr int kWgXW = 64;                   // band width (columns)
constexpr int kWgARow = 40, kWgAPitch = 260;  // aT: dwords per row (8 + 64 + 8 bf16), per channel (6 rows + 20)
constexpr int kWgZRow = 32, kWgZPitch = 132;  // zT: dwords per row (64 bf16), per channel (4 rows + 4)
constexpr int kWgBlocks = 768;              // workgroups (three per CU on 256 CUs)

__device__ __forceinline__ uint32_t bf_pair(uint32_t hi_src, uint32_t lo_src, uint32_t sel) {
    return __builtin_amdgcn_perm(hi_src, lo_src, sel);
}

// channel c = cb + 2 q + s of four positions' dwords q: s = 0 the low halves, 1 the high
__device__ constexpr uint32_t kSel[2] = {0x05040100u, 0x07060302u};

// zT[co][rr][x] = dz of band row rr (image row y0 + rr), column x0 + x, for
// rr < 4 and x < XW: items of 4 positions x 8 channels (four 16-byte loads,
// eight 8-byte channel runs), over the workgroup's 256 threads.
__device__ __forceinline__ void stage_dz_t(uint32_t *zT, const uint16_t *__restrict__ zb, int L, int y0, int x0,
                                           int XW, int tid) {
    for (int it = tid; it < 4 * XW; it += 256) {
        const int cb = (it & 3) * 8, pq = it >> 2;
        const int rr = pq / (XW / 4), x = 4 * (pq % (XW / 4));
        const uint4 *src = (const uint4 *)(zb + ((size_t)(y0 + rr) * L + x0 + x) * C + cb);
        const uint4 v[4] = {src[0], src[4], src[8], src[12]};  // positions x .. x + 3
        const uint32_t d[4][4] = {{v[0].x, v[0].y, v[0].z, v[0].w}, {v[1].x, v[1].y, v[1].z, v[1].w},
                                  {v[2].x, v[2].y, v[2].z, v[2].w}, {v[3].x, v[3].y, v[3].z, v[3].w}};
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int sh = 0; sh < 2; ++sh) {
                uint2 run;
                run.x = bf_pair(d[1][q], d[0][q], kSel[sh]);
                run.y = bf_pair(d[3][q], d[2][q], kSel[sh]);
                *(uint2 *)&zT[(cb + 2 * q + sh) * kWgZPitch + rr * kWgZRow + x / 2] = run;
            }
    }
}

constexpr int kWgRow = 9 * C * C + C;  // partial row: weights then bias
static_assert(OFF_B[2] == OFF_W[2] + 9 * C * C && OFF_B[3] == OFF_W[3] + 9 * C * C && OFF_B[4] == OFF_W[4] + 9 * C * C,
              "each conv's bias follows its weights");

__global__ void __launch_bounds__(256) conv_wgrad(const uint16_t *__restrict__ ain, const uint16_t *__restrict__ dz,
                                                  float *__restrict__ part, int n, int L) {
    __shared__ __align__(16) uint32_t lds[C * kWgAPitch + C * kWgZPitch];
    uint32_t *aT = lds, *zT = lds + C * kWgAPitch;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int N = L * L, m = L - 1;
    const int XW = L < kWgXW ? L : kWgXW, nxc = L / XW;
    const int nbands = n * (L / 4) * nxc;
    const int pl = lane & 15, kq = lane >> 4;
    const int nh = w & 1;  // ci half of this wave
    f32x4 acc[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h) acc[t][h] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum[2] = {0.f, 0.f};
    for (int band = blockIdx.x; band < nbands; band += gridDim.x) {
        const int img = band / ((L / 4) * nxc), rem = band % ((L / 4) * nxc);
        const int y0 = 4 * (rem / nxc), x0 = XW * (rem % nxc);
        const uint16_t *zb = dz + (size_t)img * N * C;
        const uint16_t *ab = ain + (size_t)img * N * C;
        __syncthreads();  // the previous band's fragments have been read
        stage_dz_t(zT, zb, L, y0, x0, XW, tid);
        // input: rows y0 - 1 .. y0 + 4 x positions x0 .. x0 + XW - 1, then the halo columns
        for (int it = tid; it < 6 * XW + 48; it += 256) {
            if (it < 6 * XW) {
                const int cb = (it & 3) * 8, pq = it >> 2;
                const int r = pq / (XW / 4), x = 4 * (pq % (XW / 4));
                const int yy = (y0 + r - 1) & m;
                const uint4 *src = (const uint4 *)(ab + ((size_t)yy * L + x0 + x) * C + cb);
                const uint4 v[4] = {src[0], src[4], src[8], src[12]};
                const uint32_t d[4][4] = {{v[0].x, v[0].y, v[0].z, v[0].w}, {v[1].x, v[1].y, v[1].z, v[1].w},
                                          {v[2].x, v[2].y, v[2].z, v[2].w}, {v[3].x, v[3].y, v[3].z, v[3].w}};
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int sh = 0; sh < 2; ++sh) {
                        uint2 run;
                        run.x = bf_pair(d[1][q], d[0][q], kSel[sh]);
                        run.y = bf_pair(d[3][q], d[2][q], kSel[sh]);
                        *(uint2 *)&aT[(cb + 2 * q + sh) * kWgAPitch + r * kWgARow + (8 + x) / 2] = run;
                    }
            } else {
                const int hq = it - 6 * XW;  // 48 items: (row, side, channel block)
                const int cb = (hq & 3) * 8, side = (hq >> 2) & 1, r = hq >> 3;
                const int yy = (y0 + r - 1) & m;
                const int xx = side ? ((x0 + XW) & m) : ((x0 - 1) & m);
                const int idx = side ? 8 + XW : 7;  // bf16 index in the row
                const uint4 v = *(const uint4 *)(ab + ((size_t)yy * L + xx) * C + cb);
                const uint32_t d[4] = {v.x, v.y, v.z, v.w};
                uint16_t *a16 = (uint16_t *)aT;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a16[(size_t)(cb + 2 * q) * 2 * kWgAPitch + r * 2 * kWgARow + idx] = (uint16_t)d[q];
                    a16[(size_t)(cb + 2 * q + 1) * 2 * kWgAPitch + r * 2 * kWgARow + idx] = (uint16_t)(d[q] >> 16);
                }
            }
        }
        __syncthreads();
        // K-chunks of 32 positions: runs g = 4 kc + kq of 8 positions (row g / (XW / 8))
        for (int kc = w >> 1; kc < XW / 8; kc += 2) {
            const int g = 4 * kc + kq;
            const int rr = g / (XW / 8), xg = 8 * (g % (XW / 8));
            uint4 A[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) A[h] = *(const uint4 *)&zT[(16 * h + pl) * kWgZPitch + rr * kWgZRow + xg / 2];
            if (nh == 0) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const uint32_t v4[4] = {A[h].x, A[h].y, A[h].z, A[h].w};
#pragma unroll
                    for (int q = 0; q < 4; ++q) bsum[h] += bf2f((uint16_t)v4[q]) + bf2f((uint16_t)(v4[q] >> 16));
                }
            }
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int base = (16 * nh + pl) * kWgAPitch + (rr + dy) * kWgARow + (8 + xg) / 2;
                const uint4 Q = *(const uint4 *)&aT[base];
                const uint32_t Dp = aT[base - 1], Dn = aT[base + 4];
                uint4 Bm, Bp;
                Bm.x = __builtin_amdgcn_alignbit(Q.x, Dp, 16);
                Bm.y = __builtin_amdgcn_alignbit(Q.y, Q.x, 16);
                Bm.z = __builtin_amdgcn_alignbit(Q.z, Q.y, 16);
                Bm.w = __builtin_amdgcn_alignbit(Q.w, Q.z, 16);
                Bp.x = Bm.y;
                Bp.y = Bm.z;
                Bp.z = Bm.w;
                Bp.w = __builtin_amdgcn_alignbit(Dn, Q.w, 16);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    acc[3 * dy][h] = mfma(A[h], Bm, acc[3 * dy][h]);
                    acc[3 * dy + 1][h] = mfma(A[h], Q, acc[3 * dy + 1][h]);
                    acc[3 * dy + 2][h] = mfma(A[h], Bp, acc[3 * dy + 2][h]);
                }
            }
        }
    }

    // the waves' sums in a fixed order through LDS: waves 0 and 1 store
    // (ci halves 0 and 1), then waves 2 and 3 add
    float *red = (float *)lds;  // [9][ci][co] then the 32 bias sums
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float v = bsum[h];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        bsum[h] = v;  // lanes pl hold channel 16 h + pl
    }
    __syncthreads();
    for (int k = 0; k < 2; ++k) {
        if ((w >> 1) == k) {
#pragma unroll
            for (int t = 0; t < 9; ++t)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int co = 16 * h + 4 * kq + j, ci = 16 * nh + pl;
                        float *d = &red[t * C * C + ci * C + co];
                        *d = (k == 0) ? acc[t][h][j] : *d + acc[t][h][j];
                    }
            // bias sums: waves 0 (stores) and 2 (adds)
            if (nh == 0 && kq == 0) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    float *d = &red[9 * C * C + 16 * h + pl];
                    *d = (k == 0) ? bsum[h] : *d + bsum[h];
                }
            }
        }
        __syncthreads();
    }
    float *out = part + (size_t)blockIdx.x * kWgRow;
    for (int i = tid; i < kWgRow; i += 256) out[i] = red[i];
}

// ------------------------------------------------------------------ conv1 weight gradient
// dW1[t][c] = sum over positions p of s[p + off_t] dz1[p][c] and db1[c] =
// sum of dz1[p][c], on MFMA as D[c][t] += dz1^T[c][p] S[p][t] with the nine
// taps as N (lanes t >= 9 hold zeros).  Workgroups take bands of 4 rows x XW
// columns as conv_wgrad does: dz1 staged transposed (stage_dz_t), the spins
// of rows y0 - 1 .. y0 + 4 as bfloat16 +-1 with a periodic halo column each
// side.  Lane t's B fragment is 8 spins of its tap's row, shifted by its dx
// with v_alignbit (a per-lane shift).  The waves take the K-chunks in turn,
// add their sums in a fixed order through LDS, and each workgroup writes one
// partial [w1 (9 x 32) | b1 (32)] to part1[blk][320].
constexpr int kW1Blocks = 1024;

__global__ void __launch_bounds__(256) conv1_wgrad(const int8_t *__restrict__ s, const uint16_t *__restrict__ dz,
                                                   float *__restrict__ part1, int n, int L) {
    __shared__ __align__(16) uint32_t zT[C * kWgZPitch];
    __shared__ __align__(16) uint32_t sT[6 * kWgARow];  // [row][8 + x] bfloat16
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int N = L * L, m = L - 1;
    const int XW = L < kWgXW ? L : kWgXW, nxc = L / XW;
    const int nbands = n * (L / 4) * nxc;
    const int pl = lane & 15, kq = lane >> 4;
    // this lane's tap (t < 9; lanes 9..15 of each group contribute zeros)
    const bool tap_ok = pl < 9;
    const int tt = tap_ok ? pl : 4;
    const int tdy = tt / 3, tdx = tt % 3 - 1;
    f32x4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    float bsum[2] = {0.f, 0.f};

    for (int band = blockIdx.x; band < nbands; band += gridDim.x) {
        const int img = band / ((L / 4) * nxc), rem = band % ((L / 4) * nxc);
        const int y0 = 4 * (rem / nxc), x0 = XW * (rem % nxc);
        __syncthreads();  // the previous band's fragments have been read
        stage_dz_t(zT, dz + (size_t)img * N * C, L, y0, x0, XW, tid);
        uint16_t *s16 = (uint16_t *)sT;
        for (int it = tid; it < 6 * (XW + 2); it += 256) {
            const int r = it / (XW + 2), xi = it % (XW + 2);  // xi = x + 1, x = -1 .. XW
            const int yy = (y0 + r - 1) & m, xx = (x0 + xi - 1) & m;
            const int8_t v = s[(size_t)img * N + (size_t)yy * L + xx];
            s16[r * 2 * kWgARow + 7 + xi] = v > 0 ? (uint16_t)0x3F80u : (uint16_t)0xBF80u;
        }
        __syncthreads();
        for (int kc = w; kc < XW / 8; kc += 4) {
            const int g = 4 * kc + kq;
            const int rr = g / (XW / 8), xg = 8 * (g % (XW / 8));
            uint4 A[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) A[h] = *(const uint4 *)&zT[(16 * h + pl) * kWgZPitch + rr * kWgZRow + xg / 2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t v4[4] = {A[h].x, A[h].y, A[h].z, A[h].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) bsum[h] += bf2f((uint16_t)v4[q]) + bf2f((uint16_t)(v4[q] >> 16));
            }
            const int base = (rr + tdy) * kWgARow + (8 + xg) / 2;
            const uint4 Q = *(const uint4 *)&sT[base];
            const uint32_t Dp = sT[base - 1], Dn = sT[base + 4];
            // dx = -1: (Q_i : Q_i-1) >> 16; 0: Q_i; +1: (Q_i+1 : Q_i) >> 16
            const uint32_t q[6] = {Dp, Q.x, Q.y, Q.z, Q.w, Dn};
            const uint32_t sh = tdx == 0 ? 0u : 16u;
            uint32_t b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t lo = tdx < 0 ? q[i] : q[i + 1];
                const uint32_t hi = tdx > 0 ? q[i + 2] : q[i + 1];
                b[i] = tap_ok ? __builtin_amdgcn_alignbit(hi, lo, sh) : 0u;
            }
            const uint4 B = make_uint4(b[0], b[1], b[2], b[3]);
            acc[0] = mfma(A[0], B, acc[0]);
            acc[1] = mfma(A[1], B, acc[1]);
        }
    }

    // the waves' sums in a fixed order (wave 0, 1, 2, 3) through LDS
    float *red = (float *)zT;  // [9][32] then the 32 bias sums
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float v = bsum[h];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        bsum[h] = v;
    }
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
        if (w == k) {
            if (tap_ok) {
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float *d = &red[pl * C + 16 * h + 4 * kq + j];
                        *d = (k == 0) ? acc[h][j] : *d + acc[h][j];
                    }
            }
            if (kq == 0) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    float *d = &red[288 + 16 * h + pl];
                    *d = (k == 0) ? bsum[h] : *d + bsum[h];
                }
            }
        }
        __syncthreads();
    }
    float *out = part1 + (size_t)blockIdx.x * 320;
    for (int i = tid; i < 320; i += 256) out[i] = red[i];
}

// ------------------------------------------------------------------ head
// One block (256 threads) per image: global average pool, fc1 + ReLU, fc2,
// sigmoid; with labels also the loss, the head's parameter gradient
// (partial per image) and dg = d loss / d (pooled features).
__global__ void __launch_bounds__(256) head(const uint16_t *__restrict__ a4, const float *__restrict__ p,
                                            const float *__restrict__ labels, float *__restrict__ pb,
                                            float *__restrict__ loss, float *__restrict__ part_head,
                                            float *__restrict__ dg, int n, int L, float inv_batch) {
    __shared__ float red[4][C];
    __shared__ float g[C], hpre[HID], h[HID], dh[HID], zsh;
    const int img = blockIdx.x, tid = threadIdx.x, N = L * L;
    // pool: thread = (position lane tid / 4, channel octet tid % 4), one
    // 16-byte load of 8 channels per position
    const int oc = tid & 3, pg = tid >> 2;
    float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const uint4 *src = (const uint4 *)(a4 + (size_t)img * N * C) + oc;
#pragma unroll 4
    for (int q = pg; q < N; q += 64) {
        const uint4 v = src[(size_t)q * (C / 8)];
        const uint32_t d[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            sum[2 * k] += bf2f((uint16_t)d[k]);
            sum[2 * k + 1] += bf2f((uint16_t)(d[k] >> 16));
        }
    }
    // lanes with the same octet: xor over the position bits of the lane
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int off = 4; off < 64; off <<= 1) sum[k] += __shfl_xor(sum[k], off, 64);
    if ((tid & 63) < 4) {
#pragma unroll
        for (int k = 0; k < 8; ++k) red[tid >> 6][8 * oc + k] = sum[k];
    }
    __syncthreads();
    if (tid < C) g[tid] = (red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]) / (float)N;
    __syncthreads();
    if (tid < HID) {
        float v = p[OFF_FB1 + tid];
        for (int k = 0; k < C; ++k) v += g[k] * p[OFF_FW1 + k * HID + tid];
        hpre[tid] = v;
        h[tid] = fmaxf(v, 0.f);
    }
    __syncthreads();
    if (tid < 64) {
        float v = h[tid] * p[OFF_FW2 + tid];
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (tid == 0) zsh = v + p[OFF_FB2];
    }
    __syncthreads();
    const float z = zsh;
    const float sig = 1.0f / (1.0f + expf(-z));
    if (tid == 0) pb[img] = sig;
    if (!labels) return;
    const float y = labels[img];
    if (tid == 0) loss[img] = fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))) - y * z;
    const float dz = (sig - y) * inv_batch;
    float *ph = part_head + (size_t)img * HEAD_N;
    if (tid < HID) {
        ph[OFF_FW2 - OFF_FW1 + tid] = h[tid] * dz;
        const float d = hpre[tid] > 0.f ? dz * p[OFF_FW2 + tid] : 0.f;
        dh[tid] = d;
        ph[OFF_FB1 - OFF_FW1 + tid] = d;
    }
    if (tid == 0) ph[OFF_FB2 - OFF_FW1] = dz;
    __syncthreads();
    for (int e = tid; e < C * HID; e += 256) ph[e] = g[e / HID] * dh[e % HID];
    if (tid < C) {
        float v = 0.f;
        for (int j = 0; j < HID; ++j) v += dh[j] * p[OFF_FW1 + tid * HID + j];
        dg[img * C + tid] = v;
    }
}

// dz4[n][p][c] = bf16(dg[n][c] / N * (a4 > 0))
__global__ void head_bwd_fill(const uint16_t *__restrict__ a4, const float *__restrict__ dg, uint16_t *__restrict__ dz,
                              int n, int L) {
    const int N = L * L;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // one per (img, pos, channel pair)
    if (idx >= (size_t)n * N * (C / 2)) return;
    const size_t e = idx * 2;
    const int c = (int)(e % C);
    const int img = (int)(e / ((size_t)N * C));
    const uint32_t av = *(const uint32_t *)(a4 + e);
    const float inv = 1.0f / (float)N;
    const float v0 = bf2f((uint16_t)av) > 0.f ? dg[img * C + c] * inv : 0.f;
    const float v1 = bf2f((uint16_t)(av >> 16)) > 0.f ? dg[img * C + c + 1] * inv : 0.f;
    *(uint32_t *)(dz + e) = (uint32_t)f2bf(v0) | ((uint32_t)f2bf(v1) << 16);
}

// out[j] = sum_i part[i * M + j], fixed order (deterministic)
__global__ void reduce_partials(const float *__restrict__ part, int P, int M, float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    float s = 0.f;
    for (int i = 0; i < P; ++i) s += part[(size_t)i * M + j];
    out[j] = s;
}

// First pass of a two-level deterministic reduction: tmp[c][j] = sum of
// part[p][j] over the chunk p in [c * chunk, (c + 1) * chunk).
__global__ void reduce_chunks(const float *__restrict__ part, int P, int M, int chunk, float *__restrict__ tmp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const int p0 = blockIdx.y * chunk, p1 = min(P, p0 + chunk);
    float s = 0.f;
    for (int i = p0; i < p1; ++i) s += part[(size_t)i * M + j];
    tmp[(size_t)blockIdx.y * M + j] = s;
}

// Adam with bias correction, float32 master weights
__global__ void adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                     int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void gather_batch(const int8_t *__restrict__ data, const float *__restrict__ labels,
                             const int *__restrict__ order, int nb, int N, int8_t *__restrict__ s,
                             float *__restrict__ y) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t words = (size_t)nb * N / 16;
    if (idx < words) {
        const size_t e = idx * 16;
        const int b = (int)(e / N), q = (int)(e % N);
        *(uint4 *)(s + e) = *(const uint4 *)(data + (size_t)order[b] * N + q);
    }
    if (idx < (size_t)nb) y[idx] = labels[order[idx]];
}

}  // namespace cnn

using namespace cnn;

// =================================================================== host side
struct gasp_cnn {
    int L = 64, N = 4096, nmax = 0;
    hipStream_t st = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *w1r = nullptr;
    uint16_t *wf = nullptr, *wb = nullptr;
    int8_t *s = nullptr;
    float *y = nullptr, *pb = nullptr, *loss = nullptr, *dg = nullptr;
    uint16_t *a[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // a[1..4]
    uint16_t *dz0 = nullptr, *dz1 = nullptr;
    float *part = nullptr, *part1 = nullptr, *part_head = nullptr;
    float *red = nullptr;  // chunk sums of the two-level reductions
    int8_t *data = nullptr;
    float *data_y = nullptr;
    int *order = nullptr;
    int ndata = 0;
    long long step = 0;
    bool weights_ready = false;
};

#define CTRY(x)                                                                                   \
    do {                                                                                          \
        hipError_t q_ = (x);                                                                      \
        if (q_ != hipSuccess)                                                                     \
            return set_error(GASP_ERR_HIP, "GPUassert: %s %s %d", hipGetErrorString(q_), __FILE__, __LINE__); \
    } while (0)

static void cfree(void *p) {
    if (p) (void)hipFree(p);
}

extern "C" int gasp_cnn_num_params(void) { return NPARAM; }

extern "C" int gasp_cnn_destroy(gasp_cnn *c) {
    if (!c) return GASP_OK;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (void *q : {(void *)c->p, (void *)c->g, (void *)c->m, (void *)c->v, (void *)c->w1r, (void *)c->wf,
                    (void *)c->wb, (void *)c->s, (void *)c->y, (void *)c->pb, (void *)c->loss, (void *)c->dg,
                    (void *)c->a[1], (void *)c->a[2], (void *)c->a[3], (void *)c->a[4], (void *)c->dz0,
                    (void *)c->dz1, (void *)c->part, (void *)c->part1, (void *)c->part_head,
                    (void *)c->red, (void *)c->data, (void *)c->data_y, (void *)c->order})
        cfree(q);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return GASP_OK;
}

extern "C" int gasp_cnn_create(int L, int max_batch, gasp_cnn **out) {
    gasp::clear_error();
    if (!out) return set_error(GASP_ERR_ARG, "gasp_cnn_create: NULL output");
    *out = nullptr;
    if (L < 16 || (L & (L - 1)) != 0 || L > 1024) return set_error(GASP_ERR_ARG, "gasp_cnn: L must be a power of two in [16, 1024]");
    if (max_batch < 1) return set_error(GASP_ERR_ARG, "gasp_cnn: max_batch must be positive");
    gasp_cnn *c = new gasp_cnn();
    c->L = L;
    c->N = L * L;
    c->nmax = max_batch;
    const size_t n = (size_t)max_batch, N = (size_t)c->N;
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    auto A = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
        if (e == hipSuccess) e = hipMemsetAsync(*p, 0, bytes, c->st);
    };
    A((void **)&c->p, NPARAM * 4);
    A((void **)&c->g, NPARAM * 4);
    A((void **)&c->m, NPARAM * 4);
    A((void **)&c->v, NPARAM * 4);
    A((void **)&c->w1r, 288 * 4);
    A((void **)&c->wf, 3 * 9 * 2 * 64 * 8 * 2);
    A((void **)&c->wb, 3 * 9 * 2 * 64 * 8 * 2);
    A((void **)&c->s, n * N);
    A((void **)&c->y, n * 4);
    A((void **)&c->pb, n * 4);
    A((void **)&c->loss, n * 4);
    A((void **)&c->dg, n * C * 4);
    for (int k = 1; k <= 4; ++k) A((void **)&c->a[k], n * N * C * 2);
    A((void **)&c->dz0, n * N * C * 2);
    A((void **)&c->dz1, n * N * C * 2);
    A((void **)&c->part, n * 4 * kWgRow * 4);
    A((void **)&c->part1, n * (size_t)L * 320 * 4);
    A((void **)&c->part_head, n * HEAD_N * 4);
    {
        const size_t q = (size_t)kRedChunk;
        const size_t r = std::max({(n * 4 + q - 1) / q * kWgRow, (n * L + q - 1) / q * 320, (n + q - 1) / q * HEAD_N});
        A((void **)&c->red, r * 4);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) {
        gasp_cnn_destroy(c);
        return set_error(e == hipErrorOutOfMemory ? GASP_ERR_NOMEM : GASP_ERR_HIP, "gasp_cnn_create: %s",
                         hipGetErrorString(e));
    }
    *out = c;
    return GASP_OK;
}

// out[j] = sum_p part[p][j] for part [P][M], in a fixed order: chunks of
// kRedChunk partials in parallel, then the chunk sums.
static int reduce_to(gasp_cnn *c, const float *part, int P, int M, float *out) {
    const unsigned gx = (unsigned)((M + 255) / 256);
    if (P <= kRedChunk) {
        hipLaunchKernelGGL(reduce_partials, dim3(gx), dim3(256), 0, c->st, part, P, M, out);
    } else {
        // at least kRedChunk per chunk (the buffer holds ceil(P / kRedChunk) x M), and at most 65535 chunks
        const int chunk = std::max(kRedChunk, (P + 65534) / 65535);
        const int nc = (P + chunk - 1) / chunk;
        hipLaunchKernelGGL(reduce_chunks, dim3(gx, (unsigned)nc), dim3(256), 0, c->st, part, P, M, chunk, c->red);
        hipLaunchKernelGGL(reduce_partials, dim3(gx), dim3(256), 0, c->st, (const float *)c->red, nc, M, out);
    }
    CTRY(hipGetLastError());
    return GASP_OK;
}

static int prep(gasp_cnn *c) {
    hipLaunchKernelGGL(prep_weights, dim3((3 * 9 * 2 * 64 * 8 + 255) / 256), dim3(256), 0, c->st, c->p, c->wf, c->wb,
                       c->w1r);
    CTRY(hipGetLastError());
    c->weights_ready = true;
    return GASP_OK;
}

extern "C" int gasp_cnn_set_params(gasp_cnn *c, const float *host_params) {
    gasp::clear_error();
    if (!c || !host_params) return set_error(GASP_ERR_ARG, "gasp_cnn_set_params: NULL argument");
    CTRY(hipMemcpyAsync(c->p, host_params, NPARAM * 4, hipMemcpyHostToDevice, c->st));
    CTRY(hipMemsetAsync(c->m, 0, NPARAM * 4, c->st));
    CTRY(hipMemsetAsync(c->v, 0, NPARAM * 4, c->st));
    c->step = 0;
    int rc = prep(c);
    if (rc) return rc;
    CTRY(hipStreamSynchronize(c->st));
    return GASP_OK;
}

extern "C" int gasp_cnn_get_params(gasp_cnn *c, float *host_params) {
    if (!c || !host_params) return set_error(GASP_ERR_ARG, "gasp_cnn_get_params: NULL argument");
    CTRY(hipMemcpyAsync(host_params, c->p, NPARAM * 4, hipMemcpyDeviceToHost, c->st));
    CTRY(hipStreamSynchronize(c->st));
    return GASP_OK;
}

// forward of n <= nmax grids already in c->s
static int forward(gasp_cnn *c, int n, const float *d_labels) {
    const int L = c->L, N = c->N;
    const size_t tot = (size_t)n * N;
    hipLaunchKernelGGL(conv1_fwd, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->st, c->s, c->w1r, c->p,
                       c->a[1], n, L);
    const int rows = n * L;
    for (int k = 2; k <= 4; ++k)
        hipLaunchKernelGGL(conv_mfma<false>, dim3(rows / kConvRows), dim3(256), 0, c->st, c->a[k - 1],
                           c->wf + (size_t)(k - 2) * 9 * 2 * 64 * 8, c->p + OFF_B[k], (const uint16_t *)nullptr, c->a[k],
                           n, L);
    hipLaunchKernelGGL(head, dim3(n), dim3(256), 0, c->st, c->a[4], c->p, d_labels, c->pb, c->loss, c->part_head,
                       c->dg, n, L, 1.0f / (float)n);
    CTRY(hipGetLastError());
    return GASP_OK;
}

// backward after forward(c, n, labels): fills c->g
static int backward(gasp_cnn *c, int n) {
    const int L = c->L, N = c->N;
    const size_t tot = (size_t)n * N;
    const int rows = n * L;
    int rc = reduce_to(c, c->part_head, n, HEAD_N, c->g + OFF_FW1);
    if (rc) return rc;
    hipLaunchKernelGGL(head_bwd_fill, dim3((unsigned)((tot * 16 + 255) / 256)), dim3(256), 0, c->st, c->a[4], c->dg,
                       c->dz0, n, L);
    uint16_t *dz = c->dz0, *dzn = c->dz1;
    for (int k = 4; k >= 2; --k) {
        // one partial per workgroup; the buffers hold 4 nmax of them
        const int nbands = n * (L / 4) * (L / std::min(L, kWgXW));
        const int nblk = std::min({nbands, kWgBlocks, 4 * c->nmax});
        hipLaunchKernelGGL(conv_wgrad, dim3(nblk), dim3(256), 0, c->st, c->a[k - 1], dz, c->part, n, L);
        // weights and bias are adjacent in the parameter vector
        if ((rc = reduce_to(c, c->part, nblk, kWgRow, c->g + OFF_W[k])) != 0) return rc;
        // d a_{k-1}, masked by a_{k-1} > 0
        hipLaunchKernelGGL(conv_mfma<true>, dim3(rows / kConvRows), dim3(256), 0, c->st, dz,
                           c->wb + (size_t)(k - 2) * 9 * 2 * 64 * 8, (const float *)nullptr, c->a[k - 1], dzn, n, L);
        std::swap(dz, dzn);
    }
    {
        const int nbands = n * (L / 4) * (L / std::min(L, kWgXW));
        const int nblk = std::min({nbands, kW1Blocks, rows});  // part1 holds n L partials
        hipLaunchKernelGGL(conv1_wgrad, dim3(nblk), dim3(256), 0, c->st, c->s, dz, c->part1, n, L);
        if ((rc = reduce_to(c, c->part1, nblk, 320, c->g)) != 0) return rc;
    }
    CTRY(hipGetLastError());
    return GASP_OK;
}

static int load_batch(gasp_cnn *c, const int8_t *spins, const float *labels, int n) {
    if (n < 1 || n > c->nmax) return set_error(GASP_ERR_ARG, "gasp_cnn: batch %d outside [1, %d]", n, c->nmax);
    CTRY(hipMemcpyAsync(c->s, spins, (size_t)n * c->N, hipMemcpyHostToDevice, c->st));
    if (labels) CTRY(hipMemcpyAsync(c->y, labels, (size_t)n * 4, hipMemcpyHostToDevice, c->st));
    return GASP_OK;
}

extern "C" int gasp_cnn_predict(gasp_cnn *c, const int8_t *spins, int n, float *pB) {
    gasp::clear_error();
    if (!c || !spins || !pB || n < 0) return set_error(GASP_ERR_ARG, "gasp_cnn_predict: bad argument");
    if (!c->weights_ready) return set_error(GASP_ERR_ARG, "gasp_cnn_predict: parameters not set");
    for (int i0 = 0; i0 < n; i0 += c->nmax) {
        const int nb = std::min(c->nmax, n - i0);
        int rc = load_batch(c, spins + (size_t)i0 * c->N, nullptr, nb);
        if (!rc) rc = forward(c, nb, nullptr);
        if (rc) return rc;
        CTRY(hipMemcpyAsync(pB + i0, c->pb, (size_t)nb * 4, hipMemcpyDeviceToHost, c->st));
    }
    CTRY(hipStreamSynchronize(c->st));
    return GASP_OK;
}

extern "C" int gasp_cnn_gradient(gasp_cnn *c, const int8_t *spins, const float *labels, int n, float *grad,
                                 float *loss) {
    gasp::clear_error();
    if (!c || !spins || !labels || !grad) return set_error(GASP_ERR_ARG, "gasp_cnn_gradient: bad argument");
    if (!c->weights_ready) return set_error(GASP_ERR_ARG, "gasp_cnn_gradient: parameters not set");
    int rc = load_batch(c, spins, labels, n);
    if (!rc) rc = forward(c, n, c->y);
    if (!rc) rc = backward(c, n);
    if (rc) return rc;
    CTRY(hipMemcpyAsync(grad, c->g, NPARAM * 4, hipMemcpyDeviceToHost, c->st));
    std::vector<float> l(n);
    CTRY(hipMemcpyAsync(l.data(), c->loss, (size_t)n * 4, hipMemcpyDeviceToHost, c->st));
    CTRY(hipStreamSynchronize(c->st));
    if (loss) {
        double s = 0;
        for (float x : l) s += x;
        *loss = (float)(s / n);
    }
    return GASP_OK;
}

static int adam_step(gasp_cnn *c, float lr) {
    c->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)c->step), c2 = 1.0f - powf(b2, (float)c->step);
    hipLaunchKernelGGL(adam, dim3((NPARAM + 255) / 256), dim3(256), 0, c->st, c->p, c->g, c->m, c->v, NPARAM, lr, b1,
                       b2, 1e-8f, c1, c2);
    CTRY(hipGetLastError());
    return prep(c);
}

extern "C" int gasp_cnn_train_step(gasp_cnn *c, const int8_t *spins, const float *labels, int n, float lr,
                                   float *loss) {
    float dummy;
    gasp::clear_error();
    if (!c || !spins || !labels) return set_error(GASP_ERR_ARG, "gasp_cnn_train_step: bad argument");
    if (!c->weights_ready) return set_error(GASP_ERR_ARG, "gasp_cnn_train_step: parameters not set");
    int rc = load_batch(c, spins, labels, n);
    if (!rc) rc = forward(c, n, c->y);
    if (!rc) rc = backward(c, n);
    if (!rc) rc = adam_step(c, lr);
    if (rc) return rc;
    std::vector<float> l(n);
    CTRY(hipMemcpyAsync(l.data(), c->loss, (size_t)n * 4, hipMemcpyDeviceToHost, c->st));
    CTRY(hipStreamSynchronize(c->st));
    double s = 0;
    for (float x : l) s += x;
    *(loss ? loss : &dummy) = (float)(s / n);
    return GASP_OK;
}

extern "C" int gasp_cnn_set_data(gasp_cnn *c, const int8_t *spins, const float *labels, int count) {
    gasp::clear_error();
    if (!c || !spins || !labels || count < 1) return set_error(GASP_ERR_ARG, "gasp_cnn_set_data: bad argument");
    cfree(c->data);
    cfree(c->data_y);
    cfree(c->order);
    c->data = nullptr;
    c->data_y = nullptr;
    c->order = nullptr;
    CTRY(hipMalloc((void **)&c->data, (size_t)count * c->N));
    CTRY(hipMalloc((void **)&c->data_y, (size_t)count * 4));
    CTRY(hipMalloc((void **)&c->order, (size_t)count * 4));
    CTRY(hipMemcpyAsync(c->data, spins, (size_t)count * c->N, hipMemcpyHostToDevice, c->st));
    CTRY(hipMemcpyAsync(c->data_y, labels, (size_t)count * 4, hipMemcpyHostToDevice, c->st));
    CTRY(hipStreamSynchronize(c->st));
    c->ndata = count;
    return GASP_OK;
}

// One pass over the stored data in the given order (indices into the data),
// in batches of `batch` (the last one may be smaller).  Mean loss returned.
extern "C" int gasp_cnn_train_epoch(gasp_cnn *c, const int *order, int count, int batch, float lr, float *mean_loss) {
    gasp::clear_error();
    if (!c || !order || count < 1 || count > c->ndata || batch < 1 || batch > c->nmax)
        return set_error(GASP_ERR_ARG, "gasp_cnn_train_epoch: bad argument");
    if (!c->weights_ready) return set_error(GASP_ERR_ARG, "gasp_cnn_train_epoch: parameters not set");
    for (int i = 0; i < count; ++i)
        if (order[i] < 0 || order[i] >= c->ndata) return set_error(GASP_ERR_ARG, "gasp_cnn_train_epoch: index out of range");
    CTRY(hipMemcpyAsync(c->order, order, (size_t)count * 4, hipMemcpyHostToDevice, c->st));
    std::vector<float> losses;
    std::vector<float> l(batch);
    double tot = 0.0;
    for (int i0 = 0; i0 < count; i0 += batch) {
        const int nb = std::min(batch, count - i0);
        const size_t words = std::max<size_t>((size_t)nb * c->N / 16, (size_t)nb);
        hipLaunchKernelGGL(gather_batch, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, c->st, c->data,
                           c->data_y, c->order + i0, nb, c->N, c->s, c->y);
        int rc = forward(c, nb, c->y);
        if (!rc) rc = backward(c, nb);
        if (!rc) rc = adam_step(c, lr);
        if (rc) return rc;
        CTRY(hipMemcpyAsync(l.data(), c->loss, (size_t)nb * 4, hipMemcpyDeviceToHost, c->st));
        CTRY(hipStreamSynchronize(c->st));
        for (int k = 0; k < nb; ++k) tot += l[k];
    }
    if (mean_loss) *mean_loss = (float)(tot / count);
    return GASP_OK;
}
