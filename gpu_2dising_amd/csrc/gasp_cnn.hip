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
//     32-position chunks staged transposed in LDS.
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
// grid (n, 9 taps), 4 waves per block, each wave a quarter of the image's
// positions.  Writes the wave's partial dW[t][ci][co] (and, for the centre
// tap, the partial bias gradient) to part[(img * 4 + wave)][...].
__global__ void __launch_bounds__(256) conv_wgrad(const uint16_t *__restrict__ ain, const uint16_t *__restrict__ dz,
                                                  float *__restrict__ part, float *__restrict__ part_b, int n, int L) {
    __shared__ uint16_t sz[4][C][40];  // [wave][channel][32 positions + pad]
    __shared__ uint16_t sa[4][C][40];
    const int img = blockIdx.x, t = blockIdx.y;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int N = L * L, m = L - 1, dy = t / 3 - 1, dx = t % 3 - 1;
    const int chunk = N / 4;  // positions per wave (multiple of 32 for L >= 16)
    const uint16_t *zb = dz + (size_t)img * N * C;
    const uint16_t *ab = ain + (size_t)img * N * C;
    f32x4 acc[2][2] = {{{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}, {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}};
    float bsum[2] = {0.f, 0.f};
    const int pl = lane & 15, kq = lane >> 4;
    for (int p0 = w * chunk; p0 < (w + 1) * chunk; p0 += 32) {
        // stage 32 positions x 32 channels of dz and of the shifted input, transposed
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int item = lane + 64 * r;  // 128 items: position (item >> 2), channel block (item & 3)
            const int k = item >> 2, cb = (item & 3) * 8;
            const int pos = p0 + k;
            const int yy = ((pos / L) + dy) & m, xx = ((pos % L) + dx) & m;
            const uint4 vz = *(const uint4 *)(zb + (size_t)pos * C + cb);
            const uint4 va = *(const uint4 *)(ab + (size_t)(yy * L + xx) * C + cb);
            const uint32_t z4[4] = {vz.x, vz.y, vz.z, vz.w}, a4[4] = {va.x, va.y, va.z, va.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                sz[w][cb + 2 * q][k] = (uint16_t)z4[q];
                sz[w][cb + 2 * q + 1][k] = (uint16_t)(z4[q] >> 16);
                sa[w][cb + 2 * q][k] = (uint16_t)a4[q];
                sa[w][cb + 2 * q + 1][k] = (uint16_t)(a4[q] >> 16);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint4 A[2], B[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            A[h] = *(const uint4 *)&sz[w][16 * h + pl][8 * kq];
            B[h] = *(const uint4 *)&sa[w][16 * h + pl][8 * kq];
        }
        if (t == 4) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t v4[4] = {A[h].x, A[h].y, A[h].z, A[h].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) bsum[h] += bf2f((uint16_t)v4[q]) + bf2f((uint16_t)(v4[q] >> 16));
            }
        }
#pragma unroll
        for (int mh = 0; mh < 2; ++mh)
#pragma unroll
            for (int nh = 0; nh < 2; ++nh) acc[mh][nh] = mfma(A[mh], B[nh], acc[mh][nh]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    float *out = part + ((size_t)(img * 4 + w) * 9 + t) * C * C;
#pragma unroll
    for (int mh = 0; mh < 2; ++mh)
#pragma unroll
        for (int nh = 0; nh < 2; ++nh)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = 16 * mh + 4 * kq + j, ci = 16 * nh + pl;
                out[ci * C + co] = acc[mh][nh][j];
            }
    if (t == 4) {
        // lanes l, l+16, l+32, l+48 hold partial sums of channel 16 h + l % 16
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float v = bsum[h];
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (kq == 0) part_b[(size_t)(img * 4 + w) * C + 16 * h + pl] = v;
        }
    }
}

// ------------------------------------------------------------------ conv1 weight gradient
// One block per (image, row), wave = channel octet, lane = column.  Partials
// [w1 (9 x 32) | b1 (32)] per row to part1[(img * L + y)][320].
__global__ void __launch_bounds__(256) conv1_wgrad(const int8_t *__restrict__ s, const uint16_t *__restrict__ dz,
                                                   float *__restrict__ part1, int n, int L) {
    const int row = blockIdx.x;
    const int img = row / L, y = row % L, m = L - 1, N = L * L;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float acc[9][8], bacc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        bacc[c] = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[t][c] = 0.f;
    }
    for (int x = lane; x < L; x += 64) {
        float sv[9];
#pragma unroll
        for (int t = 0; t < 9; ++t)
            sv[t] = (float)s[(size_t)img * N + ((y + t / 3 - 1) & m) * L + ((x + t % 3 - 1) & m)];
        const uint4 v = *(const uint4 *)(dz + ((size_t)img * N + y * L + x) * C + 8 * w);
        const uint32_t d4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float d0 = bf2f((uint16_t)d4[q]), d1 = bf2f((uint16_t)(d4[q] >> 16));
            bacc[2 * q] += d0;
            bacc[2 * q + 1] += d1;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                acc[t][2 * q] += sv[t] * d0;
                acc[t][2 * q + 1] += sv[t] * d1;
            }
        }
    }
    float *out = part1 + (size_t)row * 320;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            float v = acc[t][c];
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) out[t * C + 8 * w + c] = v;
        }
        float b = bacc[c];
        for (int off = 32; off >= 1; off >>= 1) b += __shfl_xor(b, off, 64);
        if (lane == 0) out[288 + 8 * w + c] = b;
    }
}

// ------------------------------------------------------------------ head
// One block (256 threads) per image: global average pool, fc1 + ReLU, fc2,
// sigmoid; with labels also the loss, the head's parameter gradient
// (partial per image) and dg = d loss / d (pooled features).
__global__ void __launch_bounds__(256) head(const uint16_t *__restrict__ a4, const float *__restrict__ p,
                                            const float *__restrict__ labels, float *__restrict__ pb,
                                            float *__restrict__ loss, float *__restrict__ part_head,
                                            float *__restrict__ dg, int n, int L, float inv_batch) {
    __shared__ float red[8][C];
    __shared__ float g[C], hpre[HID], h[HID], dh[HID], zsh;
    const int img = blockIdx.x, tid = threadIdx.x, N = L * L;
    const int c = tid & 31, grp = tid >> 5;
    float sum = 0.f;
    for (int q = grp; q < N; q += 8) sum += bf2f(a4[((size_t)img * N + q) * C + c]);
    red[grp][c] = sum;
    __syncthreads();
    if (tid < C) {
        float s = 0.f;
        for (int k = 0; k < 8; ++k) s += red[k][tid];
        g[tid] = s / (float)N;
    }
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
    float *part = nullptr, *part_b = nullptr, *part1 = nullptr, *part_head = nullptr;
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
                    (void *)c->dz1, (void *)c->part, (void *)c->part_b, (void *)c->part1, (void *)c->part_head,
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
    A((void **)&c->part, n * 4 * 9 * C * C * 4);
    A((void **)&c->part_b, n * 4 * C * 4);
    A((void **)&c->part1, n * (size_t)L * 320 * 4);
    A((void **)&c->part_head, n * HEAD_N * 4);
    {
        const size_t q = (size_t)kRedChunk;
        const size_t r = std::max({(n * 4 + q - 1) / q * 9 * C * C, (n * L + q - 1) / q * 320, (n + q - 1) / q * HEAD_N});
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
        hipLaunchKernelGGL(conv_wgrad, dim3(n, 9), dim3(256), 0, c->st, c->a[k - 1], dz, c->part, c->part_b, n, L);
        if ((rc = reduce_to(c, c->part, n * 4, 9 * C * C, c->g + OFF_W[k])) != 0) return rc;
        if ((rc = reduce_to(c, c->part_b, n * 4, C, c->g + OFF_B[k])) != 0) return rc;
        // d a_{k-1}, masked by a_{k-1} > 0
        hipLaunchKernelGGL(conv_mfma<true>, dim3(rows / kConvRows), dim3(256), 0, c->st, dz,
                           c->wb + (size_t)(k - 2) * 9 * 2 * 64 * 8, (const float *)nullptr, c->a[k - 1], dzn, n, L);
        std::swap(dz, dzn);
    }
    hipLaunchKernelGGL(conv1_wgrad, dim3(rows), dim3(256), 0, c->st, c->s, dz, c->part1, n, L);
    if ((rc = reduce_to(c, c->part1, rows, 320, c->g)) != 0) return rc;
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
