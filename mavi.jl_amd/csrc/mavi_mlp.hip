// mavi_mlp.hip -- the learned propagator's network: an MLP 160 -> 128 -> 128
// -> 128 -> 2 with GELU (sigmoid form), BF16 weights and activations, FP32 accumulation on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
//   inference  k_mlp_fused: one persistent kernel, all weights resident in LDS
//              (117 KB), 8 wavefronts per CU (two per SIMD, so one wave's
//              activation epilogue overlaps another's MFMAs), each running
//              16-row tiles through the four layers; activations go through
//              a per-wave LDS tile between layers
//   training   k_gemm<TA, TB, EPI>: LDS-staged 64x64x32 tiles with optional
//              transposed operands and fused epilogues (bias + GELU keeping
//              the pre-activation, GELU backward, split-K partials), plus
//              loss/dY, two-stage column sums (bias gradients), split-K reduction and
//              Adam kernels
//
// Parameters (FP32 master copy, flat): for layer l = 0..3, W_l^T (out x in,
// row-major) followed by b_l.  BF16 copies of W_l^T feed the matrix cores.
#include <math.h>

#include <mutex>
#include <string>
#include <vector>

#include "mavi_host.h"
#include "mavi_mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// Two values to packed BF16 (a in the low half), one v_cvt_pk_bf16_f32.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{a, b}, bf16x2_t));
}

// GELU in its sigmoid form, g(z) = z sigma(1.702 z) (Hendrycks & Gimpel):
// one hardware exp2 (of z times -1.702 log2 e folded into one constant) and
// one reciprocal per activation (a few ulp in FP32, far below the BF16
// rounding that follows).  g'(z) = s + 1.702 z s (1 - s).
constexpr float kGs = 1.702f;
constexpr float kGsLog2e = -1.702f * 1.44269504088896341f;
__device__ __forceinline__ float sig(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * kGsLog2e));
}
__device__ __forceinline__ float gelu(float z) { return z * sig(z); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float s = sig(z);
    return s + kGs * z * s * (1.0f - s);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Generic GEMM: C[m][n] = sum_k A(m,k) B(k,n),
//   A(m,k) = TA ? A[k*lda + m] : A[m*lda + k],  B(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n]
// BF16 operands, FP32 accumulation.  256 threads, 64x64 output tile, K steps
// of 32 through LDS (rows k-contiguous, padded to 40 elements: the 16-lane
// fragment reads fall on distinct banks).  Wave w owns the 32x32 quarter
// (w/2, w%2) as 2x2 MFMA blocks.  blockIdx.z is the split-K index.
// ---------------------------------------------------------------------------
enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_BWD = 3 };

struct Epi {
    const float *bias;  // EPI_BIAS_GELU, EPI_BIAS
    float *Z;           // pre-activation: written (BIAS_GELU) / read (GELU_BWD)
    int ldz;
    uint16_t *H;        // BF16 output (BIAS_GELU: GELU(z); GELU_BWD: dZ)
    int ldh;
};

constexpr int kGT = 64, kGK = 32, kGP = 40;

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                                              const uint16_t *__restrict__ B, int ldb, float *__restrict__ C, int ldc,
                                              int kchunk, int64_t split_stride, Epi ep) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kGT * kGP];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kGT * kGP];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kbeg; k0 < kend; k0 += kGK) {
        // A tile -> As[m][k]
        if constexpr (!TA) {
            const int r = t >> 2, kk = (t & 3) * 8;
            const int gm = m0 + r, gk = k0 + kk;
            uint16_t v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = (gm < M && gk + j < kend) ? A[(int64_t)gm * lda + gk + j] : 0;
#pragma unroll
            for (int j = 0; j < 8; j++) As[r * kGP + kk + j] = v[j];
        } else {
            const int kk = t >> 3, mm = (t & 7) * 8;
            const int gk = k0 + kk, gm = m0 + mm;
#pragma unroll
            for (int j = 0; j < 8; j++)
                As[(mm + j) * kGP + kk] = (gk < kend && gm + j < M) ? A[(int64_t)gk * lda + gm + j] : 0;
        }
        // B tile -> Bs[n][k]
        if constexpr (TB) {
            const int r = t >> 2, kk = (t & 3) * 8;
            const int gn = n0 + r, gk = k0 + kk;
#pragma unroll
            for (int j = 0; j < 8; j++)
                Bs[r * kGP + kk + j] = (gn < N && gk + j < kend) ? B[(int64_t)gn * ldb + gk + j] : 0;
        } else {
            const int kk = t >> 3, nn = (t & 7) * 8;
            const int gk = k0 + kk, gn = n0 + nn;
#pragma unroll
            for (int j = 0; j < 8; j++)
                Bs[(nn + j) * kGP + kk] = (gk < kend && gn + j < N) ? B[(int64_t)gk * ldb + gn + j] : 0;
        }
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = ld8(&As[(wm + 16 * i + fr) * kGP + fk]);
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = ld8(&Bs[(wn + 16 * j + fr) * kGP + fk]);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
        __syncthreads();
    }
    // D layout: lane holds column fr of the 16x16 block, rows 4*(lane/16)+r
    C += (int64_t)blockIdx.z * split_stride;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int n = n0 + wn + 16 * j + fr;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + (lane >> 4) * 4 + r;
                if (m >= M || n >= N) continue;
                const float x = acc[i][j][r];
                if constexpr (EPI == EPI_F32) {
                    C[(int64_t)m * ldc + n] = x;
                } else if constexpr (EPI == EPI_BIAS) {
                    C[(int64_t)m * ldc + n] = x + ep.bias[n];
                } else if constexpr (EPI == EPI_BIAS_GELU) {
                    const float z = x + ep.bias[n];
                    ep.Z[(int64_t)m * ep.ldz + n] = z;
                    ep.H[(int64_t)m * ep.ldh + n] = f2bf(gelu(z));
                } else {
                    const float z = ep.Z[(int64_t)m * ep.ldz + n];
                    ep.H[(int64_t)m * ep.ldh + n] = f2bf(x * gelu_grad(z));
                }
            }
        }
}

// out[i] = scale * sum_z ws[z * stride + i] for i < count (fixed order over z).
__global__ void k_split_reduce(const float *__restrict__ ws, int nsplit, int64_t stride, int64_t count,
                               float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int z = 0; z < nsplit; z++) s += ws[z * stride + i];
    out[i] = s;
}

// Column sums of H (BF16 [M][N], FP32 sums in a fixed order), first stage:
// block (x, y) sums rows [256 y, 256 y + 256) of columns [64 x, 64 x + 64)
// into part[y][n]; k_split_reduce then adds the row chunks in order.
constexpr int kColRows = 256;
__global__ void k_colsum_part(const uint16_t *__restrict__ H, int M, int N, int ldh, float *__restrict__ part) {
    __shared__ float red[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    const int m0 = blockIdx.y * kColRows;
    float s = 0.f;
    if (n < N)
        for (int m = m0 + g; m < min(M, m0 + kColRows); m += 4) s += bf2f(H[(int64_t)m * ldh + n]);
    red[g][c] = s;
    __syncthreads();
    if (g == 0 && n < N) part[(int64_t)blockIdx.y * N + n] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// Mean-squared error L = 1/(2M) sum_m sum_o (Y - T)^2 over the 2 outputs;
// dY = (Y - T) / M as BF16 into the 16-wide padded gradient (columns 2..15
// stay zero).  Block partial sums of the squared error go to part[].
__global__ void k_loss(const float *__restrict__ Y, const float *__restrict__ T, int M, uint16_t *__restrict__ dY,
                       float *__restrict__ part) {
    __shared__ float red[256];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.f;
    if (m < M) {
        const float inv = 1.0f / (float)M;
        for (int o = 0; o < kMlpOut; o++) {
            const float d = Y[(int64_t)m * kMlpOutPad + o] - T[(int64_t)m * kMlpOut + o];
            s += d * d;
            dY[(int64_t)m * kMlpOutPad + o] = f2bf(d * inv);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam on the FP32 master parameters (step t counted from 1).
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// FP32 master -> BF16 weights (W0t..W3t, W3t padded to 16 rows) and the
// padded output bias.
__global__ void k_sync_weights(const float *__restrict__ p, uint16_t *__restrict__ w0, uint16_t *__restrict__ w1,
                               uint16_t *__restrict__ w2, uint16_t *__restrict__ w3, float *__restrict__ b3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpHid * kMlpIn) w0[i] = f2bf(p[kOffW0 + i]);
    if (i < kMlpHid * kMlpHid) {
        w1[i] = f2bf(p[kOffW1 + i]);
        w2[i] = f2bf(p[kOffW2 + i]);
    }
    if (i < kMlpOutPad * kMlpHid) w3[i] = (i < kMlpOut * kMlpHid) ? f2bf(p[kOffW3 + i]) : (uint16_t)0;
    if (i < kMlpOutPad) b3[i] = (i < kMlpOut) ? p[kOffB3 + i] : 0.f;
}

// ---------------------------------------------------------------------------
// Fused inference.  LDS holds only the weights: W0t [128][168], W1t/W2t
// [128][136], W3t [16][136] (rows k-contiguous, +8 padding) and the biases;
// activations stay in registers.  Each wave computes 16-row tiles, every
// layer transposed: out^T[n][m] = sum_k W^T[n][k] act[m][k], weights as the
// first MFMA operand (16 output features per block f), the tile's rows as
// the second.  A lane (row m = lane % 16, g = lane / 16) then holds
// features 16 f + 4 g + i (i < 4) of its row: after bias + GELU, blocks 2 ks
// and 2 ks + 1 give exactly the 8 BF16 values the lane supplies as the next
// layer's second operand at K step ks, once that layer's weight columns are
// stored in that order (column 32 ks + 8 g + j holds hidden unit
// 32 ks + 4 g + j for j < 4, 32 ks + 16 + 4 g + j - 4 for j >= 4; a
// relabelling of the hidden units, the same network).  Rows of X: 160 BF16
// features; Y: 2 FP32 outputs per row.
// ---------------------------------------------------------------------------
constexpr int kP0 = kMlpIn + 8, kPH = kMlpHid + 8;
constexpr int kLdsW0 = kMlpHid * kP0, kLdsW1 = kMlpHid * kPH, kLdsW3 = kMlpOutPad * kPH;
constexpr int kFusedWaves = 12, kTileRows = 16;
constexpr size_t kFusedLds = sizeof(uint16_t) * (size_t)(kLdsW0 + 2 * kLdsW1 + kLdsW3) +
                             sizeof(float) * (size_t)(3 * kMlpHid + kMlpOutPad);

// One layer's MFMAs for NT 16-row tiles: acc[t][f] = W^T[16 f .. 16 f + 15][.]
// x act_t^T, the inputs h[t][ks] (the lane's 8 values at K step ks), each
// weight fragment read once from LDS for the NT tiles, one K step ahead.
template <int KIN, int PIN, int NT>
__device__ __forceinline__ void fused_layer(const bf16x8 (*h)[KIN / 32], const uint16_t *W, int lane,
                                            f32x4 (&acc)[NT][8]) {
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    constexpr int KS = KIN / 32;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int f = 0; f < 8; f++) acc[t][f] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 wf[2][8];
#pragma unroll
    for (int f = 0; f < 8; f++) wf[0][f] = ld8(W + (16 * f + fr) * PIN + fk);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const int cb = ks & 1;
        if (ks + 1 < KS) {
#pragma unroll
            for (int f = 0; f < 8; f++) wf[cb ^ 1][f] = ld8(W + (16 * f + fr) * PIN + (ks + 1) * 32 + fk);
        }
#pragma unroll
        for (int f = 0; f < 8; f++)
#pragma unroll
            for (int t = 0; t < NT; t++) acc[t][f] = mfma16(wf[cb][f], h[t][ks], acc[t][f]);
        if (ks + 1 < KS) __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 8 * NT, 0);
    }
}

// Bias + GELU of a hidden layer's accumulators into the next layer's inputs
// (the column order above).
__device__ __forceinline__ void fused_act(const f32x4 (&acc)[8], const float *bias, int lane,
                                          bf16x8 (&h)[kMlpHid / 32]) {
    const int n4 = (lane >> 4) * 4;
#pragma unroll
    for (int ks = 0; ks < kMlpHid / 32; ks++) {
        const float4 b0 = *reinterpret_cast<const float4 *>(bias + 32 * ks + n4);
        const float4 b1 = *reinterpret_cast<const float4 *>(bias + 32 * ks + 16 + n4);
        const f32x4 a0 = acc[2 * ks], a1 = acc[2 * ks + 1];
        const uint4 v = make_uint4(f2bf2(gelu(a0[0] + b0.x), gelu(a0[1] + b0.y)),
                                   f2bf2(gelu(a0[2] + b0.z), gelu(a0[3] + b0.w)),
                                   f2bf2(gelu(a1[0] + b1.x), gelu(a1[1] + b1.y)),
                                   f2bf2(gelu(a1[2] + b1.z), gelu(a1[3] + b1.w)));
        h[ks] = __builtin_bit_cast(bf16x8, v);
    }
}

// Copy a [ROWS][COLS] BF16 matrix into LDS rows of PITCH elements, 16 bytes
// per thread and 8 loads in flight (the chunk index is clamped, so every load
// is unconditional; the clamped copies store the same value twice).  PERM:
// the columns in the hidden-unit order above (each 16-byte chunk from two
// 8-byte pieces).
template <int ROWS, int COLS, int PITCH, bool PERM>
__device__ __forceinline__ void stage_rows(uint16_t *dst, const uint16_t *__restrict__ src, int t) {
    constexpr int kQ = ROWS * COLS / 8, kStride = 64 * kFusedWaves, kU = 8;
    for (int q0 = t; q0 < kQ; q0 += kU * kStride) {
        uint4 v[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int q = min(q0 + u * kStride, kQ - 1);
            const int row = q / (COLS / 8), c8 = (q % (COLS / 8)) * 8;
            if constexpr (PERM) {
                const int ks = c8 / 32, g = (c8 % 32) / 8;
                const uint2 lo = *reinterpret_cast<const uint2 *>(src + row * COLS + 32 * ks + 4 * g);
                const uint2 hi = *reinterpret_cast<const uint2 *>(src + row * COLS + 32 * ks + 16 + 4 * g);
                v[u] = make_uint4(lo.x, lo.y, hi.x, hi.y);
            } else {
                v[u] = *reinterpret_cast<const uint4 *>(src + row * COLS + c8);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int q = min(q0 + u * kStride, kQ - 1);
            *reinterpret_cast<uint4 *>(dst + (q / (COLS / 8)) * PITCH + (q % (COLS / 8)) * 8) = v[u];
        }
    }
}

__global__ __launch_bounds__(64 * kFusedWaves) void k_mlp_fused(
    const uint16_t *__restrict__ X, int M, const uint16_t *__restrict__ w0, const uint16_t *__restrict__ w1,
    const uint16_t *__restrict__ w2, const uint16_t *__restrict__ w3, const float *__restrict__ params,
    const float *__restrict__ b3, float *__restrict__ Y) {
    extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
    uint16_t *W0 = lds;
    uint16_t *W1 = W0 + kLdsW0;
    uint16_t *W2 = W1 + kLdsW1;
    uint16_t *W3 = W2 + kLdsW1;
    float *B = reinterpret_cast<float *>(W3 + kLdsW3);  // b0 | b1 | b2 | b3 (16)
    const int t = threadIdx.x;
    // stage the weights (16-byte chunks, 8 loads in flight per thread) and biases
    stage_rows<kMlpHid, kMlpIn, kP0, false>(W0, w0, t);
    stage_rows<kMlpHid, kMlpHid, kPH, true>(W1, w1, t);
    stage_rows<kMlpHid, kMlpHid, kPH, true>(W2, w2, t);
    stage_rows<kMlpOutPad, kMlpHid, kPH, true>(W3, w3, t);
    for (int q = t; q < kMlpHid; q += blockDim.x) {
        B[q] = params[kOffB0 + q];
        B[kMlpHid + q] = params[kOffB1 + q];
        B[2 * kMlpHid + q] = params[kOffB2 + q];
    }
    if (t < kMlpOutPad) B[3 * kMlpHid + t] = b3[t];
    __syncthreads();

    const int lane = t & 63, w = t >> 6;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    constexpr int NT = kTileRows / 16;
    const int ntiles = (M + kTileRows - 1) / kTileRows;
    for (int tile = blockIdx.x * kFusedWaves + w; tile < ntiles; tile += gridDim.x * kFusedWaves) {
        const int64_t row0 = (int64_t)tile * kTileRows;
        bf16x8 x[NT][kMlpIn / 32];
#pragma unroll
        for (int u = 0; u < NT; u++) {
            const int64_t r = row0 + 16 * u + fr;
#pragma unroll
            for (int ks = 0; ks < kMlpIn / 32; ks++)
                x[u][ks] =
                    r < M ? ld8(X + r * kMlpIn + ks * 32 + fk) : __builtin_bit_cast(bf16x8, uint4{0u, 0u, 0u, 0u});
        }
        f32x4 acc[NT][8];
        bf16x8 h[NT][kMlpHid / 32];
        fused_layer<kMlpIn, kP0, NT>(x, W0, lane, acc);
#pragma unroll
        for (int u = 0; u < NT; u++) fused_act(acc[u], B, lane, h[u]);
        fused_layer<kMlpHid, kPH, NT>(h, W1, lane, acc);
#pragma unroll
        for (int u = 0; u < NT; u++) fused_act(acc[u], B + kMlpHid, lane, h[u]);
        fused_layer<kMlpHid, kPH, NT>(h, W2, lane, acc);
#pragma unroll
        for (int u = 0; u < NT; u++) fused_act(acc[u], B + 2 * kMlpHid, lane, h[u]);
        // output layer, transposed: 16 padded outputs (2 used) x 16 rows per tile
#pragma unroll
        for (int u = 0; u < NT; u++) {
            f32x4 o = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < kMlpHid / 32; ks++) o = mfma16(ld8(W3 + fr * kPH + ks * 32 + fk), h[u][ks], o);
            const int64_t r = row0 + 16 * u + fr;
            if ((lane >> 4) == 0 && r < M)  // outputs 0 and 1 of row r
                *reinterpret_cast<float2 *>(Y + r * kMlpOut) =
                    make_float2(o[0] + B[3 * kMlpHid], o[1] + B[3 * kMlpHid + 1]);
        }
    }
}

// Minibatch rows drawn uniformly (with replacement) from the data set by a
// splitmix64 hash of (seed, step, row): one thread per 16-byte chunk.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_gather_batch(const uint16_t *__restrict__ dsX, const float *__restrict__ dsT, int64_t count,
                               uint64_t seed, int64_t step, int batch, uint16_t *__restrict__ bX,
                               float *__restrict__ bT) {
    constexpr int kChunks = kMlpIn / 8;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)batch * kChunks) return;
    const int j = (int)(q / kChunks), c = (int)(q % kChunks);
    const uint64_t h = mix64(seed ^ mix64(((uint64_t)step << 32) ^ (uint64_t)j));
    const int64_t src = (int64_t)(h % (uint64_t)count);
    *reinterpret_cast<uint4 *>(bX + (int64_t)j * kMlpIn + c * 8) =
        *reinterpret_cast<const uint4 *>(dsX + src * kMlpIn + c * 8);
    if (c == 0)
        for (int o = 0; o < kMlpOut; o++) bT[(int64_t)j * kMlpOut + o] = dsT[src * kMlpOut + o];
}

// loss[0] = sum of the block partials / (2M), in a fixed order.
__global__ void k_loss_finish(const float *__restrict__ part, int nb, int M, float *__restrict__ loss) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.f;
    for (int i = 0; i < nb; i++) s += part[i];
    loss[0] = s / (2.0f * (float)M);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace mavi {

template <bool TA, bool TB, int EPI>
hipError_t gemm_t(hipStream_t st, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                         float *C, int ldc, int kchunk, int64_t split_stride, const Epi &ep) {
    if (kchunk <= 0) kchunk = K;
    kchunk = (kchunk + kGK - 1) / kGK * kGK;
    const int nsplit = (K + kchunk - 1) / kchunk;
    dim3 grid((N + kGT - 1) / kGT, (M + kGT - 1) / kGT, nsplit > 0 ? nsplit : 1);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, kchunk,
                       split_stride, ep);
    return hipGetLastError();
}

// The fused kernel's dynamic LDS (150 KB) is above the default limit: raised
// once per device, at network creation (not inside a graph capture).
hipError_t mlp_fused_attr() {
    return hipFuncSetAttribute((const void *)k_mlp_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFusedLds);
}

hipError_t mlp_forward_fused(mavi_mlp *p, const uint16_t *X, int M, float *Y, hipStream_t st) {
    if (M <= 0) return hipSuccess;
    const int ntiles = (M + kTileRows - 1) / kTileRows;
    int grid = (ntiles + kFusedWaves - 1) / kFusedWaves;
    if (grid > p->fused_grid) grid = p->fused_grid;
    hipLaunchKernelGGL(k_mlp_fused, dim3(grid), dim3(64 * kFusedWaves), kFusedLds, st, X, M, p->w[0], p->w[1],
                       p->w[2], p->w[3], p->params, p->b3, Y);
    return hipGetLastError();
}

hipError_t mlp_sync_weights(mavi_mlp *p, hipStream_t st) {
    hipLaunchKernelGGL(k_sync_weights, dim3((kMlpHid * kMlpIn + 255) / 256), dim3(256), 0, st, p->params, p->w[0],
                       p->w[1], p->w[2], p->w[3], p->b3);
    return hipGetLastError();
}

hipError_t ensure_work(mavi_mlp *p, int M) {
    if (M <= p->work_cap) return hipSuccess;
    hipStream_t st = p->ctx->stream;
    (void)hipStreamSynchronize(st);
    for (int l = 0; l < 3; l++) {
        dfree(p->Z[l]);
        dfree(p->H[l]);
    }
    dfree(p->dZa); dfree(p->dZb); dfree(p->Yw); dfree(p->dY); dfree(p->loss_part); dfree(p->Xw); dfree(p->Tw);
    p->work_cap = 0;
    const int64_t m = M;
    hipError_t e = hipSuccess;
    for (int l = 0; l < 3 && e == hipSuccess; l++) {
        e = dalloc(&p->Z[l], m * kMlpHid, st);
        if (e == hipSuccess) e = dalloc(&p->H[l], m * kMlpHid, st);
    }
    if (e == hipSuccess) e = dalloc(&p->dZa, m * kMlpHid, st);
    if (e == hipSuccess) e = dalloc(&p->dZb, m * kMlpHid, st);
    if (e == hipSuccess) e = dalloc(&p->Yw, m * kMlpOutPad, st);
    if (e == hipSuccess) e = dalloc(&p->dY, m * kMlpOutPad, st);
    if (e == hipSuccess) e = dalloc(&p->loss_part, (m + 255) / 256, st);
    if (e == hipSuccess) e = dalloc(&p->Xw, m * kMlpIn, st);
    if (e == hipSuccess) e = dalloc(&p->Tw, m * kMlpOut, st);
    if (e == hipSuccess) p->work_cap = M;
    return e;
}

// Training forward over M rows of X (BF16 [M][160], device): keeps the
// pre-activations Z_l and BF16 activations H_l, output in p->Yw [M][16].
hipError_t mlp_forward_gemm(mavi_mlp *p, const uint16_t *X, int M, hipStream_t st) {
    hipError_t e = ensure_work(p, M);
    if (e != hipSuccess) return e;
    const float *prm = p->params;
    const uint16_t *in = X;
    int kin = kMlpIn;
    const int64_t boff[3] = {kOffB0, kOffB1, kOffB2};
    for (int l = 0; l < 3; l++) {
        Epi ep{prm + boff[l], p->Z[l], kMlpHid, p->H[l], kMlpHid};
        if ((e = gemm_t<false, true, EPI_BIAS_GELU>(st, M, kMlpHid, kin, in, kin, p->w[l], kin, nullptr, 0, 0, 0,
                                                    ep)) != hipSuccess)
            return e;
        in = p->H[l];
        kin = kMlpHid;
    }
    Epi ep{p->b3, nullptr, 0, nullptr, 0};
    return gemm_t<false, true, EPI_BIAS>(st, M, kMlpOutPad, kMlpHid, p->H[2], kMlpHid, p->w[3], kMlpHid, p->Yw,
                                         kMlpOutPad, 0, 0, ep);
}

// One forward + backward pass over M rows (X BF16 [M][160], T [M][2], both
// on the device): loss partials in p->loss_part, gradients in p->grad.
hipError_t mlp_loss_grad_dev(mavi_mlp *p, const uint16_t *X, const float *T, int M, hipStream_t st) {
    hipError_t e = mlp_forward_gemm(p, X, M, st);
    if (e != hipSuccess) return e;
    const int64_t boff[3] = {kOffB0, kOffB1, kOffB2};
    Epi ep{};
    hipLaunchKernelGGL(k_loss, dim3((M + 255) / 256), dim3(256), 0, st, p->Yw, T, M, p->dY, p->loss_part);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    // backward.  Weight gradients: split-K over rows, chunks of kSplitRows.
    const int kc = kSplitRows;
    const int nsplit = (M + kc - 1) / kc;
    auto wgrad = [&](const uint16_t *dZ, int ldz_, int nout, const uint16_t *Hin, int kin_, int64_t off,
                     int64_t count) -> hipError_t {
        if ((int64_t)nsplit * nout * kin_ > p->ws_cap) return hipErrorInvalidValue;
        Epi none{};
        hipError_t r = gemm_t<true, false, EPI_F32>(st, nout, kin_, M, dZ, ldz_, Hin, kin_, p->ws, kin_, kc,
                                                    (int64_t)nout * kin_, none);
        if (r != hipSuccess) return r;
        hipLaunchKernelGGL(k_split_reduce, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p->ws, nsplit,
                           (int64_t)nout * kin_, count, p->grad + off);
        return hipGetLastError();
    };
    auto bgrad = [&](const uint16_t *dZ, int ld, int n, int64_t off) -> hipError_t {
        const int nchunk = (M + kColRows - 1) / kColRows;
        if ((int64_t)nchunk * n > p->ws_cap) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_colsum_part, dim3((n + 63) / 64, nchunk), dim3(256), 0, st, dZ, M, n, ld, p->ws);
        hipLaunchKernelGGL(k_split_reduce, dim3((n + 255) / 256), dim3(256), 0, st, p->ws, nchunk, (int64_t)n,
                           (int64_t)n, p->grad + off);
        return hipGetLastError();
    };
    // output layer: dW3t = dY^T H2 (rows 0..1 of the 16-row padded product), db3 = colsum dY
    if ((e = wgrad(p->dY, kMlpOutPad, kMlpOutPad, p->H[2], kMlpHid, kOffW3, (int64_t)kMlpOut * kMlpHid)) !=
        hipSuccess)
        return e;
    if ((e = bgrad(p->dY, kMlpOutPad, kMlpOut, kOffB3)) != hipSuccess) return e;
    // dZ2 = (dY W3) * gelu'(Z2)
    ep = Epi{nullptr, p->Z[2], kMlpHid, p->dZa, kMlpHid};
    if ((e = gemm_t<false, false, EPI_GELU_BWD>(st, M, kMlpHid, kMlpOutPad, p->dY, kMlpOutPad, p->w[3], kMlpHid,
                                                nullptr, 0, 0, 0, ep)) != hipSuccess)
        return e;
    uint16_t *dZ = p->dZa, *dZn = p->dZb;
    const int64_t woff[3] = {kOffW0, kOffW1, kOffW2};
    for (int l = 2; l >= 0; l--) {
        const uint16_t *Hin = l > 0 ? p->H[l - 1] : X;
        const int kin_ = l > 0 ? kMlpHid : kMlpIn;
        if ((e = wgrad(dZ, kMlpHid, kMlpHid, Hin, kin_, woff[l], (int64_t)kMlpHid * kin_)) != hipSuccess) return e;
        if ((e = bgrad(dZ, kMlpHid, kMlpHid, boff[l])) != hipSuccess) return e;
        if (l == 0) break;
        // dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1})
        ep = Epi{nullptr, p->Z[l - 1], kMlpHid, dZn, kMlpHid};
        if ((e = gemm_t<false, false, EPI_GELU_BWD>(st, M, kMlpHid, kMlpHid, dZ, kMlpHid, p->w[l], kMlpHid, nullptr, 0,
                                                    0, 0, ep)) != hipSuccess)
            return e;
        uint16_t *tmp = dZ;
        dZ = dZn;
        dZn = tmp;
    }
    return hipSuccess;
}

hipError_t mlp_adam_step(mavi_mlp *p, float lr, hipStream_t st) {
    p->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)p->adam_t), c2 = 1.0f - powf(b2, (float)p->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, p->params, p->grad, p->m1, p->m2,
                       (int64_t)kMlpParams, lr, b1, b2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return mlp_sync_weights(p, st);
}

double mlp_loss_value(mavi_mlp *p, int M, hipStream_t st, hipError_t *err) {
    const int nb = (M + 255) / 256;
    std::vector<float> h(nb);
    *err = hipMemcpyAsync(h.data(), p->loss_part, sizeof(float) * nb, hipMemcpyDeviceToHost, st);
    if (*err == hipSuccess) *err = hipStreamSynchronize(st);
    double s = 0.0;
    for (float v : h) s += v;
    return s / (2.0 * M);
}

hipError_t mlp_dataset_reserve(mavi_mlp *p, int64_t cap) {
    if (cap <= p->ds_cap) return hipSuccess;
    hipStream_t st = p->ctx->stream;
    uint16_t *nx = nullptr;
    float *nt = nullptr;
    hipError_t e = dalloc(&nx, cap * kMlpIn, st);
    if (e == hipSuccess) e = dalloc(&nt, cap * kMlpOut, st);
    if (e == hipSuccess && p->ds_count > 0) {
        e = hipMemcpyAsync(nx, p->dsX, sizeof(uint16_t) * p->ds_count * kMlpIn, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(nt, p->dsT, sizeof(float) * p->ds_count * kMlpOut, hipMemcpyDeviceToDevice, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        dfree(nx);
        dfree(nt);
        return e;
    }
    dfree(p->dsX);
    dfree(p->dsT);
    p->dsX = nx;
    p->dsT = nt;
    p->ds_cap = cap;
    return hipSuccess;
}

}  // namespace mavi

// Parameter initialisation: W ~ U(-a, a), a = sqrt(6 / (fan_in + fan_out))
// (Glorot), biases zero; splitmix64 stream from the seed.
static void init_params(std::vector<float> &p, uint64_t seed) {
    uint64_t s = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto uni = [&]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); };
    p.assign(kMlpParams, 0.f);
    const int64_t offs[4] = {kOffW0, kOffW1, kOffW2, kOffW3};
    const int fin[4] = {kMlpIn, kMlpHid, kMlpHid, kMlpHid};
    const int fout[4] = {kMlpHid, kMlpHid, kMlpHid, kMlpOut};
    for (int l = 0; l < 4; l++) {
        const double a = sqrt(6.0 / (fin[l] + fout[l]));
        for (int64_t i = 0; i < (int64_t)fin[l] * fout[l]; i++) p[offs[l] + i] = (float)((2.0 * uni() - 1.0) * a);
    }
}

namespace {

// Staging array of mavi_mlp_forward_noalloc: MAVI_MLP_NOALLOC_ROWS feature
// rows (BF16) followed by as many output rows (FP32, padded to kMlpOutPad).
// Every GPU the library runs on holds its own copy (3.1 MB) until the process
// ends.  In host code `scratch` only names the symbol: its device address
// comes from hipGetSymbolAddress, once per device (noalloc_scratch_setup,
// called by mavi_ctx_create with the device current).
constexpr int64_t kNoAllocRows = MAVI_MLP_NOALLOC_ROWS;
constexpr size_t kScratchXBytes = (size_t)kNoAllocRows * kMlpIn * sizeof(uint16_t);
constexpr size_t kScratchBytes = kScratchXBytes + (size_t)kNoAllocRows * kMlpOutPad * sizeof(float);
static_assert(kScratchXBytes % 256 == 0, "the output rows start 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace mavi {
hipError_t noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}
}  // namespace mavi

namespace {
void *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}
}  // namespace

extern "C" {

int mavi_mlp_create(mavi_ctx *ctx, uint64_t seed, mavi_mlp **out) {
    if (!ctx || !out) return MAVI_E_BADARG;
    *out = nullptr;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    mavi_mlp *p = new mavi_mlp();
    p->ctx = ctx;
    hipStream_t st = ctx->stream;
    hipError_t e = dalloc(&p->params, kMlpParams, st);
    if (e == hipSuccess) e = dalloc(&p->grad, kMlpParams, st);
    if (e == hipSuccess) e = dalloc(&p->m1, kMlpParams, st);
    if (e == hipSuccess) e = dalloc(&p->m2, kMlpParams, st);
    if (e == hipSuccess) e = dalloc(&p->w[0], (int64_t)kMlpHid * kMlpIn, st);
    if (e == hipSuccess) e = dalloc(&p->w[1], (int64_t)kMlpHid * kMlpHid, st);
    if (e == hipSuccess) e = dalloc(&p->w[2], (int64_t)kMlpHid * kMlpHid, st);
    if (e == hipSuccess) e = dalloc(&p->w[3], (int64_t)kMlpOutPad * kMlpHid, st);
    if (e == hipSuccess) e = dalloc(&p->b3, kMlpOutPad, st);
    p->ws_cap = (int64_t)64 * kMlpHid * kMlpIn;
    if (e == hipSuccess) e = dalloc(&p->ws, p->ws_cap, st);
    int dev = 0, ncu = 256;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    p->fused_grid = ncu > 0 ? ncu : 256;
    if (e == hipSuccess) e = mavi::mlp_fused_attr();
    std::vector<float> h;
    init_params(h, seed);
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->params, h.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = mavi::mlp_sync_weights(p, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        std::string m = std::string("mlp allocation failed: ") + hipGetErrorString(e);
        mavi_mlp_destroy(p);
        return set_err(ctx, MAVI_E_HIP, m);
    }
    ctx_retain(ctx);
    p->ctx_held = true;
    *out = p;
    return MAVI_OK;
}

static void mlp_free(mavi_mlp *p);

void mlp_retain(mavi_mlp *p) { p->refs++; }

void mlp_release(mavi_mlp *p) {
    if (--p->refs == 0 && p->closing) mlp_free(p);
}

int mavi_mlp_destroy(mavi_mlp *p) {
    if (!p) return MAVI_OK;
    if (p->refs > 0) {  // a learned propagator still runs it
        p->closing = true;
        return MAVI_OK;
    }
    mlp_free(p);
    return MAVI_OK;
}

static void mlp_free(mavi_mlp *p) {
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    dfree(p->params); dfree(p->grad); dfree(p->m1); dfree(p->m2);
    for (int l = 0; l < 4; l++) dfree(p->w[l]);
    dfree(p->b3); dfree(p->ws);
    for (int l = 0; l < 3; l++) {
        dfree(p->Z[l]);
        dfree(p->H[l]);
    }
    dfree(p->dZa); dfree(p->dZb); dfree(p->Yw); dfree(p->dY); dfree(p->loss_part); dfree(p->Xw); dfree(p->Tw);
    dfree(p->dsX); dfree(p->dsT); dfree(p->bX); dfree(p->bT);
    mavi_ctx *ctx = p->ctx;
    const bool held = p->ctx_held;
    delete p;
    if (held) ctx_release(ctx);
}

int mavi_mlp_num_params(void) { return kMlpParams; }

int mavi_mlp_get_params(mavi_mlp *p, float *params) {
    if (!p || !params) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(params, p->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MAVI_OK;
}

int mavi_mlp_set_params(mavi_mlp *p, const float *params) {
    if (!p || !params) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemcpyAsync(p->params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, mavi::mlp_sync_weights(p, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MAVI_OK;
}

int mavi_mlp_forward(mavi_mlp *p, int64_t m, const uint16_t *x, float *y, int32_t fused) {
    if (!p || !x || !y || m <= 0 || m > (1 << 28)) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    uint16_t *dX = nullptr;
    float *dYo = nullptr;
    HIPCHK(ctx, dalloc(&dX, m * kMlpIn, st));
    hipError_t e = dalloc(&dYo, m * kMlpOutPad, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dX, x, sizeof(uint16_t) * m * kMlpIn, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        if (fused) {
            e = mavi::mlp_forward_fused(p, dX, (int)m, dYo, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(y, dYo, sizeof(float) * m * kMlpOut, hipMemcpyDeviceToHost, st);
        } else {
            // the training forward: GEMM chain, output read from the padded Y
            e = mavi::mlp_forward_gemm(p, dX, (int)m, st);
            if (e == hipSuccess)
                e = hipMemcpy2DAsync(y, sizeof(float) * kMlpOut, p->Yw, sizeof(float) * kMlpOutPad,
                                     sizeof(float) * kMlpOut, m, hipMemcpyDeviceToHost, st);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dfree(dX);
    dfree(dYo);
    HIPCHK(ctx, e);
    return MAVI_OK;
}

int mavi_mlp_forward_noalloc(mavi_mlp *p, int64_t m, const uint16_t *x, float *y) {
    if (!p || !x || !y || m <= 0) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    if (m > kNoAllocRows)
        return set_err(ctx, MAVI_E_BADARG, "mavi_mlp_forward_noalloc takes at most MAVI_MLP_NOALLOC_ROWS (" +
                                               std::to_string(kNoAllocRows) + ") rows, got " + std::to_string(m));
    HIPCHK(ctx, hipSetDevice(ctx->device));
    char *base = (char *)noalloc_scratch(ctx->device);
    if (!base) return set_err(ctx, MAVI_E_HIP, "no-alloc staging array not set up on this device");
    uint16_t *dX = (uint16_t *)base;
    float *dY = (float *)(base + kScratchXBytes);
    hipStream_t st = ctx->stream;
    // every staged row is written before the kernel reads it, and the kernel
    // writes every output row it is given: nothing needs clearing
    hipError_t e = hipMemcpyAsync(dX, x, sizeof(uint16_t) * m * kMlpIn, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = mavi::mlp_forward_fused(p, dX, (int)m, dY, st);
    if (e == hipSuccess) e = hipMemcpyAsync(y, dY, sizeof(float) * m * kMlpOut, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    HIPCHK(ctx, e);
    return MAVI_OK;
}

int mavi_mlp_loss_grad(mavi_mlp *p, int64_t m, const uint16_t *x, const float *t, double *loss, float *grad) {
    if (!p || !x || !t || m <= 0 || m > (1 << 24)) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if ((int64_t)((m + kSplitRows - 1) / kSplitRows) * kMlpHid * kMlpIn > p->ws_cap)
        return set_err(ctx, MAVI_E_BADARG, "batch too large for the split-K work space");
    HIPCHK(ctx, mavi::ensure_work(p, (int)m));
    HIPCHK(ctx, hipMemcpyAsync(p->Xw, x, sizeof(uint16_t) * m * kMlpIn, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(p->Tw, t, sizeof(float) * m * kMlpOut, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, mavi::mlp_loss_grad_dev(p, p->Xw, p->Tw, (int)m, st));
    hipError_t e = hipSuccess;
    const double l = mavi::mlp_loss_value(p, (int)m, st, &e);
    HIPCHK(ctx, e);
    if (loss) *loss = l;
    if (grad) {
        HIPCHK(ctx, hipMemcpyAsync(grad, p->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
    }
    return MAVI_OK;
}

int mavi_mlp_dataset_clear(mavi_mlp *p) {
    if (!p) return MAVI_E_BADARG;
    p->ds_count = 0;
    return MAVI_OK;
}

int mavi_mlp_dataset_append(mavi_mlp *p, int64_t n, const uint16_t *x, const float *t) {
    if (!p || !x || !t || n <= 0) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (p->ds_count + n > p->ds_cap) {
        int64_t cap = p->ds_cap > 0 ? p->ds_cap : 4096;
        while (cap < p->ds_count + n) cap *= 2;
        HIPCHK(ctx, mavi::mlp_dataset_reserve(p, cap));
    }
    HIPCHK(ctx, hipMemcpyAsync(p->dsX + p->ds_count * kMlpIn, x, sizeof(uint16_t) * n * kMlpIn,
                               hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(p->dsT + p->ds_count * kMlpOut, t, sizeof(float) * n * kMlpOut,
                               hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    p->ds_count += n;
    return MAVI_OK;
}

int mavi_mlp_dataset_size(mavi_mlp *p, int64_t *n) {
    if (!p || !n) return MAVI_E_BADARG;
    *n = p->ds_count;
    return MAVI_OK;
}

int mavi_mlp_dataset_get(mavi_mlp *p, int64_t first, int64_t n, uint16_t *x, float *t) {
    if (!p || first < 0 || n < 0 || first + n > p->ds_count) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (x)
        HIPCHK(ctx, hipMemcpyAsync(x, p->dsX + first * kMlpIn, sizeof(uint16_t) * n * kMlpIn, hipMemcpyDeviceToHost,
                                   st));
    if (t)
        HIPCHK(ctx, hipMemcpyAsync(t, p->dsT + first * kMlpOut, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost,
                                   st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return MAVI_OK;
}

int mavi_mlp_train(mavi_mlp *p, int64_t steps, int32_t batch, float lr, uint64_t seed, float *losses) {
    if (!p || steps < 0 || batch <= 0 || batch > (1 << 20) || !(lr > 0.f)) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    if (p->ds_count <= 0) return set_err(ctx, MAVI_E_BADARG, "empty data set");
    if ((int64_t)((batch + kSplitRows - 1) / kSplitRows) * kMlpHid * kMlpIn > p->ws_cap)
        return set_err(ctx, MAVI_E_BADARG, "batch too large for the split-K work space");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (batch > p->b_cap) {
        dfree(p->bX);
        dfree(p->bT);
        p->b_cap = 0;
        HIPCHK(ctx, dalloc(&p->bX, (int64_t)batch * kMlpIn, st));
        HIPCHK(ctx, dalloc(&p->bT, (int64_t)batch * kMlpOut, st));
        p->b_cap = batch;
    }
    if (steps > p->loss_hist_cap) {
        dfree(p->loss_hist);
        p->loss_hist_cap = 0;
        HIPCHK(ctx, dalloc(&p->loss_hist, steps, st));
        p->loss_hist_cap = steps;
    }
    HIPCHK(ctx, mavi::ensure_work(p, batch));
    const int nchunk = batch * (kMlpIn / 8);
    for (int64_t s = 0; s < steps; s++) {
        hipLaunchKernelGGL(k_gather_batch, dim3((nchunk + 255) / 256), dim3(256), 0, st, p->dsX, p->dsT, p->ds_count,
                           seed, s, batch, p->bX, p->bT);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, mavi::mlp_loss_grad_dev(p, p->bX, p->bT, batch, st));
        hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(64), 0, st, p->loss_part, (batch + 255) / 256, batch,
                           p->loss_hist + s);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, mavi::mlp_adam_step(p, lr, st));
    }
    if (losses && steps > 0)
        HIPCHK(ctx, hipMemcpyAsync(losses, p->loss_hist, sizeof(float) * steps, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    return MAVI_OK;
}

int mavi_mlp_gemm_check(mavi_ctx *ctx, int32_t m, int32_t n, int32_t k, int32_t ta, int32_t tb, const uint16_t *a,
                        const uint16_t *b, float *c) {
    if (!ctx || !a || !b || !c || m <= 0 || n <= 0 || k <= 0 || m > 4096 || n > 4096 || k > 4096)
        return MAVI_E_BADARG;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    uint16_t *dA = nullptr, *dB = nullptr;
    float *dC = nullptr;
    hipError_t e = dalloc(&dA, (int64_t)m * k, st);
    if (e == hipSuccess) e = dalloc(&dB, (int64_t)k * n, st);
    if (e == hipSuccess) e = dalloc(&dC, (int64_t)m * n, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dA, a, sizeof(uint16_t) * m * k, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dB, b, sizeof(uint16_t) * k * n, hipMemcpyHostToDevice, st);
    // A is [m][k] (or [k][m] when ta), B is [k][n] (or [n][k] when tb)
    const int lda = ta ? m : k, ldb = tb ? k : n;
    if (e == hipSuccess) {
        if (!ta && !tb) e = mavi::gemm_t<false, false, EPI_F32>(st, m, n, k, dA, lda, dB, ldb, dC, n, 0, 0, Epi{});
        else if (!ta && tb) e = mavi::gemm_t<false, true, EPI_F32>(st, m, n, k, dA, lda, dB, ldb, dC, n, 0, 0, Epi{});
        else if (ta && !tb) e = mavi::gemm_t<true, false, EPI_F32>(st, m, n, k, dA, lda, dB, ldb, dC, n, 0, 0, Epi{});
        else e = mavi::gemm_t<true, true, EPI_F32>(st, m, n, k, dA, lda, dB, ldb, dC, n, 0, 0, Epi{});
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c, dC, sizeof(float) * m * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dfree(dA);
    dfree(dB);
    dfree(dC);
    HIPCHK(ctx, e);
    return MAVI_OK;
}

int mavi_mlp_time_forward(mavi_mlp *p, int64_t m, int32_t reps, double *us_per_call) {
    if (!p || m <= 0 || m > (1 << 26) || reps <= 0 || !us_per_call) return MAVI_E_BADARG;
    mavi_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    uint16_t *dX = nullptr;
    float *dY = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = dalloc(&dX, m * kMlpIn, st);
    if (e == hipSuccess) e = dalloc(&dY, m * kMlpOut, st);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = mavi::mlp_forward_fused(p, dX, (int)m, dY, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(a, st);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = mavi::mlp_forward_fused(p, dX, (int)m, dY, st);
    if (e == hipSuccess) e = hipEventRecord(b, st);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    dfree(dX);
    dfree(dY);
    HIPCHK(ctx, e);
    *us_per_call = 1e3 * ms / reps;
    return MAVI_OK;
}

}  // extern "C"
