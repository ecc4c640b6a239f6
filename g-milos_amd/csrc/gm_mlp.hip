/*
 * gm_mlp.hip -- learned initial guess: a 4-layer perceptron trained and run
 * with hand-written BF16 MFMA kernels for gfx950.
 *
 * Network: input 4 x nlambda Stokes samples normalised by the continuum
 * (padded to a multiple of 64), three hidden layers of 256 with GELU, 64
 * outputs of which 11 are used: B, cos/sin(gamma), cos/sin(2 phi), vlos,
 * dopp, a, eta0, S0/Ic, Ic (standardised).  BF16 weights and activations,
 * FP32 accumulation, FP32 master weights updated by AdamW.
 *
 * GEMM tiles: a workgroup of 4 waves computes a 64 x 64 tile of C = A.B with
 * v_mfma_f32_16x16x16_bf16 (each wave a 32 x 32 quarter = 2 x 2 MFMA
 * blocks), staging 64 x 32 slices of A and B through LDS.
 */
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "gm_mlp.h"

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

/* FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950) */
__device__ __forceinline__ uint16_t f2bf(float x)
{
    const __bf16 b = (__bf16)x;
    return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

/* GELU, tanh form (Hendrycks & Gimpel 2016), and its derivative, through
 * 0.5 x (1 + tanh u) = x s with s = 1 / (1 + exp(-2u)),
 * u = sqrt(2/pi) (x + 0.044715 x^3): one exp2 and one reciprocal
 * (v_exp_f32, v_rcp_f32, about 1 ulp each, far below the BF16 rounding of
 * the activations).  For large |x|, exp2 overflows to inf or underflows to
 * 0 and s saturates to 0 or 1. */
__device__ __forceinline__ float gelu_s(float x)
{
    constexpr float k = 0.7978845608028654f;          /* sqrt(2/pi) */
    constexpr float l2e = 1.4426950408889634f;        /* log2(e)    */
    const float m = x * fmaf(x * x, -2.f * k * 0.044715f * l2e, -2.f * k * l2e);   /* -2u log2 e */
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(m));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
__device__ __forceinline__ float gelu_grad(float x)
{
    constexpr float k = 0.7978845608028654f;
    const float s = gelu_s(x);
    const float du = k * fmaf(3.f * 0.044715f, x * x, 1.f);
    return fmaf(2.f * x * s * (1.f - s), du, s);
}

/* C[M x N] (+)= op(A)[M x K] . op(B)[K x N].
 * A is row-major M x K (lda) or, with TA, stored K x M (A^T row-major).
 * B is row-major K x N (ldb) or, with TB, stored N x K.
 * M, N multiples of 64; K multiple of 32.
 * Epilogues: EPI_BIAS_GELU (store pre-activation f32 and GELU bf16),
 *            EPI_BIAS (store f32 and bf16), EPI_GELU_GRAD (multiply by
 *            GELU'(Zprev) and store bf16 + f32), EPI_STORE (f32). */
constexpr int EPI_BIAS_GELU = GM_EPI_BIAS_GELU, EPI_BIAS = GM_EPI_BIAS,
              EPI_GELU_GRAD = GM_EPI_GELU_GRAD, EPI_STORE = GM_EPI_STORE;

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const uint16_t *__restrict__ A, int lda,
                                              const uint16_t *__restrict__ B, int ldb,
                                              int M, int N, int K,
                                              const float *__restrict__ bias,
                                              const float *__restrict__ zprev, int ldz,
                                              float *__restrict__ Cf, uint16_t *__restrict__ Cb,
                                              int ldc)
{
    __shared__ uint16_t sA[64][32 + 8];   /* [m][k] */
    __shared__ uint16_t sB[64][32 + 8];   /* [n][k] */
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += 32) {
        /* stage 64 x 32 of A and B (2048 elements each, 8 per thread) */
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid * 8 + e;
            const int r = idx >> 5, c = idx & 31;          /* r: m or n, c: k */
            sA[r][c] = TA ? A[(size_t)(k0 + c) * lda + (m0 + r)] : A[(size_t)(m0 + r) * lda + (k0 + c)];
            sB[r][c] = TB ? B[(size_t)(n0 + r) * ldb + (k0 + c)] : B[(size_t)(k0 + c) * ldb + (n0 + r)];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 16) {
            short4_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int row = wm + i * 16 + (lane & 15);
                const int kb = kk + 4 * (lane >> 4);
                a[i] = short4_t{(short)sA[row][kb], (short)sA[row][kb + 1], (short)sA[row][kb + 2],
                                (short)sA[row][kb + 3]};
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int col = wn + j * 16 + (lane & 15);
                const int kb = kk + 4 * (lane >> 4);
                b[j] = short4_t{(short)sB[col][kb], (short)sB[col][kb + 1], (short)sB[col][kb + 2],
                                (short)sB[col][kb + 3]};
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    /* D layout: lane holds rows 4*(lane/16)+r, column lane%16 */
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + j * 16 + (lane & 15);
                float v = acc[i][j][r];
                const size_t o = (size_t)m * ldc + n;
                if (EPI == EPI_BIAS_GELU) {
                    v += bias[n];
                    if (Cf) Cf[o] = v;
                    Cb[o] = f2bf(gelu(v));
                } else if (EPI == EPI_BIAS) {
                    v += bias[n];
                    Cf[o] = v;
                    if (Cb) Cb[o] = f2bf(v);
                } else if (EPI == EPI_GELU_GRAD) {
                    v *= gelu_grad(zprev[(size_t)m * ldz + n]);
                    if (Cf) Cf[o] = v;
                    Cb[o] = f2bf(v);
                } else {
                    Cf[o] = v;
                }
            }
}

/* C[M x N] = A[M x K] . B[K x N], A and B row-major BF16, for the forward
 * pass: 128 x TN tiles per 256-thread workgroup, each wave a 64 x TN/2
 * block of v_mfma_f32_16x16x32_bf16 (gfx950), K staged 32 at a time
 * through LDS with 16-byte global loads; B is stored transposed in LDS so
 * that both MFMA operands are 16-byte LDS reads.  Operand layout: lane l
 * supplies row (A) / column (B) l % 16 and the eight k = 8 (l / 16) .. +7;
 * the accumulator holds rows 4 (l / 16) .. +3 of column l % 16.
 * M % 128 == 0, N % TN == 0, K % 32 == 0.                                  */
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <int TN, int EPI>
__global__ __launch_bounds__(256) void k_gemm_nn(const uint16_t *__restrict__ A, int lda,
                                                 const uint16_t *__restrict__ B, int ldb, int K,
                                                 const float *__restrict__ bias,
                                                 float *__restrict__ Cf, uint16_t *__restrict__ Cb,
                                                 int ldc)
{
    constexpr int FM = 4, FN = TN / 32;            /* 16 x 16 fragments per wave */
    constexpr int LS = 32 + 8;                     /* LDS row stride (bf16)       */
    __shared__ __attribute__((aligned(16))) uint16_t sA[128 * LS];
    __shared__ __attribute__((aligned(16))) uint16_t sB[TN * LS];   /* [n][k] */
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * 128, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * (TN / 2);
    float4_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int kq8 = 8 * (lane >> 4), l16 = lane & 15;
    for (int k0 = 0; k0 < K; k0 += 32) {
#pragma unroll
        for (int it = 0; it < 2; ++it) {                /* A: 512 16-byte pieces */
            const int idx = tid + 256 * it;
            const int r = idx >> 2, q = idx & 3;
            *reinterpret_cast<uint4 *>(&sA[r * LS + q * 8]) =
                *reinterpret_cast<const uint4 *>(&A[(size_t)(m0 + r) * lda + k0 + q * 8]);
        }
#pragma unroll
        for (int it = 0; it < TN / 64; ++it) {          /* B: 4 TN pieces */
            const int idx = tid + 256 * it;
            const int kr = idx / (TN / 8), q = idx % (TN / 8);
            const uint4 v = *reinterpret_cast<const uint4 *>(&B[(size_t)(k0 + kr) * ldb + n0 + q * 8]);
            const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
            for (int c = 0; c < 8; ++c) sB[(q * 8 + c) * LS + kr] = e[c];
        }
        __syncthreads();
        bf16x8_t a[FM], b[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
            a[i] = *reinterpret_cast<const bf16x8_t *>(&sA[(wm + i * 16 + l16) * LS + kq8]);
#pragma unroll
        for (int j = 0; j < FN; ++j)
            b[j] = *reinterpret_cast<const bf16x8_t *>(&sB[(wn + j * 16 + l16) * LS + kq8]);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn + j * 16 + l16;
                float v = acc[i][j][r] + bias[n];
                const size_t o = (size_t)m * ldc + n;
                if (Cf) Cf[o] = v;                 /* pre-activations: training only */
                if (EPI == EPI_BIAS_GELU) Cb[o] = f2bf(gelu(v));
                else if (Cb) Cb[o] = f2bf(v);
            }
}

/* Input: obs float [npix][4][nl] -> bf16 [npix][kin], normalised by the
 * continuum Ic = max(I[0], I[nl-1]) (as x * (1/Ic)); I -> 2 (I/Ic - 1),
 * Q, U, V -> 10 x/Ic.  One thread per 8 consecutive features (one 16-byte
 * store). */
__global__ void k_prep_input(const float *__restrict__ obs, int64_t npix, int nl, int kin,
                             uint16_t *__restrict__ x)
{
    const int per = kin / 8;
    const int64_t total = npix * per;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = i / per;
        const int f0 = (int)(i % per) * 8;
        const float *o = obs + p * 4 * nl;
        const float rc = 1.f / fmaxf(fmaxf(o[0], o[nl - 1]), 1e-6f);
        uint16_t h[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int f = f0 + e;
            float v = 0.f;
            if (f < 4 * nl) {
                v = o[f] * rc;
                v = (f < nl) ? (v - 1.f) * 2.f : v * 10.f;
            }
            h[e] = f2bf(v);
        }
        *reinterpret_cast<uint4 *>(&x[p * kin + f0]) = *reinterpret_cast<const uint4 *>(h);
    }
}

/* Targets from true models (standardised; gm_mlp.h documents the map). */
__global__ void k_prep_target(const double *__restrict__ models, int64_t npix,
                              float *__restrict__ t)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x) {
        const double *m = models + p * 11;
        const double g = m[5] * (3.14159265358979323846 / 180.0);
        const double ph = 2.0 * m[6] * (3.14159265358979323846 / 180.0);
        const double ic = m[7] + m[8];
        float *o = t + p * GM_MLP_NOUT;
        o[0] = (float)((m[1] - 1500.0) / 866.0);
        o[1] = (float)cos(g);
        o[2] = (float)(sin(g) * 2.0 - 1.0);
        o[3] = (float)cos(ph);
        o[4] = (float)sin(ph);
        o[5] = (float)(m[2] / 1.732);
        o[6] = (float)((m[3] - 0.04) / 0.01155);
        o[7] = (float)((m[4] - 0.255) / 0.1414);
        o[8] = (float)((m[0] - 10.5) / 5.485);
        o[9] = (float)((m[7] / ic - 0.25) / 0.1);
        o[10] = (float)((ic - 1.0) / 0.2);
        for (int k = 11; k < GM_MLP_NOUT; ++k) o[k] = 0.f;
    }
}

/* Output decode into a start model (init fields untouched where the network
 * has no output), clamped to the LM's physical ranges.  FP32 arithmetic:
 * the result is a starting point for the LM. */
__device__ __forceinline__ void decode_row(const float *__restrict__ o, const double *__restrict__ base,
                                           double *__restrict__ m)
{
    const float rad = 57.29577951308232f;
    const float B = o[0] * 866.f + 1500.f;
    const float sg = fmaxf((o[2] + 1.f) * 0.5f, 0.f);
    const float gm = atan2f(sg, o[1]) * rad;
    float az = 0.5f * atan2f(o[4], o[3]) * rad;
    if (az < 0.f) az += 180.f;
    const float ic = fmaxf(o[10] * 0.2f + 1.f, 0.05f);
    const float s0f = fminf(fmaxf(o[9] * 0.1f + 0.25f, 0.01f), 0.99f);
    m[0] = fminf(fmaxf(o[8] * 5.485f + 10.5f, 1.f), 2500.f);
    m[1] = fminf(fmaxf(B, 0.f), 5000.f);
    m[2] = fminf(fmaxf(o[5] * 1.732f, -20.f), 20.f);
    m[3] = fminf(fmaxf(o[6] * 0.01155f + 0.04f, 0.005f), 0.6f);
    m[4] = fminf(fmaxf(o[7] * 0.1414f + 0.255f, 0.0001f), 10.f);
    m[5] = fminf(fmaxf(gm, 0.f), 180.f);
    m[6] = fminf(fmaxf(az, 0.f), 180.f);
    m[7] = s0f * ic;
    m[8] = (1.f - s0f) * ic;
    m[9] = base[9];
    m[10] = base[10];
}

__global__ void k_decode(const float *__restrict__ y, int64_t npix, const double *__restrict__ base,
                         double *__restrict__ models)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
         p += (int64_t)gridDim.x * blockDim.x)
        decode_row(y + p * GM_MLP_NOUT, base, models + p * 11);
}

/* Fused inference for kin <= 256: input normalisation (as k_prep_input),
 * the four layers and the decode in one kernel, so that no activation
 * leaves the CU.  A workgroup takes 64 pixels; their activations stay in LDS
 * (64 x 256 BF16, overwritten layer by layer) and each layer's weights,
 * transposed to [out][in] so that a B fragment is one 16-byte read, are
 * staged through LDS 64 k at a time and shared by the four waves.  Wave w
 * computes output columns [w N/4, (w+1) N/4) of all 64 rows: 4 x N/64
 * fragments of v_mfma_f32_16x16x32_bf16.  The accumulation order over k,
 * the bias/GELU epilogue and the BF16 rounding are those of k_gemm_nn, so
 * the outputs equal the layered path's.                                    */
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
constexpr int FX_ROWS = 64;
constexpr int FX_AS = GM_MLP_HIDDEN + 8;   /* activation row stride (bf16) */
constexpr int FX_WS = 64 + 8;              /* weight row stride (bf16)      */

/* One layer of k_mlp_infer.  `pre` holds this layer's first 64-k weight
 * chunk on entry (16-byte pieces, PER per thread); the last chunk's
 * iteration loads the next layer's first chunk (wtx, row length KX, FNX x 64
 * outputs) into it, so no layer starts by waiting for memory. */
template <int FN, int FNX>
__device__ __forceinline__ void fused_layer(uint16_t *__restrict__ sX, uint16_t *__restrict__ sW,
                                            const uint16_t *__restrict__ wt, const float *__restrict__ bias,
                                            int K, float *__restrict__ sY, u32x4_t (&pre)[8],
                                            const uint16_t *__restrict__ wtx, int KX)
{
    constexpr int N = 64 * FN, PER = N * 8 / 256, PERX = 64 * FNX * 8 / 256;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wn = wave * (N / 4), kq8 = 8 * (lane >> 4), l16 = lane & 15;
    float4_t acc[4][FN];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 64) {
        __syncthreads();                       /* previous reads of sW (and sX writes) done */
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int q = tid + 256 * e, n = q >> 3, c = q & 7;
            *reinterpret_cast<u32x4_t *>(&sW[n * FX_WS + c * 8]) = pre[e];
        }
        __syncthreads();
        if (k0 + 64 < K) {
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int q = tid + 256 * e, n = q >> 3, c = q & 7;
                pre[e] = *reinterpret_cast<const u32x4_t *>(&wt[(size_t)n * K + k0 + 64 + c * 8]);
            }
        } else {
#pragma unroll
            for (int e = 0; e < PERX; ++e) {
                const int q = tid + 256 * e, n = q >> 3, c = q & 7;
                pre[e] = *reinterpret_cast<const u32x4_t *>(&wtx[(size_t)n * KX + c * 8]);
            }
        }
#pragma unroll
        for (int kk = 0; kk < 64; kk += 32) {
            bf16x8_t a[4], b[FN];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                a[i] = *reinterpret_cast<const bf16x8_t *>(&sX[(i * 16 + l16) * FX_AS + k0 + kk + kq8]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                b[j] = *reinterpret_cast<const bf16x8_t *>(&sW[(wn + j * 16 + l16) * FX_WS + kk + kq8]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();                           /* every wave has read this layer's input */
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = i * 16 + 4 * (lane >> 4) + r;
                const int col = wn + j * 16 + l16;
                const float v = acc[i][j][r] + bias[col];
                if (sY) {
                    if (col < 11) sY[row * 12 + col] = v;
                } else {
                    sX[row * FX_AS + col] = f2bf(gelu(v));
                }
            }
}

/* obs must be 16-byte aligned (4 nl floats per pixel keep every row so). */
__global__ __launch_bounds__(256, 2) void k_mlp_infer(const float *__restrict__ obs, int64_t npix, int nl,
                                                      int kin, const uint16_t *__restrict__ wt,
                                                      const float *__restrict__ w, gm_mlp_offsets of,
                                                      const double *__restrict__ base,
                                                      double *__restrict__ models)
{
    __shared__ __attribute__((aligned(16))) uint16_t sX[FX_ROWS * FX_AS];
    __shared__ __attribute__((aligned(16))) uint16_t sW[GM_MLP_HIDDEN * FX_WS];
    __shared__ float sIc[FX_ROWS];          /* 1 / Ic per row */
    __shared__ float sY[FX_ROWS * 12];
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * FX_ROWS;
    const int rows = (npix - p0 < FX_ROWS) ? (int)(npix - p0) : FX_ROWS;
    const int nf = 4 * nl, n4 = rows * nf / 4;
    /* first weight chunk and the whole input tile in flight together */
    u32x4_t pre[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int q = tid + 256 * e, n = q >> 3, c = q & 7;
        pre[e] = *reinterpret_cast<const u32x4_t *>(&wt[of.woff[0] + (size_t)n * kin + c * 8]);
    }
    const float4 *src = reinterpret_cast<const float4 *>(obs + p0 * nf);
    float4 v[FX_ROWS * GM_MLP_HIDDEN / 4 / 256];
#pragma unroll
    for (int j = 0; j < FX_ROWS * GM_MLP_HIDDEN / 4 / 256; ++j) {
        const int idx = tid + 256 * j;
        if (idx < n4) v[j] = src[idx];
    }
    if (tid < FX_ROWS) {
        float rc = 1.f;
        if (tid < rows) {
            const float *o = obs + (p0 + tid) * nf;
            rc = 1.f / fmaxf(fmaxf(o[0], o[nl - 1]), 1e-6f);
        }
        sIc[tid] = rc;
    }
    /* zero the pad columns and the rows past the end of the cube */
    for (int e = tid; e < FX_ROWS * (kin - nf); e += 256) {
        const int r = e / (kin - nf);
        sX[r * FX_AS + nf + (e - r * (kin - nf))] = 0;
    }
    for (int e = tid; e < (FX_ROWS - rows) * nf; e += 256) {
        const int r = rows + e / nf;
        sX[r * FX_AS + (e - (r - rows) * nf)] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < FX_ROWS * GM_MLP_HIDDEN / 4 / 256; ++j) {
        const int idx = tid + 256 * j;
        if (idx < n4) {
            const int r = (4 * idx) / nf, f = 4 * idx - r * nf;
            const float rc = sIc[r];
            const float x[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
            uint16_t h[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float y = x[c] * rc;
                y = (f + c < nl) ? (y - 1.f) * 2.f : y * 10.f;
                h[c] = f2bf(y);
            }
            *reinterpret_cast<uint2 *>(&sX[r * FX_AS + f]) = *reinterpret_cast<const uint2 *>(h);
        }
    }
    fused_layer<4, 4>(sX, sW, wt + of.woff[0], w + of.boff[0], kin, nullptr, pre, wt + of.woff[1],
                      GM_MLP_HIDDEN);
    fused_layer<4, 4>(sX, sW, wt + of.woff[1], w + of.boff[1], GM_MLP_HIDDEN, nullptr, pre,
                      wt + of.woff[2], GM_MLP_HIDDEN);
    fused_layer<4, 1>(sX, sW, wt + of.woff[2], w + of.boff[2], GM_MLP_HIDDEN, nullptr, pre,
                      wt + of.woff[3], GM_MLP_HIDDEN);
    fused_layer<1, 0>(sX, sW, wt + of.woff[3], w + of.boff[3], GM_MLP_HIDDEN, sY, pre, nullptr, 0);
    __syncthreads();
    if (tid < rows) decode_row(&sY[tid * 12], base, models + (p0 + tid) * 11);
}

/* W [in][out] (row-major, BF16) -> WT [out][in] for one layer. */
__global__ void k_transpose_w(const uint16_t *__restrict__ w, int in, int out, uint16_t *__restrict__ wt)
{
    const int64_t n = (int64_t)in * out;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int o = (int)(i / in), k = (int)(i % in);
        wt[i] = w[(int64_t)k * out + o];
    }
}

/* Mean-squared error over the used outputs; dY = 2 (Y - T) / (B * 11). */
__global__ void k_loss(const float *__restrict__ y, const float *__restrict__ t, int64_t b,
                       uint16_t *__restrict__ dy, float *__restrict__ loss_acc)
{
    float local = 0.f;
    const float sc = 2.f / (float)(b * 11);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < b * GM_MLP_NOUT;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(i % GM_MLP_NOUT);
        float d = (k < 11) ? (y[i] - t[i]) : 0.f;
        local += d * d;
        dy[i] = f2bf(d * sc);
    }
    for (int o = 32; o >= 1; o >>= 1) local += __shfl_xor(local, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss_acc, local / (float)(b * 11));
}

/* db[n] = sum_b dY[b][n] (dY bf16): a block sums 64 columns over a slice
 * of rows (four row groups, coalesced 128-byte row reads), reduces the
 * groups in LDS and adds its partial sum to db (zeroed beforehand). */
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ dy, int64_t b, int n,
                                                   int64_t rows_per_block, float *__restrict__ db)
{
    __shared__ float part[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + c;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_block;
    const int64_t r1 = (r0 + rows_per_block < b) ? r0 + rows_per_block : b;
    float s = 0.f;
    if (col < n)
        for (int64_t r = r0 + g; r < r1; r += 4) s += bf2f(dy[r * n + col]);
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && col < n) atomicAdd(&db[col], part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

/* AdamW (Loshchilov & Hutter 2019) on FP32 master weights; refreshes the
 * BF16 copy. */
__global__ void k_adamw(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                        float *__restrict__ v, uint16_t *__restrict__ wb, int64_t n, float lr,
                        float b1, float b2, float eps, float wd, float bc1, float bc2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        float wi = w[i] * (1.f - lr * wd);
        wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

} // namespace

/* ---- launchers ------------------------------------------------------------ */
#define GEMM_CASE(TA, TB, EPI)                                                              \
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, A, lda, B, ldb, M, N, K, \
                       bias, zprev, ldz, Cf, Cb, ldc)

void gm_mlp_gemm(int ta, int tb, int epi, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                 int M, int N, int K, const float *bias, const float *zprev, int ldz, float *Cf,
                 uint16_t *Cb, int ldc, hipStream_t s)
{
    if (!ta && !tb && (epi == EPI_BIAS_GELU || epi == EPI_BIAS) && M % 128 == 0 && K % 32 == 0 &&
        (lda % 8) == 0 && (ldb % 8) == 0) {
        if (N % 128 == 0) {
            dim3 g(N / 128, M / 128);
            if (epi == EPI_BIAS_GELU)
                hipLaunchKernelGGL((k_gemm_nn<128, EPI_BIAS_GELU>), g, dim3(256), 0, s, A, lda, B, ldb,
                                   K, bias, Cf, Cb, ldc);
            else
                hipLaunchKernelGGL((k_gemm_nn<128, EPI_BIAS>), g, dim3(256), 0, s, A, lda, B, ldb, K,
                                   bias, Cf, Cb, ldc);
            return;
        }
        if (N % 64 == 0) {
            dim3 g(N / 64, M / 128);
            if (epi == EPI_BIAS_GELU)
                hipLaunchKernelGGL((k_gemm_nn<64, EPI_BIAS_GELU>), g, dim3(256), 0, s, A, lda, B, ldb,
                                   K, bias, Cf, Cb, ldc);
            else
                hipLaunchKernelGGL((k_gemm_nn<64, EPI_BIAS>), g, dim3(256), 0, s, A, lda, B, ldb, K,
                                   bias, Cf, Cb, ldc);
            return;
        }
    }
    dim3 grid(N / 64, M / 64);
    if (!ta && !tb) {
        if (epi == EPI_BIAS_GELU) GEMM_CASE(false, false, EPI_BIAS_GELU);
        else if (epi == EPI_BIAS) GEMM_CASE(false, false, EPI_BIAS);
        else if (epi == EPI_GELU_GRAD) GEMM_CASE(false, false, EPI_GELU_GRAD);
        else GEMM_CASE(false, false, EPI_STORE);
    } else if (!ta && tb) {
        if (epi == EPI_GELU_GRAD) GEMM_CASE(false, true, EPI_GELU_GRAD);
        else GEMM_CASE(false, true, EPI_STORE);
    } else if (ta && !tb) {
        GEMM_CASE(true, false, EPI_STORE);
    } else {
        GEMM_CASE(true, true, EPI_STORE);
    }
}

static int grid1d(int64_t n, int b) { int64_t g = (n + b - 1) / b; return (int)(g > 4096 ? 4096 : (g < 1 ? 1 : g)); }

void gm_mlp_prep_input(const float *obs, int64_t npix, int nl, int kin, uint16_t *x, hipStream_t s)
{
    hipLaunchKernelGGL(k_prep_input, dim3(grid1d(npix * (kin / 8), 256)), dim3(256), 0, s, obs, npix,
                       nl, kin, x);
}
void gm_mlp_prep_target(const double *models, int64_t npix, float *t, hipStream_t s)
{
    hipLaunchKernelGGL(k_prep_target, dim3(grid1d(npix, 256)), dim3(256), 0, s, models, npix, t);
}
void gm_mlp_decode(const float *y, int64_t npix, const double *base, double *models, hipStream_t s)
{
    hipLaunchKernelGGL(k_decode, dim3(grid1d(npix, 256)), dim3(256), 0, s, y, npix, base, models);
}
void gm_mlp_loss(const float *y, const float *t, int64_t b, uint16_t *dy, float *loss, hipStream_t s)
{
    hipLaunchKernelGGL(k_loss, dim3(grid1d(b * GM_MLP_NOUT, 256)), dim3(256), 0, s, y, t, b, dy, loss);
}
void gm_mlp_bias_grad(const uint16_t *dy, int64_t b, int n, float *db, hipStream_t s)
{
    const int64_t rpb = 64;
    (void)hipMemsetAsync(db, 0, sizeof(float) * n, s);
    hipLaunchKernelGGL(k_bias_grad, dim3((n + 63) / 64, (unsigned)((b + rpb - 1) / rpb)), dim3(256), 0,
                       s, dy, b, n, rpb, db);
}
void gm_mlp_adamw(float *w, const float *g, float *m, float *v, uint16_t *wb, int64_t n, float lr,
                  float b1, float b2, float eps, float wd, float bc1, float bc2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adamw, dim3(grid1d(n, 256)), dim3(256), 0, s, w, g, m, v, wb, n, lr, b1, b2,
                       eps, wd, bc1, bc2);
}
void gm_mlp_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1d(n, 256)), dim3(256), 0, s, x, y, n);
}

void gm_mlp_infer_fused(const float *obs, int64_t npix, int nl, int kin, const uint16_t *wt,
                        const float *w, const gm_mlp_offsets &of, const double *base, double *models,
                        hipStream_t s)
{
    const int64_t tiles = (npix + FX_ROWS - 1) / FX_ROWS;
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)tiles), dim3(256), 0, s, obs, npix, nl, kin, wt, w, of,
                       base, models);
}
void gm_mlp_transpose(const uint16_t *w, int in, int out, uint16_t *wt, hipStream_t s)
{
    hipLaunchKernelGGL(k_transpose_w, dim3(grid1d((int64_t)in * out, 256)), dim3(256), 0, s, w, in, out, wt);
}
