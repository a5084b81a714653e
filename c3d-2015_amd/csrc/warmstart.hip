// warmstart.hip -- the learned warm start of the closure: a multilayer perceptron
// 25 -> 128 -> 128 -> 25 with GELU activations that maps the normalised moments u/u0 of a
// cell to an initial alpha of the u0-scaled dual problem.
//
// Training (FP32, all on the device): tiled GEMMs with fused bias+GELU epilogues, the GELU
// derivative fused into the backward GEMMs, deterministic column and loss reductions (fixed
// order, no atomics), minibatches drawn by a counter RNG, Adam.  Inference: FP16 weights and
// activations on the matrix cores (V_MFMA_F32_16X16X16_F16, FP32 accumulation), 16 cells per
// wavefront, the weights of all three layers resident in LDS.
//
// Parameter vector (FP32, 23065 entries): W1[25][128] b1[128] W2[128][128] b2[128]
// W3[128][25] b3[25]; a layer computes z_j = b_j + sum_i x_i W[i][j].
// The network has no counterpart in the reference; it only proposes a starting point, and
// the closure keeps the caller's alpha whenever the proposal does not lower ||grad h||.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "c3d_internal.h"

namespace c3d {
namespace {

constexpr int kIn = C3D_MLP_IN, kHid = C3D_MLP_HID;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kHid, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid,
              oW3 = oB2 + kHid, oB3 = oW3 + kHid * kIn, kNP = oB3 + kIn;
static_assert(kNP == C3D_MLP_NPARAM, "parameter count");

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// ------------------------------------------------------------------ FP32 GEMM
// C(i,j) = sum_k A(i,k) B(k,j) for i < M, j < N with A(i,k) = A[i sai + k sak],
// B(k,j) = B[k sbk + j sbj], C[i ldc + j].  64x64 tile per 256-thread block, 4x4 per
// thread, K in steps of 16 through LDS; each output is summed in k order by one thread.
// Epilogues: 0 store; 1 Z = acc + bias_j, store Z and gelu(Z) (second array); 2 store
// acc * gelu'(Zin(i,j)) (Zin with the same layout as C).  Split K: block z sums k in
// [z kLen, (z+1) kLen) and stores at C + z M ldc (epilogue 0 only); k_sum_splits adds the
// partial products in z order.
constexpr int kBM = 64, kBN = 64, kBK = 16;

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float *__restrict__ A,
                                              int64_t sai, int64_t sak,
                                              const float *__restrict__ B, int64_t sbk,
                                              int64_t sbj, float *C, int64_t ldc,
                                              const float *__restrict__ bias, float *C2,
                                              const float *__restrict__ Zin, int kLen)
{
    __shared__ float As[kBK][kBM + 4];
    __shared__ float Bs[kBK][kBN + 4];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int i0 = blockIdx.y * kBM, j0 = blockIdx.x * kBN;
    const int kBeg = blockIdx.z * kLen;
    const int kEnd = min(K, kBeg + kLen);
    C += (int64_t)blockIdx.z * M * ldc;
    float acc[4][4] = {};
    for (int k0 = kBeg; k0 < kEnd; k0 += kBK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = tid + 256 * r;  // 1024 = kBK * kBM
            const int kk = e / kBM, ii = e % kBM;
            const int gi = i0 + ii, gk = k0 + kk;
            As[kk][ii] = (gi < M && gk < kEnd) ? A[gi * sai + gk * sak] : 0.0f;
            const int kb = e / kBN, jj = e % kBN;
            const int gj = j0 + jj, gkb = k0 + kb;
            Bs[kb][jj] = (gj < N && gkb < kEnd) ? B[gkb * sbk + gj * sbj] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                a[r] = As[kk][ty * 4 + r];
                b[r] = Bs[kk][tx * 4 + r];
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    acc[r][c] = fmaf(a[r], b[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = i0 + ty * 4 + r;
        if (i >= M)
            continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = j0 + tx * 4 + c;
            if (j >= N)
                continue;
            const int64_t o = (int64_t)i * ldc + j;
            if constexpr (EPI == 0) {
                C[o] = acc[r][c];
            } else if constexpr (EPI == 1) {
                const float z = acc[r][c] + bias[j];
                C[o] = z;
                C2[o] = gelu(z);
            } else {
                C[o] = acc[r][c] * gelu_d(Zin[o]);
            }
        }
    }
}

// C[e] = sum_z P[z][e], z in order (the split-K partial products)
__global__ void k_sum_splits(int n, int splits, const float *__restrict__ Pp, float *C)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z)
        s += Pp[(int64_t)z * n + e];
    C[e] = s;
}

// out[j] = sum_i X[i ld + j], i < rows, in a fixed tree order (one block per column)
__global__ __launch_bounds__(256) void k_colsum(int rows, int ld, const float *__restrict__ X,
                                                float *out)
{
    __shared__ float red[256];
    const int j = blockIdx.x;
    float s = 0.0f;
    for (int i = threadIdx.x; i < rows; i += 256)
        s += X[(int64_t)i * ld + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[j] = red[0];
}

// dY = 2 (Y - T) / (n nm) on outputs k < nm (0 beyond); per-block partial sums of (Y - T)^2
__global__ __launch_bounds__(256) void k_mse(int64_t n, int nm, const float *__restrict__ Y,
                                             const float *__restrict__ T, float *dY,
                                             float *partial)
{
    __shared__ float red[256];
    const float scale = 2.0f / ((float)n * (float)nm);
    float s = 0.0f;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n * kIn;
         e += (int64_t)gridDim.x * 256) {
        const int k = (int)(e % kIn);
        float d = 0.0f;
        if (k < nm) {
            d = Y[e] - T[e];
            s += d * d;
        }
        if (dY)
            dY[e] = d * scale;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = red[0];
}

__global__ void k_finish_loss(int parts, const float *__restrict__ partial, float denom,
                              float *loss)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.0f;
        for (int p = 0; p < parts; ++p)
            s += partial[p];
        *loss = s / denom;
    }
}

__global__ void k_adam(int np, float *p, const float *__restrict__ g, float *m, float *v, float lr,
                       float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Glorot-uniform weights from a counter RNG, zero biases
__global__ void k_init_params(uint64_t seed, float *p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kNP)
        return;
    float a = 0.0f;
    if (i < oB1)
        a = sqrtf(6.0f / (kIn + kHid));
    else if (i >= oW2 && i < oB2)
        a = sqrtf(6.0f / (kHid + kHid));
    else if (i >= oW3 && i < oB3)
        a = sqrtf(6.0f / (kHid + kIn));
    const uint64_t r = mix64(seed ^ mix64((uint64_t)i));
    const float u = (float)((r >> 40) * (1.0 / 16777216.0));  // [0,1)
    p[i] = a * (2.0f * u - 1.0f);
}

// rows of the minibatch: idx = hash(seed, step, i) mod n
__global__ void k_gather(int64_t n, int b, uint64_t seed, uint64_t step,
                         const float *__restrict__ X, const float *__restrict__ T, float *Xb,
                         float *Tb, int64_t *idx)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)b * kIn)
        return;
    const int64_t i = e / kIn;
    const int k = (int)(e % kIn);
    const uint64_t r = mix64(seed ^ mix64(step * 0x100000001B3ull + (uint64_t)i));
    const int64_t src = (int64_t)(r % (uint64_t)n);
    Xb[e] = X[src * kIn + k];
    Tb[e] = T[src * kIn + k];
    if (k == 0)
        idx[i] = src;
}

// features x_k = (u_k/u0 - xm_k)/xs_k and targets t_k = (alpha_k - ym_k)/ys_k of the
// u0-scaled problem (alpha_0 - 2 sqrt(pi) log u0, optBatch.cpp:312-331), k < nm
__global__ void k_features(int64_t n, int nm, const double *__restrict__ u,
                           const double *__restrict__ alpha, const float *__restrict__ norm,
                           float *X, float *T)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * kIn)
        return;
    const int64_t c = e / kIn;
    const int k = (int)(e % kIn);
    float x = 0.0f, t = 0.0f;
    if (k < nm) {
        const double u0 = u[c * nm];
        x = ((float)(u[c * nm + k] / u0) - norm[k]) / norm[kIn + k];
        if (alpha) {
            double a = alpha[c * nm + k];
            if (k == 0)
                a = a - log(u0) * 2.0 * sqrt(M_PI);
            t = ((float)a - norm[2 * kIn + k]) / norm[3 * kIn + k];
        }
    }
    X[e] = x;
    if (T)
        T[e] = t;
}

// ------------------------------------------------------------------ dual-objective loss
// For each row i (one wavefront): alpha = alpha_mean + alpha_std * y_i (FP64; or the pair's
// own alpha when aPair is given), the u0-scaled dual objective f(alpha) = sum_q w_q
// e^{alpha.m_q} - alpha.u~ and its gradient g(alpha) = sum_q w_q e^{alpha.m_q} m_q - u~
// (the moment residual), u~ = u[src]/u0[src].  dY = g * alpha_std * scale (the gradient of
// scale * sum_i f_i with respect to the network's standardised outputs); fout = f - fstar.
// Exponents are capped at 80 so that a wild early prediction stays finite.
template <int NM>
__global__ __launch_bounds__(256) void k_dual_loss(int b, int nqPad, const double *__restrict__ Mg,
                                                   const double *__restrict__ wg,
                                                   const float *__restrict__ Y,
                                                   const double *__restrict__ u,
                                                   const double *__restrict__ aPair,
                                                   const int64_t *__restrict__ idx,
                                                   const float *__restrict__ norm,
                                                   const double *__restrict__ fstar, float scale,
                                                   float *dY, double *fout)
{
    constexpr int NMS = NM | 1;
    __shared__ double sM[64 * NMS];
    __shared__ double sW[64];
    __shared__ double sA[4][NM];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nBlocks = (b + 3) / 4;
    for (int blk = blockIdx.x; blk < nBlocks; blk += gridDim.x) {
        const int row = blk * 4 + wv;
        const bool have = row < b;
        const int64_t src = have ? (idx ? idx[row] : (int64_t)row) : 0;
        const double u0 = u[src * NM];
        if (have && lane < NM) {
            double a;
            if (aPair) {
                a = aPair[src * NM + lane];
                if (lane == 0)
                    a = a - log(u0) * 2.0 * sqrt(M_PI);
            } else {
                a = (double)norm[2 * kIn + lane] + (double)norm[3 * kIn + lane] * (double)Y[(int64_t)row * kIn + lane];
            }
            sA[wv][lane] = a;
        }
        double f = 0.0, g[NM];
#pragma unroll
        for (int i = 0; i < NM; ++i)
            g[i] = 0.0;
        for (int ch = 0; ch < nqPad / 64; ++ch) {
            __syncthreads();
            for (int e = tid; e < 64 * NM; e += 256) {
                const int rr = e / NM, cc = e - rr * NM;
                sM[rr * NMS + cc] = Mg[(size_t)ch * 64 * NM + e];
            }
            if (tid < 64)
                sW[tid] = wg[(size_t)ch * 64 + tid];
            __syncthreads();
            if (have) {
                const double *rowp = &sM[lane * NMS];
                double x = 0.0;
#pragma unroll
                for (int i = 0; i < NM; ++i)
                    x = fma(sA[wv][i], rowp[i], x);
                const double e = sW[lane] * exp(fmin(x, 80.0));
                f += e;
#pragma unroll
                for (int i = 0; i < NM; ++i)
                    g[i] = fma(e, rowp[i], g[i]);
            }
        }
        if (!have)
            continue;
        double fv = 0.0;
        for (int o = 32; o >= 1; o >>= 1)
            f += __shfl_xor(f, o);
        fv = f;
        double gl = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            double t = g[i];
            for (int o = 32; o >= 1; o >>= 1)
                t += __shfl_xor(t, o);
            const double ui = u[src * NM + i] / u0;
            fv -= sA[wv][i] * ui;
            if (lane == i)
                gl = t - ui;
        }
        if (dY && lane < kIn)
            dY[(int64_t)row * kIn + lane] =
                lane < NM ? (float)(gl * (double)norm[3 * kIn + lane]) * scale : 0.0f;
        if (fout && lane == 0)
            fout[row] = fv - (fstar ? fstar[src] : 0.0);
    }
}

__global__ void k_sum_f64(int n, const double *__restrict__ x, double scale, float *out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256)
        s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = (float)(red[0] * scale);
}

// ------------------------------------------------------------------ FP16 MFMA inference
// Operand layout of V_MFMA_F32_16X16X16_F16: lane l holds A[l%16][4(l/16) + e] and
// B[4(l/16) + e][l%16], e = 0..3; accumulator register r holds D[4(l/16) + r][l%16].
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kS1 = 36, kS2 = kHid + 4;  // LDS row strides (halves) of W1^T and W2^T/W3^T/H

struct InferArgs {
    int64_t n;
    int nm;
    const _Float16 *w1t;  // [128][kS1]  W1 transposed, inputs >= 25 zero
    const _Float16 *w2t;  // [128][kS2]
    const _Float16 *w3t;  // [32][kS2]    outputs >= 25 zero
    const float *bias;    // b1[128] b2[128] b3[25]
    const float *norm;    // xm xs ym ys, 25 each
};

__device__ __forceinline__ h4 lds_h4(const _Float16 *p)
{
    return *reinterpret_cast<const h4 *>(p);
}

// X: n cells' standardised features [n][25] (FP32); out: alpha [n][nm] (FP64), u0-scaled.
__global__ __launch_bounds__(256) void k_mlp_infer(InferArgs a, const float *__restrict__ X,
                                                   double *out)
{
    __shared__ __attribute__((aligned(16))) _Float16 sW1[kHid * kS1];
    __shared__ __attribute__((aligned(16))) _Float16 sW2[kHid * kS2];
    __shared__ __attribute__((aligned(16))) _Float16 sW3[32 * kS2];
    __shared__ __attribute__((aligned(16))) _Float16 sH[4][16 * kS2];
    __shared__ float sB[kHid * 2 + 32];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int e = tid; e < kHid * kS1; e += 256)
        sW1[e] = a.w1t[e];
    for (int e = tid; e < kHid * kS2; e += 256)
        sW2[e] = a.w2t[e];
    for (int e = tid; e < 32 * kS2; e += 256)
        sW3[e] = a.w3t[e];
    for (int e = tid; e < kHid * 2 + 32; e += 256)
        sB[e] = e < kHid * 2 + kIn ? a.bias[e] : 0.0f;
    __syncthreads();
    const int col = lane & 15, quad = lane >> 4;
    _Float16 *H = sH[wv];
    const int64_t groups = (a.n + 15) / 16;
    for (int64_t g = (int64_t)blockIdx.x * 4 + wv; g < groups; g += (int64_t)gridDim.x * 4) {
        const int64_t cell = g * 16 + col;  // A-operand row of this lane
        h4 x[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = 16 * s + 4 * quad + e;
                x[s][e] = (_Float16)((cell < a.n && k < kIn) ? X[cell * kIn + k] : 0.0f);
            }
        // layer 1
        f4v acc[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 2; ++s)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(
                    x[s], lds_h4(&sW1[(16 * t + col) * kS1 + 16 * s + 4 * quad]), acc[t], 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                H[(4 * quad + r) * kS2 + 16 * t + col] = (_Float16)gelu(acc[t][r] + sB[16 * t + col]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // layer 2
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 8; ++s)
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(
                    lds_h4(&H[col * kS2 + 16 * s + 4 * quad]),
                    lds_h4(&sW2[(16 * t + col) * kS2 + 16 * s + 4 * quad]), acc[t], 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                H[(4 * quad + r) * kS2 + 16 * t + col] =
                    (_Float16)gelu(acc[t][r] + sB[kHid + 16 * t + col]);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // layer 3 and de-standardisation
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            f4v y = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 8; ++s)
                y = __builtin_amdgcn_mfma_f32_16x16x16f16(
                    lds_h4(&H[col * kS2 + 16 * s + 4 * quad]),
                    lds_h4(&sW3[(16 * t + col) * kS2 + 16 * s + 4 * quad]), y, 0, 0, 0);
            const int j = 16 * t + col;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t c = g * 16 + 4 * quad + r;
                if (c < a.n && j < a.nm)
                    out[c * a.nm + j] =
                        (double)(a.norm[2 * kIn + j] + a.norm[3 * kIn + j] * (y[r] + sB[2 * kHid + j]));
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// FP16 transposed weight images for k_mlp_infer
__global__ void k_pack_f16(const float *__restrict__ p, _Float16 *w1t, _Float16 *w2t,
                           _Float16 *w3t, float *bias)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kHid * kS1) {
        const int j = e / kS1, k = e % kS1;
        w1t[e] = (_Float16)(k < kIn ? p[oW1 + k * kHid + j] : 0.0f);
    }
    if (e < kHid * kS2) {
        const int j = e / kS2, k = e % kS2;
        w2t[e] = (_Float16)(k < kHid ? p[oW2 + k * kHid + j] : 0.0f);
    }
    if (e < 32 * kS2) {
        const int j = e / kS2, k = e % kS2;
        w3t[e] = (_Float16)((k < kHid && j < kIn) ? p[oW3 + k * kIn + j] : 0.0f);
    }
    if (e < kHid)
        bias[e] = p[oB1 + e];
    if (e < kHid)
        bias[kHid + e] = p[oB2 + e];
    if (e < kIn)
        bias[2 * kHid + e] = p[oB3 + e];
}

template <int EPI>
void gemm(hipStream_t s, int M, int N, int K, const float *A, int64_t sai, int64_t sak,
          const float *B, int64_t sbk, int64_t sbj, float *C, int64_t ldc,
          const float *bias = nullptr, float *C2 = nullptr, const float *Zin = nullptr)
{
    dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM);
    hipLaunchKernelGGL(k_gemm<EPI>, grid, dim3(256), 0, s, M, N, K, A, sai, sak, B, sbk, sbj, C,
                       ldc, bias, C2, Zin, K);
}

// C[M][N] (ldc = N) with the reduction over K split into pieces of kLen (weight gradients:
// small M, N and K = the batch); partial products in Pp [splits][M][N]
constexpr int kSplitLen = 256, kMaxSplits = 64;
void gemm_splitk(hipStream_t s, int M, int N, int K, const float *A, int64_t sai, int64_t sak,
                 const float *B, int64_t sbk, int64_t sbj, float *C, float *Pp)
{
    const int splits = std::min(kMaxSplits, std::max(1, (K + kSplitLen - 1) / kSplitLen));
    const int kLen = (K + splits - 1) / splits;
    dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM, splits);
    hipLaunchKernelGGL(k_gemm<0>, grid, dim3(256), 0, s, M, N, K, A, sai, sak, B, sbk, sbj, Pp,
                       (int64_t)N, nullptr, nullptr, nullptr, kLen);
    hipLaunchKernelGGL(k_sum_splits, dim3((M * N + 255) / 256), dim3(256), 0, s, M * N, splits,
                       Pp, C);
}

// Workspace of one forward/backward pass over b rows
struct Acts {
    float *Z1, *A1, *Z2, *A2, *Y, *dY, *dA, *dB, *part, *loss, *splitk;
};

size_t acts_floats(int64_t b)
{
    return (size_t)b * (4 * kHid + 2 * kIn + 2 * kHid) + 1024 + 1 + (size_t)kMaxSplits * kHid * kHid;
}

Acts carve(float *w, int64_t b)
{
    Acts a;
    a.Z1 = w;
    a.A1 = a.Z1 + b * kHid;
    a.Z2 = a.A1 + b * kHid;
    a.A2 = a.Z2 + b * kHid;
    a.Y = a.A2 + b * kHid;
    a.dY = a.Y + b * kIn;
    a.dA = a.dY + b * kIn;
    a.dB = a.dA + b * kHid;
    a.part = a.dB + b * kHid;
    a.loss = a.part + 1024;
    a.splitk = a.loss + 1;
    return a;
}

void forward(hipStream_t s, int64_t b, const float *p, const float *X, const Acts &a)
{
    const int M = (int)b;
    gemm<1>(s, M, kHid, kIn, X, kIn, 1, p + oW1, kHid, 1, a.Z1, kHid, p + oB1, a.A1);
    gemm<1>(s, M, kHid, kHid, a.A1, kHid, 1, p + oW2, kHid, 1, a.Z2, kHid, p + oB2, a.A2);
    gemm<0>(s, M, kIn, kHid, a.A2, kHid, 1, p + oW3, kIn, 1, a.Y, kIn);
    // + b3: folded in by a rank-1 update through the GEMM would cost a launch as well
}

__global__ void k_add_bias(int64_t rows, int cols, float *Y, const float *__restrict__ b)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < rows * cols)
        Y[e] += b[e % cols];
}

void forward_full(hipStream_t s, int64_t b, const float *p, const float *X, const Acts &a)
{
    forward(s, b, p, X, a);
    hipLaunchKernelGGL(k_add_bias, dim3((unsigned)((b * kIn + 255) / 256)), dim3(256), 0, s, b,
                       kIn, a.Y, p + oB3);
}

// parameter gradient from dY (after forward_full on the same rows)
void backward(hipStream_t s, int64_t b, const float *p, const float *X, float *grad,
              const Acts &a)
{
    const int M = (int)b;
    // layer 3
    gemm_splitk(s, kHid, kIn, M, a.A2, 1, kHid, a.dY, kIn, 1, grad + oW3, a.splitk);
    hipLaunchKernelGGL(k_colsum, dim3(kIn), dim3(256), 0, s, M, kIn, a.dY, grad + oB3);
    gemm<2>(s, M, kHid, kIn, a.dY, kIn, 1, p + oW3, 1, kIn, a.dA, kHid, nullptr, nullptr, a.Z2);
    // layer 2 (dA holds dZ2)
    gemm_splitk(s, kHid, kHid, M, a.A1, 1, kHid, a.dA, kHid, 1, grad + oW2, a.splitk);
    hipLaunchKernelGGL(k_colsum, dim3(kHid), dim3(256), 0, s, M, kHid, a.dA, grad + oB2);
    gemm<2>(s, M, kHid, kHid, a.dA, kHid, 1, p + oW2, 1, kHid, a.dB, kHid, nullptr, nullptr, a.Z1);
    // layer 1 (dB holds dZ1)
    gemm_splitk(s, kIn, kHid, M, X, 1, kIn, a.dB, kHid, 1, grad + oW1, a.splitk);
    hipLaunchKernelGGL(k_colsum, dim3(kHid), dim3(256), 0, s, M, kHid, a.dB, grad + oB1);
}

// loss (device scalar) and gradient of the mean squared error over outputs k < nm
void loss_grad(hipStream_t s, int64_t b, int nm, const float *p, const float *X, const float *T,
               float *grad, const Acts &a)
{
    forward_full(s, b, p, X, a);
    const int parts = (int)std::min<int64_t>(1024, (b * kIn + 255) / 256);
    hipLaunchKernelGGL(k_mse, dim3(parts), dim3(256), 0, s, b, nm, a.Y, T, a.dY, a.part);
    hipLaunchKernelGGL(k_finish_loss, dim3(1), dim3(64), 0, s, parts, a.part,
                       (float)b * (float)nm, a.loss);
    backward(s, b, p, X, grad, a);
}

// the dual-objective loss: mean over rows of f(alpha(net(X))) - fstar and its gradient
hipError_t dual_loss_grad(c3d_ctx_s *ctx, hipStream_t s, int64_t b, const float *p,
                          const float *X, const double *u, const int64_t *idx,
                          const float *normDev, const double *fstar, float *grad, double *fbuf,
                          const Acts &a)
{
    forward_full(s, b, p, X, a);
    const unsigned grid = (unsigned)std::min<int64_t>((b + 3) / 4, (int64_t)ctx->numCU * 8);
    const float scale = 1.0f / (float)b;
#define C3D_DL(NMV) hipLaunchKernelGGL(k_dual_loss<NMV>, dim3(grid), dim3(256), 0, s, (int)b, ctx->nqPad, ctx->d_M, ctx->d_wpad, a.Y, u, nullptr, idx, normDev, fstar, scale, a.dY, fbuf)
    switch (ctx->t.nm) {
    case 4: C3D_DL(4); break;
    case 9: C3D_DL(9); break;
    case 16: C3D_DL(16); break;
    case 25: C3D_DL(25); break;
    default: return hipErrorInvalidValue;
    }
#undef C3D_DL
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(256), 0, s, (int)b, fbuf, 1.0 / (double)b, a.loss);
    backward(s, b, p, X, grad, a);
    return hipGetLastError();
}

// fstar[i] = f(alpha_i) of the training pairs themselves
hipError_t pair_fstar(c3d_ctx_s *ctx, hipStream_t s, int64_t n, const double *u,
                      const double *alpha, double *fstar)
{
    const unsigned grid = (unsigned)std::min<int64_t>((n + 3) / 4, (int64_t)ctx->numCU * 8);
#define C3D_FS(NMV) hipLaunchKernelGGL(k_dual_loss<NMV>, dim3(grid), dim3(256), 0, s, (int)n, ctx->nqPad, ctx->d_M, ctx->d_wpad, nullptr, u, alpha, nullptr, nullptr, nullptr, 0.0f, nullptr, fstar)
    switch (ctx->t.nm) {
    case 4: C3D_FS(4); break;
    case 9: C3D_FS(9); break;
    case 16: C3D_FS(16); break;
    case 25: C3D_FS(25); break;
    default: return hipErrorInvalidValue;
    }
#undef C3D_FS
    return hipGetLastError();
}

}  // namespace

// ------------------------------------------------------------------ host entry points
hipError_t mlp_pack(c3d_ctx_s *ctx, hipStream_t s)
{
    Mlp &m = ctx->mlp;
    hipError_t e = grow(m.w16, sizeof(_Float16) * (kHid * kS1 + kHid * kS2 + 32 * kS2));
    if (e == hipSuccess)
        e = grow(m.bias, sizeof(float) * (2 * kHid + kIn));
    if (e != hipSuccess)
        return e;
    _Float16 *w1t = (_Float16 *)m.w16.p, *w2t = w1t + kHid * kS1, *w3t = w2t + kHid * kS2;
    hipLaunchKernelGGL(k_pack_f16, dim3((kHid * kS2 + 255) / 256), dim3(256), 0, s,
                       (const float *)m.params.p, w1t, w2t, w3t, (float *)m.bias.p);
    return hipGetLastError();
}

hipError_t mlp_infer(c3d_ctx_s *ctx, int64_t n, const float *X, double *out, hipStream_t s)
{
    Mlp &m = ctx->mlp;
    if (n <= 0)
        return hipSuccess;
    InferArgs a;
    a.n = n;
    a.nm = ctx->t.nm;
    a.w1t = (const _Float16 *)m.w16.p;
    a.w2t = a.w1t + kHid * kS1;
    a.w3t = a.w2t + kHid * kS2;
    a.bias = (const float *)m.bias.p;
    a.norm = (const float *)m.normDev.p;
    const int64_t groups = (n + 15) / 16;
    const unsigned grid = (unsigned)std::min<int64_t>((groups + 3) / 4, (int64_t)ctx->numCU * 2);
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(256), 0, s, a, X, out);
    return hipGetLastError();
}

hipError_t mlp_features(c3d_ctx_s *ctx, int64_t n, const double *u, const double *alpha,
                        const float *normDev, float *X, float *T, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((n * kIn + 255) / 256)), dim3(256), 0, s, n,
                       ctx->t.nm, u, alpha, normDev, X, T);
    return hipGetLastError();
}

hipError_t mlp_forward(c3d_ctx_s *ctx, int64_t n, const float *p, const float *X, float *Y,
                       hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipError_t e = grow(ctx->mlp.work, acts_floats(n) * sizeof(float));
    if (e != hipSuccess)
        return e;
    Acts a = carve((float *)ctx->mlp.work.p, n);
    a.Y = Y;
    forward_full(s, n, p, X, a);
    return hipGetLastError();
}

hipError_t mlp_loss_grad(c3d_ctx_s *ctx, int64_t n, const float *p, const float *X,
                         const float *T, float *grad, float *loss, hipStream_t s)
{
    if (n <= 0)
        return hipErrorInvalidValue;
    hipError_t e = grow(ctx->mlp.work, acts_floats(n) * sizeof(float));
    if (e != hipSuccess)
        return e;
    Acts a = carve((float *)ctx->mlp.work.p, n);
    loss_grad(s, n, ctx->t.nm, p, X, T, grad, a);
    return hipMemcpyAsync(loss, a.loss, sizeof(float), hipMemcpyDeviceToDevice, s);
}

// Normalisation from a sample of (u, alpha) (host statistics in FP64)
static void fit_norm(int nm, int64_t n, const double *uh, const double *ah, float *norm)
{
    for (int k = 0; k < kIn; ++k) {
        norm[k] = 0.0f;
        norm[kIn + k] = 1.0f;
        norm[2 * kIn + k] = 0.0f;
        norm[3 * kIn + k] = 1.0f;
    }
    for (int k = 0; k < nm; ++k) {
        double sx = 0, sxx = 0, sy = 0, syy = 0;
        for (int64_t c = 0; c < n; ++c) {
            const double u0 = uh[c * nm];
            const double x = uh[c * nm + k] / u0;
            double y = ah[c * nm + k];
            if (k == 0)
                y -= log(u0) * 2.0 * sqrt(M_PI);
            sx += x;
            sxx += x * x;
            sy += y;
            syy += y * y;
        }
        const double mx = sx / n, my = sy / n;
        const double vx = std::max(sxx / n - mx * mx, 0.0), vy = std::max(syy / n - my * my, 0.0);
        norm[k] = (float)mx;
        norm[kIn + k] = (float)(vx > 1e-12 ? sqrt(vx) : 1.0);
        norm[2 * kIn + k] = (float)my;
        norm[3 * kIn + k] = (float)(vy > 1e-12 ? sqrt(vy) : 1.0);
    }
}

int mlp_train(c3d_ctx_s *ctx, int64_t n, const double *u, const double *alpha, int mseSteps,
              int dualSteps, int batch, float lr, uint64_t seed, float *lossFirst,
              float *lossLast, hipStream_t s)
{
    Mlp &m = ctx->mlp;
    const int nm = ctx->t.nm;
    // normalisation from the first (up to) 65536 pairs
    const int64_t ns = std::min<int64_t>(n, 65536);
    std::vector<double> uh(ns * nm), ah(ns * nm);
    if (hipMemcpyAsync(uh.data(), u, uh.size() * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(ah.data(), alpha, ah.size() * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return C3D_ERR_HIP;
    fit_norm(nm, ns, uh.data(), ah.data(), m.normHost);
    hipError_t e = grow(m.normDev, sizeof(m.normHost));
    if (e == hipSuccess)
        e = hipMemcpyAsync(m.normDev.p, m.normHost, sizeof(m.normHost), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = grow(m.params, kNP * sizeof(float));
    if (e == hipSuccess)
        e = grow(m.opt, 3 * (size_t)kNP * sizeof(float));  // grad, m, v
    // features, targets, f* per pair; minibatch rows, their indices and objectives
    const size_t xb = (size_t)n * kIn * sizeof(float);
    const size_t bytes = 2 * xb + (size_t)n * sizeof(double) +
                         (size_t)batch * (2 * kIn * sizeof(float) + sizeof(int64_t) + sizeof(double));
    if (e == hipSuccess)
        e = grow(m.data, bytes);
    if (e == hipSuccess)
        e = grow(m.work, acts_floats(batch) * sizeof(float));
    if (e != hipSuccess)
        return C3D_ERR_HIP;
    float *P = (float *)m.params.p, *G = (float *)m.opt.p, *Mo = G + kNP, *Vo = Mo + kNP;
    char *base = (char *)m.data.p;
    float *X = (float *)base, *T = (float *)(base + xb);
    double *fstar = (double *)(base + 2 * xb);
    float *Xb = (float *)(fstar + n), *Tb = Xb + (size_t)batch * kIn;
    int64_t *idx = (int64_t *)(Tb + (size_t)batch * kIn);
    double *fbuf = (double *)(idx + batch);
    e = mlp_features(ctx, n, u, alpha, (const float *)m.normDev.p, X, T, s);
    if (e == hipSuccess && dualSteps > 0)
        e = pair_fstar(ctx, s, n, u, alpha, fstar);
    hipLaunchKernelGGL(k_init_params, dim3((kNP + 255) / 256), dim3(256), 0, s, seed, P);
    if (e != hipSuccess)
        return C3D_ERR_HIP;
    Acts a = carve((float *)m.work.p, batch);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    // phase 1: mean squared error on the standardised alpha; phase 2: the mean dual gap
    // f(alpha_net) - f(alpha*), whose gradient with respect to alpha is the moment residual
    for (int phase = 0; phase < 2; ++phase) {
        const int steps = phase ? dualSteps : mseSteps;
        if (steps <= 0)
            continue;
        const float lr0 = phase ? 0.3f * lr : lr;
        if (hipMemsetAsync(Mo, 0, 2 * (size_t)kNP * sizeof(float), s) != hipSuccess)
            return C3D_ERR_HIP;
        float c1 = 1.0f, c2 = 1.0f;
        for (int it = 0; it < steps; ++it) {
            const uint64_t key = (uint64_t)phase * 0x9E3779B9ull + (uint64_t)it;
            hipLaunchKernelGGL(k_gather, dim3((unsigned)(((int64_t)batch * kIn + 255) / 256)),
                               dim3(256), 0, s, n, batch, seed, key, X, T, Xb, Tb, idx);
            if (phase == 0)
                loss_grad(s, batch, nm, P, Xb, Tb, G, a);
            else if ((e = dual_loss_grad(ctx, s, batch, P, Xb, u, idx,
                                         (const float *)m.normDev.p, fstar, G, fbuf, a)) !=
                     hipSuccess)
                return C3D_ERR_HIP;
            if (it == 0 && lossFirst && (phase == 0 || mseSteps <= 0))
                (void)hipMemcpyAsync(lossFirst, a.loss, sizeof(float), hipMemcpyDeviceToHost, s);
            c1 *= b1;
            c2 *= b2;
            // cosine decay of the step size over the phase
            const float lrt = lr0 * 0.5f * (1.0f + cosf(3.14159265f * (float)it / (float)steps));
            hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, s, kNP, P, G, Mo,
                               Vo, lrt, b1, b2, eps, 1.0f - c1, 1.0f - c2);
        }
    }
    if (lossLast)
        (void)hipMemcpyAsync(lossLast, a.loss, sizeof(float), hipMemcpyDeviceToHost, s);
    e = mlp_pack(ctx, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return C3D_ERR_HIP;
    m.have = true;
    return C3D_OK;
}

// the dual-objective loss and gradient on given rows (for checking the kernels)
hipError_t mlp_dual_loss_grad(c3d_ctx_s *ctx, int64_t n, const float *p, const float *X,
                              const double *u, const float *normDev, float *grad, float *loss,
                              hipStream_t s)
{
    if (n <= 0)
        return hipErrorInvalidValue;
    hipError_t e = grow(ctx->mlp.work, acts_floats(n) * sizeof(float) + n * sizeof(double));
    if (e != hipSuccess)
        return e;
    Acts a = carve((float *)ctx->mlp.work.p, n);
    double *fbuf = (double *)((char *)ctx->mlp.work.p + acts_floats(n) * sizeof(float));
    e = dual_loss_grad(ctx, s, n, p, X, u, nullptr, normDev, nullptr, grad, fbuf, a);
    if (e != hipSuccess)
        return e;
    return hipMemcpyAsync(loss, a.loss, sizeof(float), hipMemcpyDeviceToDevice, s);
}

int mlp_set(c3d_ctx_s *ctx, const float *params, const float *norm)
{
    Mlp &m = ctx->mlp;
    std::memcpy(m.normHost, norm, sizeof(m.normHost));
    hipError_t e = grow(m.normDev, sizeof(m.normHost));
    if (e == hipSuccess)
        e = hipMemcpy(m.normDev.p, m.normHost, sizeof(m.normHost), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = grow(m.params, kNP * sizeof(float));
    if (e == hipSuccess)
        e = hipMemcpy(m.params.p, params, kNP * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = mlp_pack(ctx, nullptr);
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (e != hipSuccess)
        return C3D_ERR_HIP;
    m.have = true;
    return C3D_OK;
}

int mlp_get(c3d_ctx_s *ctx, float *params, float *norm)
{
    Mlp &m = ctx->mlp;
    if (!m.have)
        return C3D_ERR_ARG;
    if (params && hipMemcpy(params, m.params.p, kNP * sizeof(float), hipMemcpyDeviceToHost) !=
                      hipSuccess)
        return C3D_ERR_HIP;
    if (norm)
        std::memcpy(norm, m.normHost, sizeof(m.normHost));
    return C3D_OK;
}

}  // namespace c3d
