// Training and inference kernels of the learned preconditioner.
//
// The network (feature vector of one parameter point -> diagonal and
// low-rank factor of a preconditioner for the metric solve) is a stack of
// dense layers with BF16 weights, BF16 MFMA operands and FP32 accumulation
// (v_mfma_f32_32x32x16_bf16); activations, gradients and the Adam state are
// FP32; the Kaporin loss and its gradient per sample are FP64.  QGTL has no
// learned component: these kernels precondition the conjugate-gradient
// solve of compute_natural_gradient_cpu
// (src/quantum_geometric/hybrid/classical_optimization_engine.c:543-610).
#include <math.h>

#include "qgtl_hip_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr int kTM = 64, kTN = 64, kTK = 32;
constexpr int kLdsRow = kTK + 8;                 // bf16 per LDS row: 80 B, 16-B aligned

__device__ __forceinline__ uint16_t bf16_rn(float x) {
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float load_op(const void *p, int bf16, size_t i) {
    if (bf16) return __uint_as_float((uint32_t)reinterpret_cast<const uint16_t *>(p)[i] << 16);
    return reinterpret_cast<const float *>(p)[i];
}

// C (M x N, fp32, row-major) = op(A) op(B) [+ bias] [+ C]:
//   op(A)[m][k] = ta ? A[k][m] : A[m][k]  (A row-major with lda columns)
//   op(B)[k][n] = tb ? B[n][k] : B[k][n]
// Operands are rounded to bf16 when staged into LDS.  One 64 x 64 tile per
// workgroup, one 32 x 32 MFMA block per wave.
__global__ __launch_bounds__(kThreads) void gemm_bf16_kernel(int M, int N, int K, const void *__restrict__ A, int a_bf16,
                                                             int ta, int lda, const void *__restrict__ B, int b_bf16,
                                                             int tb, int ldb, float *C, int ldc,
                                                             const float *__restrict__ bias, int accumulate) {
    __shared__ uint16_t As[kTM * kLdsRow];   // [m][k]
    __shared__ uint16_t Bs[kTN * kLdsRow];   // [n][k]
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.f;

    for (int k0 = 0; k0 < K; k0 += kTK) {
        for (int i = threadIdx.x; i < kTM * kTK; i += kThreads) {
            const int mm = i / kTK, kk = i % kTK;
            const int m = m0 + mm, k = k0 + kk;
            float v = 0.f;
            if (m < M && k < K) v = load_op(A, a_bf16, ta ? (size_t)k * lda + m : (size_t)m * lda + k);
            As[mm * kLdsRow + kk] = bf16_rn(v);
            const int nn = i / kTK;
            const int n = n0 + nn;
            float w = 0.f;
            if (n < N && k < K) w = load_op(B, b_bf16, tb ? (size_t)n * ldb + k : (size_t)k * ldb + n);
            Bs[nn * kLdsRow + kk] = bf16_rn(w);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            const int koff = kk + 8 * (lane >> 5);
            const bf16x8 fa = *reinterpret_cast<const bf16x8 *>(&As[(wm * 32 + (lane & 31)) * kLdsRow + koff]);
            const bf16x8 fb = *reinterpret_cast<const bf16x8 *>(&Bs[(wn * 32 + (lane & 31)) * kLdsRow + koff]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int row = m0 + wm * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
        const int col = n0 + wn * 32 + (lane & 31);
        if (row < M && col < N) {
            float v = acc[e];
            if (bias) v += bias[col];
            float *c = C + (size_t)row * ldc + col;
            *c = accumulate ? *c + v : v;
        }
    }
}

constexpr float kGeluC = 0.7978845608028654f;   // sqrt(2 / pi)

__global__ void gelu_fwd_kernel(const float *__restrict__ z, float *a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = z[i];
        a[i] = 0.5f * x * (1.f + tanhf(kGeluC * (x + 0.044715f * x * x * x)));
    }
}

// dz = da * g'(z), in place on da
__global__ void gelu_bwd_kernel(const float *__restrict__ z, float *da, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float x = z[i];
        const float t = tanhf(kGeluC * (x + 0.044715f * x * x * x));
        const float g = 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * kGeluC * (1.f + 3.f * 0.044715f * x * x);
        da[i] *= g;
    }
}

// out[j] = sum_i X[i][j] (rows in a fixed order)
__global__ void colsum_kernel(const float *__restrict__ X, int rows, int cols, float *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int i = 0; i < rows; i++) s += X[(size_t)i * cols + j];
    out[j] = s;
}

__global__ void gather_rows_kernel(const float *__restrict__ src, const int *__restrict__ idx, int rows, int cols,
                                   float *dst) {
    const size_t total = (size_t)rows * cols;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t r = e / cols, c = e % cols;
        dst[e] = src[(size_t)idx[r] * cols + c];
    }
}

// Adam on the FP32 master copy; the BF16 mirror is what the GEMMs read.
__global__ void adam_kernel(float *w, uint16_t *w_bf16, float *m, float *v, const float *__restrict__ g, size_t n,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        w_bf16[i] = bf16_rn(wi);
    }
}

__global__ void to_bf16_kernel(const float *__restrict__ w, uint16_t *out, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = bf16_rn(w[i]);
}

constexpr int kMaxRank = 32;
constexpr int kMaxP = 1024;
constexpr int kMaxPR = 8192;    // P * r: the low-rank factor held in LDS as the network's fp32 outputs (32 KB)

__device__ double block_sum(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < kThreads / 64; w++) s += red[w];
    return s;
}

// Columns jb .. jb + 7 of the low-rank factor are handled together, in registers (compile-time indices).
constexpr int kJB = 8;

// (A_s u)_k,j for row k and the column block j = jb + i, i < kJB (zero past r), FP64; u_lj = yu[l r + j] / sqrt(P)
__device__ __forceinline__ void row_times_u(const float *__restrict__ As, const float *yu, int P, int r, int k, int jb,
                                            double us, double (&acc)[kJB]) {
#pragma unroll
    for (int i = 0; i < kJB; i++) acc[i] = 0.0;
    for (int l = 0; l < P; l++) {
        const double a = As[(size_t)k * P + l];
#pragma unroll
        for (int i = 0; i < kJB; i++)
            if (jb + i < r) acc[i] += a * ((double)yu[l * r + jb + i] * us);
    }
}

// (u W^-1)_k,j for the same column block
__device__ __forceinline__ void row_times_winv(const float *yu, const double *Wi, int r, int k, int jb, double us,
                                               double (&uw)[kJB]) {
#pragma unroll
    for (int i = 0; i < kJB; i++) uw[i] = 0.0;
    for (int a = 0; a < r; a++) {
        const double ua = (double)yu[k * r + a] * us;
#pragma unroll
        for (int i = 0; i < kJB; i++)
            if (jb + i < r) uw[i] += ua * Wi[a * r + jb + i];
    }
}

// Kaporin loss of sample b and dL/dy (scaled by `scale`), FP64.
//   s_k = y[k(1+r)], u_kj = y[k(1+r)+1+j] / sqrt(P), Mt = diag(e^s) + u u^T,
//   L = P log(tr(Mt As)/P) - sum s - log det(I + u^T e^-s u).
// The low-rank outputs stay fp32 in LDS (each use scales them by 1/sqrt(P) in FP64, the value an FP64 copy
// of u would hold).  A_s u is formed twice, for the trace and again for the gradient, eight columns at a
// time in registers, rather than held (P r doubles do not fit beside u at rank 32).
__global__ __launch_bounds__(kThreads) void kaporin_kernel(int P, int r, const float *__restrict__ Y,
                                                           const float *__restrict__ As_all, const int *__restrict__ idx,
                                                           double *loss, float *dY, float scale) {
    const int b = blockIdx.x;
    const float *y = Y + (size_t)b * P * (1 + r);
    const float *As = As_all + (size_t)idx[b] * P * P;
    float *dy = dY + (size_t)b * P * (1 + r);
    __shared__ float yu[kMaxPR];
    __shared__ double W[kMaxRank * kMaxRank], Wi[kMaxRank * kMaxRank];
    __shared__ double red[kThreads / 64];
    __shared__ double sh_logdet;
    const double us = 1.0 / sqrt((double)P);
    const int st = 1 + r;
    for (int e = threadIdx.x; e < P * r; e += kThreads) yu[e] = y[(size_t)(e / r) * st + 1 + e % r];
    __syncthreads();
    double trp = 0.0, ssum = 0.0;
    for (int k = threadIdx.x; k < P; k += kThreads) {
        const double s = y[(size_t)k * st];
        ssum += s;
        trp += exp(s) * (double)As[(size_t)k * P + k];
        for (int jb = 0; jb < r; jb += kJB) {
            double acc[kJB];
            row_times_u(As, yu, P, r, k, jb, us, acc);
#pragma unroll
            for (int i = 0; i < kJB; i++)
                if (jb + i < r) trp += ((double)yu[k * r + jb + i] * us) * acc[i];
        }
    }
    const double tr = block_sum(trp, red);
    const double S = block_sum(ssum, red);
    for (int e = threadIdx.x; e < r * r; e += kThreads) {
        const int a = e / r, c = e % r;
        double v = (a == c) ? 1.0 : 0.0;
        for (int k = 0; k < P; k++)
            v += exp(-(double)y[(size_t)k * st]) * ((double)yu[k * r + a] * us) * ((double)yu[k * r + c] * us);
        W[e] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ld = 0.0;
        for (int j = 0; j < r; j++) {   // Cholesky, lower, in place
            double d = W[j * r + j];
            for (int k = 0; k < j; k++) d -= W[j * r + k] * W[j * r + k];
            d = sqrt(d);
            W[j * r + j] = d;
            ld += 2.0 * log(d);
            for (int i = j + 1; i < r; i++) {
                double v = W[i * r + j];
                for (int k = 0; k < j; k++) v -= W[i * r + k] * W[j * r + k];
                W[i * r + j] = v / d;
            }
        }
        for (int c = 0; c < r; c++) {    // W^-1 column c (symmetric, stored as row c)
            double *x = Wi + c * r;
            for (int i = 0; i < r; i++) {
                double v = (i == c) ? 1.0 : 0.0;
                for (int k = 0; k < i; k++) v -= W[i * r + k] * x[k];
                x[i] = v / W[i * r + i];
            }
            for (int i = r - 1; i >= 0; i--) {
                double v = x[i];
                for (int k = i + 1; k < r; k++) v -= W[k * r + i] * x[k];
                x[i] = v / W[i * r + i];
            }
        }
        sh_logdet = ld;
    }
    __syncthreads();
    if (threadIdx.x == 0) loss[b] = (double)P * log(tr / P) - S - sh_logdet;
    const double g = (double)P / tr;
    for (int k = threadIdx.x; k < P; k += kThreads) {
        const double s = y[(size_t)k * st], ems = exp(-s);
        double q = 0.0;
        for (int jb = 0; jb < r; jb += kJB) {
            double uw[kJB];
            row_times_winv(yu, Wi, r, k, jb, us, uw);
#pragma unroll
            for (int i = 0; i < kJB; i++)
                if (jb + i < r) q += uw[i] * ((double)yu[k * r + jb + i] * us);
        }
        dy[(size_t)k * st] = (float)(scale * (g * exp(s) * (double)As[(size_t)k * P + k] - 1.0 + ems * q));
        for (int jb = 0; jb < r; jb += kJB) {
            double acc[kJB], uw[kJB];
            row_times_u(As, yu, P, r, k, jb, us, acc);
            row_times_winv(yu, Wi, r, k, jb, us, uw);
#pragma unroll
            for (int i = 0; i < kJB; i++)
                if (jb + i < r)
                    dy[(size_t)k * st + 1 + jb + i] = (float)(scale * (2.0 * g * acc[i] - 2.0 * ems * uw[i]) * us);
        }
    }
}

unsigned grid_1d(size_t n) {
    size_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return (unsigned)g;
}

}  // namespace

extern "C" {

int qgtl_hip_mlp_gemm(int M, int N, int K, const void *A, int a_bf16, int ta, int lda, const void *B, int b_bf16,
                      int tb, int ldb, float *C, int ldc, const float *bias, int accumulate) {
    if (M < 1 || N < 1 || K < 1) return qgtl::set_error_msg(-118, "mlp_gemm: empty shape");
    const dim3 grid((unsigned)((N + kTN - 1) / kTN), (unsigned)((M + kTM - 1) / kTM));
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(kThreads), 0, qgtl::stream(), M, N, K, A, a_bf16, ta, lda, B,
                       b_bf16, tb, ldb, C, ldc, bias, accumulate);
    QGTL_LAUNCH_CHECK("mlp_gemm");
    return 0;
}

int qgtl_hip_gelu_fwd(const float *z, float *a, size_t n) {
    hipLaunchKernelGGL(gelu_fwd_kernel, dim3(grid_1d(n)), dim3(256), 0, qgtl::stream(), z, a, n);
    QGTL_LAUNCH_CHECK("gelu_fwd");
    return 0;
}

int qgtl_hip_gelu_bwd(const float *z, float *da, size_t n) {
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(grid_1d(n)), dim3(256), 0, qgtl::stream(), z, da, n);
    QGTL_LAUNCH_CHECK("gelu_bwd");
    return 0;
}

int qgtl_hip_colsum(const float *X, int rows, int cols, float *out) {
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, qgtl::stream(), X, rows,
                       cols, out);
    QGTL_LAUNCH_CHECK("colsum");
    return 0;
}

int qgtl_hip_gather_rows(const float *src, const int *idx, int rows, int cols, float *dst) {
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_1d((size_t)rows * cols)), dim3(256), 0, qgtl::stream(), src,
                       idx, rows, cols, dst);
    QGTL_LAUNCH_CHECK("gather_rows");
    return 0;
}

int qgtl_hip_adam(float *w, uint16_t *w_bf16, float *m, float *v, const float *g, size_t n, float lr, float b1,
                  float b2, float eps, int step) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_1d(n)), dim3(256), 0, qgtl::stream(), w, w_bf16, m, v, g, n, lr, b1,
                       b2, eps, c1, c2);
    QGTL_LAUNCH_CHECK("adam");
    return 0;
}

int qgtl_hip_to_bf16(const float *w, uint16_t *out, size_t n) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_1d(n)), dim3(256), 0, qgtl::stream(), w, out, n);
    QGTL_LAUNCH_CHECK("to_bf16");
    return 0;
}

int qgtl_hip_kaporin(int P, int r, int B, const float *Y, const float *As_all, const int *idx, double *loss,
                     float *dY, float scale) {
    if (P < 1 || P > kMaxP || r < 0 || r > kMaxRank || P * r > kMaxPR || B < 1)
        return qgtl::set_error_msg(-118, "kaporin: needs P <= 1024, r <= 32, P r <= 8192");
    hipLaunchKernelGGL(kaporin_kernel, dim3((unsigned)B), dim3(kThreads), 0, qgtl::stream(), P, r, Y, As_all, idx,
                       loss, dY, scale);
    QGTL_LAUNCH_CHECK("kaporin");
    return 0;
}

}  // extern "C"
