// Whole-vector operations on complex64 amplitude vectors of any length:
// FP64 inner products and norms, scaling, diagonal expectation values and a
// dense row-major complex matrix-vector product.  They back the device side
// of QGTL's quantum_state_* container (quantum_state.h:93-98 and the state
// functions it keeps on the device) and the HIP ComputeBackendOps binding
// (compute_backend.h:179-263).
//
// Reductions are deterministic: a grid-stride pass writes one FP64 partial
// per workgroup (a fixed grid for a given length), and a single workgroup
// adds the partials in a fixed order.
#include "qgtl_hip_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxParts = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum a value over the workgroup in a fixed order; the result in thread 0.
__device__ __forceinline__ void block_sums4(double s[4], double (*red)[4]) {
    for (int j = 0; j < 4; j++) s[j] = wave_sum(s[j]);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < 4; j++) red[wave][j] = s[j];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kThreads / 64; w++)
            for (int j = 0; j < 4; j++) red[0][j] += red[w][j];
}

// partials[4 g + 0..3] = Re <a|b>, Im <a|b>, |a|^2, |b|^2 over the elements
// of workgroup g's grid-stride range (b may equal a).
__global__ __launch_bounds__(kThreads) void dotc_partials_kernel(const float2 *__restrict__ a,
                                                                 const float2 *__restrict__ b, size_t n,
                                                                 double *__restrict__ partials) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float2 x = a[i], y = b[i];
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
        s[0] += xr * yr + xi * yi;
        s[1] += xr * yi - xi * yr;
        s[2] += xr * xr + xi * xi;
        s[3] += yr * yr + yi * yi;
    }
    __shared__ double red[kThreads / 64][4];
    block_sums4(s, red);
    if (threadIdx.x == 0)
        for (int j = 0; j < 4; j++) partials[4 * blockIdx.x + j] = red[0][j];
}

// partials[4 g] = sum |psi_i|^2 obs_i (FP64), the other three slots |psi|^2, 0, 0.
__global__ __launch_bounds__(kThreads) void expect_diag_partials_kernel(const float2 *__restrict__ psi,
                                                                        const float *__restrict__ obs, size_t n,
                                                                        double *__restrict__ partials) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float2 x = psi[i];
        const double p = (double)x.x * x.x + (double)x.y * x.y;
        s[0] += p * (double)obs[i];
        s[1] += p;
    }
    __shared__ double red[kThreads / 64][4];
    block_sums4(s, red);
    if (threadIdx.x == 0)
        for (int j = 0; j < 4; j++) partials[4 * blockIdx.x + j] = red[0][j];
}

// out[0..3] = sums of the `parts` partial quadruples, added in a fixed order.
__global__ __launch_bounds__(kThreads) void reduce4_kernel(const double *__restrict__ partials, int parts,
                                                           double *__restrict__ out) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int g = threadIdx.x; g < parts; g += kThreads)
        for (int j = 0; j < 4; j++) s[j] += partials[4 * g + j];
    __shared__ double red[kThreads / 64][4];
    block_sums4(s, red);
    if (threadIdx.x == 0)
        for (int j = 0; j < 4; j++) out[j] = red[0][j];
}

__global__ __launch_bounds__(kThreads) void scale_c64_kernel(float2 *__restrict__ x, size_t n, float s) {
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        float2 v = x[i];
        v.x *= s;
        v.y *= s;
        x[i] = v;
    }
}

// y = U x, U row-major m x ncols complex64.  One wave per row: the lanes
// stride over the row (coalesced 8-byte loads), FP64 accumulation, a wave sum.
__global__ __launch_bounds__(kThreads) void cgemv_rows_kernel(const float2 *__restrict__ U,
                                                              const float2 *__restrict__ x, float2 *__restrict__ y,
                                                              size_t m, size_t ncols) {
    const size_t waves = (size_t)gridDim.x * (kThreads / 64);
    const int lane = threadIdx.x & 63;
    for (size_t r = blockIdx.x * (size_t)(kThreads / 64) + (threadIdx.x >> 6); r < m; r += waves) {
        const float2 *row = U + r * ncols;
        double sr = 0.0, si = 0.0;
        for (size_t j = lane; j < ncols; j += 64) {
            const float2 u = row[j], v = x[j];
            sr += (double)u.x * v.x - (double)u.y * v.y;
            si += (double)u.x * v.y + (double)u.y * v.x;
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) y[r] = make_float2((float)sr, (float)si);
    }
}

// C = A B, row-major complex64: A m x n, B n x k, C m x k.  16 x 16 output
// tiles, 16-wide slabs of A and B staged through LDS, FP32 accumulation (the
// arithmetic of QGTL's simd_complex_matrix_multiply).
constexpr int kGemmTile = 16;
__global__ __launch_bounds__(kGemmTile * kGemmTile) void cgemm_c64_kernel(const float2 *__restrict__ A,
                                                                        const float2 *__restrict__ B,
                                                                        float2 *__restrict__ C, size_t m, size_t n,
                                                                        size_t k) {
    __shared__ float2 As[kGemmTile][kGemmTile + 1];
    __shared__ float2 Bs[kGemmTile][kGemmTile + 1];
    const int tx = threadIdx.x % kGemmTile, ty = threadIdx.x / kGemmTile;
    const size_t row = (size_t)blockIdx.y * kGemmTile + ty, col = (size_t)blockIdx.x * kGemmTile + tx;
    float accr = 0.f, acci = 0.f;
    for (size_t l0 = 0; l0 < n; l0 += kGemmTile) {
        As[ty][tx] = (row < m && l0 + tx < n) ? A[row * n + l0 + tx] : make_float2(0.f, 0.f);
        Bs[ty][tx] = (l0 + ty < n && col < k) ? B[(l0 + ty) * k + col] : make_float2(0.f, 0.f);
        __syncthreads();
#pragma unroll
        for (int l = 0; l < kGemmTile; l++) {
            const float2 a = As[ty][l], b = Bs[l][tx];
            accr = fmaf(a.x, b.x, fmaf(-a.y, b.y, accr));
            acci = fmaf(a.x, b.y, fmaf(a.y, b.x, acci));
        }
        __syncthreads();
    }
    if (row < m && col < k) C[row * k + col] = make_float2(accr, acci);
}

unsigned parts_for(size_t n) {
    size_t g = (n + (size_t)kThreads * 8 - 1) / ((size_t)kThreads * 8);
    if (g < 1) g = 1;
    if (g > (size_t)kMaxParts) g = kMaxParts;
    return (unsigned)g;
}

}  // namespace

extern "C" {

int qgtl_hip_vec_partials_len(void) { return 4 * kMaxParts; }

int qgtl_hip_vec_dotc_c64(const void *a, const void *b, size_t n, double *partials, double *out4) {
    if (n == 0) return qgtl::set_error_msg(-118, "qgtl_hip_vec_dotc_c64: empty vector");
    const unsigned g = parts_for(n);
    hipLaunchKernelGGL(dotc_partials_kernel, dim3(g), dim3(kThreads), 0, qgtl::stream(), (const float2 *)a,
                       (const float2 *)b, n, partials);
    QGTL_LAUNCH_CHECK("dotc_partials");
    hipLaunchKernelGGL(reduce4_kernel, dim3(1), dim3(kThreads), 0, qgtl::stream(), partials, (int)g, out4);
    QGTL_LAUNCH_CHECK("reduce4");
    return 0;
}

int qgtl_hip_vec_expect_diag_c64(const void *psi, const float *obs, size_t n, double *partials, double *out4) {
    if (n == 0) return qgtl::set_error_msg(-118, "qgtl_hip_vec_expect_diag_c64: empty vector");
    const unsigned g = parts_for(n);
    hipLaunchKernelGGL(expect_diag_partials_kernel, dim3(g), dim3(kThreads), 0, qgtl::stream(), (const float2 *)psi,
                       obs, n, partials);
    QGTL_LAUNCH_CHECK("expect_diag_partials");
    hipLaunchKernelGGL(reduce4_kernel, dim3(1), dim3(kThreads), 0, qgtl::stream(), partials, (int)g, out4);
    QGTL_LAUNCH_CHECK("reduce4");
    return 0;
}

int qgtl_hip_vec_scale_c64(void *x, size_t n, float s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(scale_c64_kernel, dim3(parts_for(n)), dim3(kThreads), 0, qgtl::stream(), (float2 *)x, n, s);
    QGTL_LAUNCH_CHECK("scale_c64");
    return 0;
}

int qgtl_hip_vec_cgemv_c64(const void *U, const void *x, void *y, size_t m, size_t ncols) {
    if (m == 0 || ncols == 0) return qgtl::set_error_msg(-118, "qgtl_hip_vec_cgemv_c64: empty matrix");
    size_t blocks = (m + (kThreads / 64) - 1) / (kThreads / 64);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(cgemv_rows_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, qgtl::stream(),
                       (const float2 *)U, (const float2 *)x, (float2 *)y, m, ncols);
    QGTL_LAUNCH_CHECK("cgemv_rows");
    return 0;
}

int qgtl_hip_vec_cgemm_c64(const void *A, const void *B, void *C, size_t m, size_t n, size_t k) {
    if (m == 0 || n == 0 || k == 0) return qgtl::set_error_msg(-118, "qgtl_hip_vec_cgemm_c64: empty matrix");
    if ((m + kGemmTile - 1) / kGemmTile > 65535)
        return qgtl::set_error_msg(-118, "qgtl_hip_vec_cgemm_c64: more than 65535 x 16 rows");
    const dim3 grid((unsigned)((k + kGemmTile - 1) / kGemmTile), (unsigned)((m + kGemmTile - 1) / kGemmTile));
    hipLaunchKernelGGL(cgemm_c64_kernel, grid, dim3(kGemmTile * kGemmTile), 0, qgtl::stream(), (const float2 *)A,
                       (const float2 *)B, (float2 *)C, m, n, k);
    QGTL_LAUNCH_CHECK("cgemm_c64");
    return 0;
}

}  // extern "C"
