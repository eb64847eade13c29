// Fixed per-GPU working memory for qgtl_solve_f64_noalloc: the variant of the
// natural-gradient solve for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops).  Its device buffers (staged
// matrices and right-hand sides, solutions, preconditioner, per-system status)
// are carved out of one file-scope array that every GPU the library runs on
// holds until the process ends.
//
// In host code the array's name is only a host-side handle: its device address
// is looked up once per device, with that device current, at library init
// (qgtl_hip_noalloc_init, called by qgtl_hip_init), and only that address is
// handed to kernels and copies.
#include "qgtl_hip_common.h"

namespace {

constexpr size_t kScratchBytes = QGTL_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

void *g_scratch_ptr[kMaxDevices];   // device address of `scratch` on each device, once looked up

// d[s][k] = 1 / (A[s][k][k] + lambda): the Jacobi preconditioner of system s
__global__ void jacobi_diag_kernel(const double *__restrict__ A, int P, int batch, double lambda, double *d) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)P * batch) return;
    const size_t s = i / P, k = i % P;
    d[i] = 1.0 / (A[s * P * P + k * P + k] + lambda);
}

}  // namespace

extern "C" {

int qgtl_hip_noalloc_init(void) {
    int dev = -1;
    QGTL_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return qgtl::set_error_msg(-118, "qgtl_hip_noalloc_init: device index out of range");
    if (g_scratch_ptr[dev]) return 0;
    void *p = nullptr;
    QGTL_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
    return 0;
}

int qgtl_hip_noalloc_scratch(void **ptr, size_t *bytes) {
    *ptr = nullptr;
    *bytes = 0;
    int dev = -1;
    QGTL_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices || !g_scratch_ptr[dev])
        return qgtl::set_error_msg(-120, "qgtl_hip_noalloc_scratch: scratch address not looked up on this device");
    *ptr = g_scratch_ptr[dev];
    *bytes = kScratchBytes;
    return 0;
}

int qgtl_hip_jacobi_diag_f64(const double *A, int P, int batch, double lambda, double *d) {
    const size_t total = (size_t)P * batch;
    const unsigned grid = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(jacobi_diag_kernel, dim3(grid), dim3(256), 0, qgtl::stream(), A, P, batch, lambda, d);
    QGTL_LAUNCH_CHECK("jacobi_diag");
    return 0;
}

}  // extern "C"
