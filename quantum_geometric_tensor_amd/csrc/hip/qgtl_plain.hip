// Plain FP64 GPU path: complex128 state, one kernel launch per gate,
// derivative states generated one at a time, a tiled FP64 Gram and the FP64
// correction.  It is the denominator of the mixed path's speed-up and the
// GPU-side FP64 cross-check; it follows QGTL's gate loops one to one
// (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:174-210) and
// the FP64 QGT of src/quantum_geometric/distributed/differential_geometry.c:2832-2858.
#include "qgtl_hip_common.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ uint32_t insert_zero(uint32_t x, int b) {
    return ((x >> b) << (b + 1)) | (x & ((1u << b) - 1u));
}

struct M8 { double v[8]; };

__device__ __forceinline__ void apply2(double2 &a, double2 &b, const M8 &m) {
    const double2 na = make_double2(m.v[0] * a.x - m.v[1] * a.y + m.v[2] * b.x - m.v[3] * b.y,
                                    m.v[0] * a.y + m.v[1] * a.x + m.v[2] * b.y + m.v[3] * b.x);
    const double2 nb = make_double2(m.v[4] * a.x - m.v[5] * a.y + m.v[6] * b.x - m.v[7] * b.y,
                                    m.v[4] * a.y + m.v[5] * a.x + m.v[6] * b.y + m.v[7] * b.x);
    a = na;
    b = nb;
}

__global__ void fill_zero_c128_kernel(double2 *psi, uint32_t N) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        psi[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
}

__global__ void gate1q_c128_kernel(double2 *psi, uint32_t half, int q, M8 m) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < half; t += gridDim.x * blockDim.x) {
        const uint32_t i = insert_zero(t, q), j = i | (1u << q);
        double2 a = psi[i], b = psi[j];
        apply2(a, b, m);
        psi[i] = a;
        psi[j] = b;
    }
}

__global__ void cgate_c128_kernel(double2 *psi, uint32_t quarter, int c, int t, M8 m) {
    const int lo = min(c, t), hi = max(c, t);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < quarter; x += gridDim.x * blockDim.x) {
        const uint32_t i = insert_zero(insert_zero(x, lo), hi) | (1u << c), j = i | (1u << t);
        double2 a = psi[i], b = psi[j];
        apply2(a, b, m);
        psi[i] = a;
        psi[j] = b;
    }
}

__global__ void swap_c128_kernel(double2 *psi, uint32_t quarter, int qa, int qb) {
    const int lo = min(qa, qb), hi = max(qa, qb);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < quarter; x += gridDim.x * blockDim.x) {
        const uint32_t b0 = insert_zero(insert_zero(x, lo), hi);
        const uint32_t i = b0 | (1u << qa), j = b0 | (1u << qb);
        const double2 a = psi[i];
        psi[i] = psi[j];
        psi[j] = a;
    }
}

// Tiled FP64 Gram: a 256-thread workgroup owns a 32 x 32 block of G over one
// K chunk; each thread accumulates 2 x 2 complex outputs.
constexpr int kGT = 32;
constexpr int kGK = 16;

__device__ __forceinline__ void tile_coords(int t, int nt, int &ti, int &tj) {
    int i = 0, rem = t;
    while (rem >= nt - i) { rem -= nt - i; i++; }
    ti = i;
    tj = i + rem;
}

__global__ __launch_bounds__(kThreads) void gram_f64_kernel(const double2 *__restrict__ D, int P, size_t N,
                                                            int nt, int nchunks, size_t chunk_len,
                                                            double2 *__restrict__ partials) {
    __shared__ double2 sa[kGT][kGK + 1], sb[kGT][kGK + 1];
    const int T = nt * (nt + 1) / 2;
    const int tile = blockIdx.x % T, chunk = blockIdx.x / T;
    int ti, tj;
    tile_coords(tile, nt, ti, tj);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double2 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = make_double2(0.0, 0.0);
    const size_t kbeg = (size_t)chunk * chunk_len;
    size_t kend = kbeg + chunk_len;
    if (kend > N) kend = N;
    for (size_t k0 = kbeg; k0 < kend; k0 += kGK) {
        for (int e = threadIdx.x; e < kGT * kGK; e += kThreads) {
            const int r = e / kGK, kk = e % kGK;
            const int sA = ti * kGT + r, sB = tj * kGT + r;
            const size_t k = k0 + kk;
            sa[r][kk] = (sA < P && k < kend) ? D[(size_t)sA * N + k] : make_double2(0.0, 0.0);
            sb[r][kk] = (sB < P && k < kend) ? D[(size_t)sB * N + k] : make_double2(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < kGK; kk++) {
            double2 a[2], b[2];
            for (int i = 0; i < 2; i++) { a[i] = sa[ty * 2 + i][kk]; b[i] = sb[tx * 2 + i][kk]; }
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    // conj(a) * b
                    acc[i][j].x = fma(a[i].x, b[j].x, fma(a[i].y, b[j].y, acc[i][j].x));
                    acc[i][j].y = fma(a[i].x, b[j].y, fma(-a[i].y, b[j].x, acc[i][j].y));
                }
        }
        __syncthreads();
    }
    double2 *out = partials + ((size_t)chunk * T + tile) * (kGT * kGT);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) out[(ty * 2 + i) * kGT + tx * 2 + j] = acc[i][j];
}

__global__ __launch_bounds__(kThreads) void gram_f64_reduce_kernel(const double2 *__restrict__ partials, int P,
                                                                   int nt, int nchunks, double2 *__restrict__ G) {
    const int T = nt * (nt + 1) / 2;
    const size_t total = (size_t)T * kGT * kGT;
    for (size_t idx = blockIdx.x * (size_t)kThreads + threadIdx.x; idx < total; idx += (size_t)gridDim.x * kThreads) {
        const int tile = (int)(idx / (kGT * kGT)), e = (int)(idx % (kGT * kGT));
        int ti, tj;
        tile_coords(tile, nt, ti, tj);
        const int r = ti * kGT + e / kGT, c = tj * kGT + e % kGT;
        if (r >= P || c >= P) continue;
        double2 s = make_double2(0.0, 0.0);
        for (int ch = 0; ch < nchunks; ch++) {
            const double2 v = partials[((size_t)ch * T + tile) * (kGT * kGT) + e];
            s.x += v.x;
            s.y += v.y;
        }
        G[(size_t)r * P + c] = s;
    }
}

// a_k = <D_k|psi>: one workgroup per (tile of 4096 amplitudes, state).
constexpr int kDotTile = 4096;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(kThreads) void dotc_c128_kernel(const double2 *__restrict__ D, size_t N,
                                                             const double2 *__restrict__ psi, double *partials) {
    const int k = blockIdx.y;
    const double2 *d = D + (size_t)k * N;
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < kDotTile / kThreads; i++) {
        const size_t y = (size_t)blockIdx.x * kDotTile + i * kThreads + threadIdx.x;
        if (y < N) {
            const double2 a = d[y], b = psi[y];
            sr += a.x * b.x + a.y * b.y;
            si += a.x * b.y - a.y * b.x;
        }
    }
    __shared__ double red[kThreads / 64][2];
    sr = wave_sum(sr);
    si = wave_sum(si);
    if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6][0] = sr; red[threadIdx.x >> 6][1] = si; }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        if (threadIdx.x < 2)
            for (int w = 0; w < kThreads / 64; w++) s += red[w][threadIdx.x];
        partials[((size_t)k * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = s;
    }
}

__global__ void finalize_plain_kernel(const double2 *__restrict__ G, const double *__restrict__ a4, int P,
                                      double2 *__restrict__ Q) {
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)P * P;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(idx / P), l = (int)(idx % P);
        if (k > l) continue;
        const double akr = a4[4 * k], aki = a4[4 * k + 1], alr = a4[4 * l], ali = a4[4 * l + 1];
        const double pr = akr * alr + aki * ali, pi = aki * alr - akr * ali;   // a_k conj(a_l)
        const double2 g = G[(size_t)k * P + l];
        const double2 q = make_double2(g.x - pr, g.y - pi);
        Q[(size_t)k * P + l] = q;
        if (k != l) Q[(size_t)l * P + k] = make_double2(q.x, -q.y);
    }
}

unsigned grid_for(size_t work, unsigned cap = 16384) {
    size_t g = (work + kThreads - 1) / kThreads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

M8 to_m8(const double *m) {
    M8 r;
    for (int i = 0; i < 8; i++) r.v[i] = m[i];
    return r;
}

int gram_f64_chunks(int P, size_t N) {
    const int nt = (P + kGT - 1) / kGT;
    const int T = nt * (nt + 1) / 2;
    size_t want = (2048 + T - 1) / T;
    size_t maxc = (N + 1023) / 1024;
    if (want > maxc) want = maxc;
    if (want < 1) want = 1;
    return (int)want;
}

}  // namespace

extern "C" {

int qgtl_hip_fill_zero_state_c128(void *psi, int n) {
    const uint32_t N = 1u << n;
    hipLaunchKernelGGL(fill_zero_c128_kernel, dim3(grid_for(N)), dim3(kThreads), 0, qgtl::stream(), (double2 *)psi, N);
    QGTL_LAUNCH_CHECK("fill_zero_c128");
    return 0;
}

int qgtl_hip_gate1q_c128(void *psi, int n, int q, const double *m8) {
    const uint32_t half = 1u << (n - 1);
    hipLaunchKernelGGL(gate1q_c128_kernel, dim3(grid_for(half)), dim3(kThreads), 0, qgtl::stream(),
                       (double2 *)psi, half, q, to_m8(m8));
    QGTL_LAUNCH_CHECK("gate1q_c128");
    return 0;
}

int qgtl_hip_cgate_c128(void *psi, int n, int c, int t, const double *m8) {
    if (n < 2) return qgtl::set_error_msg(-118, "controlled gate needs two qubits");
    const uint32_t quarter = 1u << (n - 2);
    hipLaunchKernelGGL(cgate_c128_kernel, dim3(grid_for(quarter)), dim3(kThreads), 0, qgtl::stream(),
                       (double2 *)psi, quarter, c, t, to_m8(m8));
    QGTL_LAUNCH_CHECK("cgate_c128");
    return 0;
}

int qgtl_hip_swap_c128(void *psi, int n, int a, int b) {
    if (n < 2) return qgtl::set_error_msg(-118, "swap needs two qubits");
    const uint32_t quarter = 1u << (n - 2);
    hipLaunchKernelGGL(swap_c128_kernel, dim3(grid_for(quarter)), dim3(kThreads), 0, qgtl::stream(),
                       (double2 *)psi, quarter, a, b);
    QGTL_LAUNCH_CHECK("swap_c128");
    return 0;
}

size_t qgtl_hip_gram_f64_partials(int P, size_t N) {
    const int nt = (P + kGT - 1) / kGT;
    const size_t T = (size_t)nt * (nt + 1) / 2;
    return T * gram_f64_chunks(P, N) * kGT * kGT * 2;
}

int qgtl_hip_gram_f64(const void *D, int P, size_t N, double *partials, double *G) {
    const int nt = (P + kGT - 1) / kGT;
    const int T = nt * (nt + 1) / 2;
    const int nchunks = gram_f64_chunks(P, N);
    size_t chunk_len = (N + nchunks - 1) / nchunks;
    chunk_len = (chunk_len + kGK - 1) / kGK * kGK;
    hipLaunchKernelGGL(gram_f64_kernel, dim3((unsigned)(T * nchunks)), dim3(kThreads), 0, qgtl::stream(),
                       (const double2 *)D, P, N, nt, nchunks, chunk_len, (double2 *)partials);
    QGTL_LAUNCH_CHECK("gram_f64");
    const size_t total = (size_t)T * kGT * kGT;
    hipLaunchKernelGGL(gram_f64_reduce_kernel, dim3(grid_for(total)), dim3(kThreads), 0, qgtl::stream(),
                       (const double2 *)partials, P, nt, nchunks, (double2 *)G);
    QGTL_LAUNCH_CHECK("gram_f64_reduce");
    return 0;
}

int qgtl_hip_dotc_c128(const void *D, int P, size_t N, const void *psi, double *partials, double *a) {
    const unsigned tiles = (unsigned)((N + kDotTile - 1) / kDotTile);
    hipLaunchKernelGGL(dotc_c128_kernel, dim3(tiles, (unsigned)P), dim3(kThreads), 0, qgtl::stream(),
                       (const double2 *)D, N, (const double2 *)psi, partials);
    QGTL_LAUNCH_CHECK("dotc_c128");
    return qgtl_hip_reduce_partials_f64(partials, P, (int)tiles, a);
}

int qgtl_hip_qgt_finalize_plain(const double *G, const double *a, int P, double *Q) {
    hipLaunchKernelGGL(finalize_plain_kernel, dim3(grid_for((size_t)P * P, 4096)), dim3(kThreads), 0, qgtl::stream(),
                       (const double2 *)G, a, P, (double2 *)Q);
    QGTL_LAUNCH_CHECK("finalize_plain");
    return 0;
}

}  // extern "C"
