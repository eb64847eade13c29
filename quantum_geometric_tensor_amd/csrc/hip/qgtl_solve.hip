// FP64 natural-gradient solves on the GPU, one workgroup per system (batched
// over blockIdx.x), every loop bounded.
//
//  - lu_solve_c128: (A) x = b by LU with partial pivoting, singular when
//    |pivot|^2 < 1e-10 on columns 0..P-2 -- the rule of QGTL's
//    lu_decomposition / solve_linear_system
//    (src/quantum_geometric/core/matrix_operations.c:60-160) that
//    compute_regularized_natural_gradient relies on
//    (src/quantum_geometric/core/quantum_geometric_gradient.c:2887-2964).
//  - herm_eig_c128: parallel cyclic Jacobi (round-robin pairs) for Hermitian
//    matrices: eigenvalues ascending and eigenvectors, for the condition
//    number and the pseudo-inverse fallback (:2733-2884).
//  - cholesky_f64: (A + lambda I) x = b for real SPD A, the exact yardstick.
//  - pcg_f64: preconditioned CG from x0 = 0 to ||r||/||b|| <= tol with
//    M^-1 = diag(d) + U U^T (U = 0 gives Jacobi-PCG with d = 1/diag(A)),
//    QGTL's CG (src/quantum_geometric/hybrid/classical_optimization_engine.c:543-610)
//    with a preconditioner.
#include "qgtl_hip_common.h"

namespace {

constexpr int kT = 1024;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// Block-wide sum in a fixed order (deterministic).
__device__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = kT / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------- LU (complex)
__global__ __launch_bounds__(kT) void lu_solve_c128_kernel(double2 *Aall, const double2 *ball, double2 *xall,
                                                           int *pivall, int P, int *status) {
    double2 *A = Aall + (size_t)blockIdx.x * P * P;
    const double2 *b = ball + (size_t)blockIdx.x * P;
    double2 *x = xall + (size_t)blockIdx.x * P;
    int *piv = pivall + (size_t)blockIdx.x * P;
    __shared__ double red_v[kT];
    __shared__ int red_i[kT];
    __shared__ int s_fail;
    for (int i = threadIdx.x; i < P; i += kT) piv[i] = i;
    if (threadIdx.x == 0) s_fail = 0;
    __syncthreads();
    for (int i = 0; i < P - 1; i++) {
        // pivot: max |a_ji|^2 over j >= i, first index on ties
        double best = -1.0;
        int bi = i;
        for (int j = i + threadIdx.x; j < P; j += kT) {
            const double2 a = A[(size_t)j * P + i];
            const double v = a.x * a.x + a.y * a.y;
            if (v > best) { best = v; bi = j; }
        }
        red_v[threadIdx.x] = best;
        red_i[threadIdx.x] = bi;
        __syncthreads();
        for (int h = kT / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) {
                const double o = red_v[threadIdx.x + h];
                const int oi = red_i[threadIdx.x + h];
                if (o > red_v[threadIdx.x] || (o == red_v[threadIdx.x] && oi < red_i[threadIdx.x])) {
                    red_v[threadIdx.x] = o;
                    red_i[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        const double mx = red_v[0];
        const int p = red_i[0];
        __syncthreads();
        if (mx < 1e-10) {
            if (threadIdx.x == 0) s_fail = 1;
            break;
        }
        if (p != i) {
            for (int k = threadIdx.x; k < P; k += kT) {
                const double2 t = A[(size_t)i * P + k];
                A[(size_t)i * P + k] = A[(size_t)p * P + k];
                A[(size_t)p * P + k] = t;
            }
            if (threadIdx.x == 0) { const int t = piv[i]; piv[i] = piv[p]; piv[p] = t; }
        }
        __syncthreads();
        const double2 d = A[(size_t)i * P + i];
        for (int j = i + 1 + threadIdx.x; j < P; j += kT) A[(size_t)j * P + i] = cdiv(A[(size_t)j * P + i], d);
        __syncthreads();
        const int m = P - i - 1;
        for (int e = threadIdx.x; e < m * m; e += kT) {
            const int j = i + 1 + e / m, k = i + 1 + e % m;
            const double2 l = A[(size_t)j * P + i], u = A[(size_t)i * P + k];
            double2 a = A[(size_t)j * P + k];
            const double2 lu = cmul(l, u);
            a.x -= lu.x;
            a.y -= lu.y;
            A[(size_t)j * P + k] = a;
        }
        __syncthreads();
    }
    if (s_fail) {
        if (threadIdx.x == 0) status[blockIdx.x] = -1;
        return;
    }
    // forward (unit lower) and back substitution by one wave through LDS (an
    // LDS write by lane 0 is seen by the wave's later LDS reads: in order)
    __shared__ double2 sx[1024];
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int i = lane; i < P; i += 64) sx[i] = b[piv[i]];
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        for (int i = 0; i < P; i++) {
            double sr = 0.0, si = 0.0;
            for (int j = lane; j < i; j += 64) {
                const double2 t = cmul(A[(size_t)i * P + j], sx[j]);
                sr += t.x;
                si += t.y;
            }
            for (int o = 32; o > 0; o >>= 1) { sr += __shfl_xor(sr, o, 64); si += __shfl_xor(si, o, 64); }
            if (lane == 0) sx[i] = make_double2(sx[i].x - sr, sx[i].y - si);
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
        }
        for (int i = P - 1; i >= 0; i--) {
            double sr = 0.0, si = 0.0;
            for (int j = i + 1 + lane; j < P; j += 64) {
                const double2 t = cmul(A[(size_t)i * P + j], sx[j]);
                sr += t.x;
                si += t.y;
            }
            for (int o = 32; o > 0; o >>= 1) { sr += __shfl_xor(sr, o, 64); si += __shfl_xor(si, o, 64); }
            if (lane == 0) sx[i] = cdiv(make_double2(sx[i].x - sr, sx[i].y - si), A[(size_t)i * P + i]);
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
        }
        for (int i = lane; i < P; i += 64) x[i] = sx[i];
        if (lane == 0) status[blockIdx.x] = 0;
    }
}

// ------------------------------------------------------- Hermitian Jacobi
// A (in/out, destroyed), V (out, columns = eigenvectors), ev (out, ascending).
__global__ __launch_bounds__(kT) void herm_eig_c128_kernel(double2 *Aall, double2 *Vall, double *evall, int P,
                                                           int max_sweeps, int *sweeps_out) {
    double2 *A = Aall + (size_t)blockIdx.x * P * P;
    double2 *V = Vall + (size_t)blockIdx.x * P * P;
    double *ev = evall + (size_t)blockIdx.x * P;
    __shared__ double red[kT];
    __shared__ int pair_p[512], pair_q[512];
    __shared__ double rot[512][6];   // c, s, er, ei, active
    const int M = (P + 1) & ~1;      // even count; index P is a dummy when P is odd
    for (int e = threadIdx.x; e < P * P; e += kT) V[e] = make_double2((e / P) == (e % P) ? 1.0 : 0.0, 0.0);
    double fro = 0.0;
    for (int e = threadIdx.x; e < P * P; e += kT) fro += A[e].x * A[e].x + A[e].y * A[e].y;
    fro = sqrt(block_sum(fro, red));
    int sweep = 0;
    for (; sweep < max_sweeps; sweep++) {
        double off = 0.0;
        for (int e = threadIdx.x; e < P * P; e += kT)
            if (e / P != e % P) off += A[e].x * A[e].x + A[e].y * A[e].y;
        off = sqrt(block_sum(off, red));
        if (off <= 1e-15 * (fro > 0.0 ? fro : 1.0)) break;
        for (int step = 0; step < M - 1; step++) {
            // round-robin: position 0 fixed, others rotate
            for (int k = threadIdx.x; k < M / 2; k += kT) {
                int a = (k == 0) ? 0 : 1 + (k - 1 + step) % (M - 1);
                int bidx = 1 + (M - 2 - k + step) % (M - 1);
                int p = a < bidx ? a : bidx, q = a < bidx ? bidx : a;
                pair_p[k] = p;
                pair_q[k] = q;
                double c = 1.0, s = 0.0, er = 1.0, ei = 0.0, act = 0.0;
                if (q < P) {
                    const double2 apq = A[(size_t)p * P + q];
                    const double mag = sqrt(apq.x * apq.x + apq.y * apq.y);
                    if (mag > 0.0) {
                        er = apq.x / mag;
                        ei = apq.y / mag;
                        const double tau = (A[(size_t)q * P + q].x - A[(size_t)p * P + p].x) / (2.0 * mag);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        act = 1.0;
                    }
                }
                rot[k][0] = c; rot[k][1] = s; rot[k][2] = er; rot[k][3] = ei; rot[k][4] = act;
            }
            __syncthreads();
            // columns: A <- A J, V <- V J with J = [[c, s], [-s conj(e), c conj(e)]]
            for (int e = threadIdx.x; e < P * (M / 2); e += kT) {
                const int k = e / P, r = e % P;
                if (rot[k][4] == 0.0) continue;
                const int p = pair_p[k], q = pair_q[k];
                const double c = rot[k][0], s = rot[k][1], er = rot[k][2], ei = rot[k][3];
                const double2 jqp = make_double2(-s * er, s * ei), jqq = make_double2(c * er, -c * ei);
                double2 *Ar = A + (size_t)r * P;
                const double2 ap = Ar[p], aq = Ar[q];
                const double2 t1 = cmul(aq, jqp), t2 = cmul(aq, jqq);
                Ar[p] = make_double2(c * ap.x + t1.x, c * ap.y + t1.y);
                Ar[q] = make_double2(s * ap.x + t2.x, s * ap.y + t2.y);
                double2 *Vr = V + (size_t)r * P;
                const double2 vp = Vr[p], vq = Vr[q];
                const double2 u1 = cmul(vq, jqp), u2 = cmul(vq, jqq);
                Vr[p] = make_double2(c * vp.x + u1.x, c * vp.y + u1.y);
                Vr[q] = make_double2(s * vp.x + u2.x, s * vp.y + u2.y);
            }
            __syncthreads();
            // rows: A <- J^H A
            for (int e = threadIdx.x; e < P * (M / 2); e += kT) {
                const int k = e / P, col = e % P;
                if (rot[k][4] == 0.0) continue;
                const int p = pair_p[k], q = pair_q[k];
                const double c = rot[k][0], s = rot[k][1], er = rot[k][2], ei = rot[k][3];
                const double2 cjqp = make_double2(-s * er, -s * ei), cjqq = make_double2(c * er, c * ei);
                const double2 pk = A[(size_t)p * P + col], qk = A[(size_t)q * P + col];
                const double2 t1 = cmul(cjqp, qk), t2 = cmul(cjqq, qk);
                A[(size_t)p * P + col] = make_double2(c * pk.x + t1.x, c * pk.y + t1.y);
                A[(size_t)q * P + col] = make_double2(s * pk.x + t2.x, s * pk.y + t2.y);
            }
            __syncthreads();
            for (int k = threadIdx.x; k < M / 2; k += kT) {
                if (rot[k][4] == 0.0) continue;
                const int p = pair_p[k], q = pair_q[k];
                A[(size_t)p * P + q] = make_double2(0.0, 0.0);
                A[(size_t)q * P + p] = make_double2(0.0, 0.0);
                A[(size_t)p * P + p].y = 0.0;
                A[(size_t)q * P + q].y = 0.0;
            }
            __syncthreads();
        }
    }
    // eigenvalues ascending (rank by value, index breaks ties), permute V's columns
    for (int i = threadIdx.x; i < P; i += kT) {
        const double vi = A[(size_t)i * P + i].x;
        int rank = 0;
        for (int j = 0; j < P; j++) {
            const double vj = A[(size_t)j * P + j].x;
            rank += (vj < vi) || (vj == vi && j < i);
        }
        ev[rank] = vi;
        red[i] = rank;   // P <= kT
    }
    __syncthreads();
    // copy V into A (scratch) then scatter columns
    for (int e = threadIdx.x; e < P * P; e += kT) A[e] = V[e];
    __syncthreads();
    for (int e = threadIdx.x; e < P * P; e += kT) {
        const int r = e / P, c = e % P;
        V[(size_t)r * P + (int)red[c]] = A[e];
    }
    if (threadIdx.x == 0 && sweeps_out) sweeps_out[blockIdx.x] = sweep;
}

// ------------------------------------------------------------- Cholesky (real)
__global__ __launch_bounds__(kT) void cholesky_f64_kernel(double *Lall, const double *ball, double *xall, int P,
                                                          double lambda, int *status) {
    double *L = Lall + (size_t)blockIdx.x * P * P;    // holds A on entry (row-major, symmetric)
    const double *b = ball + (size_t)blockIdx.x * P;
    double *x = xall + (size_t)blockIdx.x * P;
    __shared__ int s_fail;
    if (threadIdx.x == 0) s_fail = 0;
    for (int i = threadIdx.x; i < P; i += kT) L[(size_t)i * P + i] += lambda;
    __syncthreads();
    for (int j = 0; j < P; j++) {
        const double d = L[(size_t)j * P + j];
        if (!(d > 0.0)) {
            if (threadIdx.x == 0) s_fail = 1;
            break;
        }
        const double ljj = sqrt(d);
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < P; i += kT) L[(size_t)i * P + j] /= ljj;
        if (threadIdx.x == 0) L[(size_t)j * P + j] = ljj;
        __syncthreads();
        const int m = P - j - 1;
        for (int e = threadIdx.x; e < m * (m + 1) / 2; e += kT) {
            // lower triangle of the trailing block: (i, k) with k <= i
            int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) / 2.0);
            while ((i + 1) * (i + 2) / 2 <= e) i++;
            while (i * (i + 1) / 2 > e) i--;
            const int k = e - i * (i + 1) / 2;
            const int ii = j + 1 + i, kk = j + 1 + k;
            L[(size_t)ii * P + kk] -= L[(size_t)ii * P + j] * L[(size_t)kk * P + j];
        }
        __syncthreads();
    }
    __syncthreads();
    if (s_fail) {
        if (threadIdx.x == 0) status[blockIdx.x] = -1;
        return;
    }
    __shared__ double sx[1024];
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int i = 0; i < P; i++) {
            double s = 0.0;
            for (int k = lane; k < i; k += 64) s += L[(size_t)i * P + k] * sx[k];
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) sx[i] = (b[i] - s) / L[(size_t)i * P + i];
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
        }
        for (int i = P - 1; i >= 0; i--) {
            double s = 0.0;
            for (int k = i + 1 + lane; k < P; k += 64) s += L[(size_t)k * P + i] * sx[k];
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) sx[i] = (sx[i] - s) / L[(size_t)i * P + i];
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
        }
        for (int i = lane; i < P; i += 64) x[i] = sx[i];
        if (lane == 0) status[blockIdx.x] = 0;
    }
}

// ------------------------------------------------------------------- PCG
// A symmetric (row-major, so A[j][i] = A[i][j] gives coalesced column reads).

// PCG A-column cache: each thread keeps kPcgRegs of its matrix-vector columns in registers
// and the next kPcgLds in LDS; the rest (P above about 170) come from global memory.
constexpr int kPcgRegs = 18;
constexpr int kPcgLds = 14;
// U (P x r) is copied to LDS when P * r fits.
constexpr int kPcgUCache = 2048;

// One lane of a double moved by a DPP pattern (two 32-bit DPP moves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sum over the 64 lanes of a wave, identical in every lane and independent of timing:
// pairs, quads (quad_perm), half rows and rows of 16 (row_half_mirror, row_mirror) by DPP,
// each step adding two equal-shaped partial sums so every lane of a row holds the same bits;
// then the four row sums (lanes 0, 16, 32, 48) in the order (r0 + r1) + (r2 + r3).
__device__ __forceinline__ double wave_sum_f64(double v) {
    v += dpp_f64<0xB1>(v);                            // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);                            // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);                           // row_half_mirror
    v += dpp_f64<0x140>(v);                           // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// W block sums with one barrier: wave sums into wp[wave][c], barrier, every thread adds the
// 16 wave sums in wave order.  wp must not be rewritten before the next barrier, so callers
// alternate between two buffers.
template <int W>
__device__ __forceinline__ void block_sums(double (&v)[W], double (*wp)[8]) {
    static_assert(W >= 1 && W <= 2, "one or two sums per call");
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = wave_sum_f64(v[c]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < W; c++) wp[threadIdx.x >> 6][c] = v[c];
    __syncthreads();
    double t[W];
#pragma unroll
    for (int c = 0; c < W; c++) t[c] = 0.0;
#pragma unroll
    for (int w = 0; w < kT / 64; w++)
#pragma unroll
        for (int c = 0; c < W; c++) t[c] += wp[w][c];
#pragma unroll
    for (int c = 0; c < W; c++) v[c] = t[c];
}

// PCG with M^-1 = diag(d) + U U^T (d = NULL: identity diagonal; U = NULL: no low-rank part),
// one workgroup per system, three barriers per iteration (four with a low-rank part).  Matrix-vector product: thread t
// takes row t % P and the column segment t / P (kT / P segments), its A entries cached in
// registers and LDS at the start; the segments are added in a fixed order, and p^T A p comes
// from the same partial sums (sum over threads of p_row * partial) with no extra pass.  One
// more reduction gives |r|^2, r^T D r and w = U^T r together, so z = D r + U w and
// r^T z = r^T D r + |w|^2.
__global__ __launch_bounds__(kT) void pcg_f64_kernel(const double *Aall, const double *ball, double *xall, int P,
                                                     double lambda, const double *dall, const double *Uall, int r,
                                                     double tol, int maxit, int *iters, double *relres) {
    const double *A = Aall + (size_t)blockIdx.x * P * P;
    const double *b = ball + (size_t)blockIdx.x * P;
    double *x = xall + (size_t)blockIdx.x * P;
    const double *d = dall ? dall + (size_t)blockIdx.x * P : nullptr;
    const double *U = Uall ? Uall + (size_t)blockIdx.x * P * r : nullptr;
    __shared__ double sp[1024], seg_sum[kT], wbuf[2][kT / 64][8], tot[72], sd[1024];
    __shared__ double sa[kPcgLds * kT], su[kPcgUCache];
    const int i = threadIdx.x;
    const bool own = i < P;
    const int rk = (U && r > 0) ? r : 0;
    const int seg = kT / P;
    const int row = i % P, sg = i / P;
    const int seg_len = (P + seg - 1) / seg;
    const int j0 = min(P, sg * seg_len);
    const int len = sg < seg ? min(P, j0 + seg_len) - j0 : 0;
    double areg[kPcgRegs];
#pragma unroll
    for (int k = 0; k < kPcgRegs; k++) areg[k] = (k < len) ? A[(size_t)(j0 + k) * P + row] : 0.0;
    for (int k = kPcgRegs; k < min(len, kPcgRegs + kPcgLds); k++)
        sa[(k - kPcgRegs) * kT + i] = A[(size_t)(j0 + k) * P + row];
    const bool u_lds = rk > 0 && P * rk <= kPcgUCache;
    if (u_lds)
        for (int e = i; e < P * rk; e += kT) su[(e % rk) * P + e / rk] = U[e];
    const double di = own ? (d ? d[i] : 1.0) : 0.0;
    if (own) sd[i] = di;
    __syncthreads();
    auto uval = [&](int c) -> double { return u_lds ? su[c * P + i] : U[(size_t)i * rk + c]; };
    int ph = 0;
    double xi = 0.0, ri = own ? b[i] : 0.0;
    // Reduction of r: returns |r|^2, sets z = D r + U (U^T r) and rz = r^T z.
    auto reduce_r = [&](double &z, double &rz) -> double {
        double rr;
        z = di * ri;
        if (rk == 0) {
            double v[2] = {ri * ri, di * ri * ri};
            block_sums<2>(v, wbuf[ph]);
            ph ^= 1;
            rr = v[0];
            rz = v[1];
        } else {
            // the 2 + rk dot products r.r, r.(D r), (U^T r)_c share the waves: wave c % 16 takes
            // value c as a lane-strided sum over r in LDS and a wave sum.  sp is free here: the
            // matrix-vector reads of p ended at the previous barrier.
            const int lane = i & 63, wv = i >> 6;
            if (own) sp[i] = ri;
            __syncthreads();
            for (int c = wv; c < 2 + rk; c += kT / 64) {
                double acc = 0.0;
                for (int k = lane; k < P; k += 64) {
                    const double rv = sp[k];
                    const double f = (c == 0) ? rv
                                   : (c == 1) ? sd[k] * rv
                                   : (u_lds ? su[(c - 2) * P + k] : U[(size_t)k * rk + c - 2]);
                    acc += f * rv;
                }
                acc = wave_sum_f64(acc);
                if (lane == 0) tot[c] = acc;
            }
            __syncthreads();
            rr = tot[0];
            rz = tot[1];
            for (int c = 0; c < rk; c++) {
                const double wc = tot[2 + c];
                if (own) z += uval(c) * wc;
                rz += wc * wc;
            }
        }
        if (!own) z = 0.0;
        return rr;
    };
    double zi, rz;
    const double bn = sqrt(reduce_r(zi, rz));
    if (bn == 0.0) {
        if (own) x[i] = 0.0;
        if (i == 0) { iters[blockIdx.x] = 0; relres[blockIdx.x] = 0.0; }
        return;
    }
    double pi = zi;
    int it = 0;
    double rel = 1.0;
    for (; it < maxit; it++) {
        if (own) sp[i] = pi;
        __syncthreads();
        double partial = 0.0;
#pragma unroll
        for (int k = 0; k < kPcgRegs; k++)
            if (k < len) partial += areg[k] * sp[j0 + k];
        for (int k = kPcgRegs; k < min(len, kPcgRegs + kPcgLds); k++) partial += sa[(k - kPcgRegs) * kT + i] * sp[j0 + k];
        for (int k = kPcgRegs + kPcgLds; k < len; k++) partial += A[(size_t)(j0 + k) * P + row] * sp[j0 + k];
        seg_sum[i] = partial;
        // p^T (A + lambda I) p = sum over threads of p_row * partial, plus lambda |p|^2
        double v[1] = {(len > 0 ? partial * sp[row] : 0.0) + (own ? lambda * pi * pi : 0.0)};
        block_sums<1>(v, wbuf[ph]);                   // its barrier also publishes seg_sum
        ph ^= 1;
        double api = 0.0;
        if (own) {
            for (int q = 0; q < seg; q++) api += seg_sum[q * P + i];
            api += lambda * pi;
        }
        const double alpha = rz / v[0];
        xi += alpha * pi;
        ri -= alpha * api;
        double rz_new;
        rel = sqrt(reduce_r(zi, rz_new)) / bn;
        if (rel <= tol) { it++; break; }
        const double beta = rz_new / rz;
        rz = rz_new;
        pi = zi + beta * pi;
    }
    if (own) x[i] = xi;
    if (i == 0) { iters[blockIdx.x] = it; relres[blockIdx.x] = rel; }
}

// B = M^H M for the singular values of a general complex matrix
__global__ void gram_small_c128_kernel(const double2 *M, double2 *B, int P) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < P * P; e += gridDim.x * blockDim.x) {
        const int i = e / P, j = e % P;
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < P; k++) {
            const double2 a = M[(size_t)k * P + i], c = M[(size_t)k * P + j];
            sr += a.x * c.x + a.y * c.y;
            si += a.x * c.y - a.y * c.x;
        }
        B[e] = make_double2(sr, si);
    }
}

// x = V diag(1/ev_i for |ev_i| > cutoff, else 0) V^H g   (one workgroup)
__global__ __launch_bounds__(kT) void pinv_apply_c128_kernel(const double2 *V, const double *ev, const double2 *g,
                                                             double2 *x, int P, double cutoff) {
    __shared__ double2 t[1024];
    for (int i = threadIdx.x; i < P; i += kT) {
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < P; k++) {
            const double2 v = V[(size_t)k * P + i], gk = g[k];
            sr += v.x * gk.x + v.y * gk.y;       // conj(v) g
            si += v.x * gk.y - v.y * gk.x;
        }
        const double e = ev[i];
        t[i] = (fabs(e) > cutoff) ? make_double2(sr / e, si / e) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < P; k += kT) {
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < P; i++) {
            const double2 p = cmul(V[(size_t)k * P + i], t[i]);
            sr += p.x;
            si += p.y;
        }
        x[k] = make_double2(sr, si);
    }
}

// ---- device-resident natural-gradient step ----
// A[k][l] = Re Q[s(k)][s(l)] + lambda delta_kl (Q: P x P complex in D-slot order),
// D[k] = A[k][k], g[k] = eg[1 + s(k)]: parameter order from slot order.
__global__ void ng_assemble_kernel(const double *__restrict__ Q, const int *__restrict__ slot, int P, double lambda,
                                   const double *__restrict__ eg, double *A, double *D, double *g) {
    const size_t total = (size_t)P * P;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(idx / P), l = (int)(idx % P);
        double v = Q[2 * ((size_t)slot[k] * P + slot[l])];
        if (k == l) {
            v += lambda;
            D[k] = v;
            g[k] = eg[1 + slot[k]];
        }
        A[idx] = v;
    }
}

// learned-preconditioner features: sin theta, cos theta, log D, field
__global__ void pc_features_kernel(const double *__restrict__ theta, const double *__restrict__ D, double field,
                                   int P, float *x) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= P; k += gridDim.x * blockDim.x) {
        if (k == P) {
            x[3 * P] = (float)field;
        } else {
            x[k] = (float)sin(theta[k]);
            x[P + k] = (float)cos(theta[k]);
            x[2 * P + k] = (float)log(D[k]);
        }
    }
}

// network output y (P (1 + r)) -> d = e^s / D, U = u / sqrt(P D)
__global__ void pc_transform_kernel(const float *__restrict__ y, const double *__restrict__ D, int P, int r,
                                    double *d, double *U) {
    const int st = 1 + r;
    const double us = 1.0 / sqrt((double)P);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) {
        d[k] = exp((double)y[(size_t)k * st]) / D[k];
        const double sd = us / sqrt(D[k]);
        for (int j = 0; j < r; j++) U[(size_t)k * r + j] = (double)y[(size_t)k * st + 1 + j] * sd;
    }
}

__global__ void reciprocal_kernel(const double *__restrict__ D, int P, double *d) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < P; k += gridDim.x * blockDim.x) d[k] = 1.0 / D[k];
}

}  // namespace

extern "C" {

int qgtl_hip_ng_assemble(const double *Q, const int *slot, int P, double lambda, const double *eg, double *A,
                         double *D, double *g) {
    const size_t total = (size_t)P * P;
    unsigned grid = (unsigned)((total + 255) / 256);
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(ng_assemble_kernel, dim3(grid), dim3(256), 0, qgtl::stream(), Q, slot, P, lambda, eg, A, D, g);
    QGTL_LAUNCH_CHECK("ng_assemble");
    return 0;
}

int qgtl_hip_pc_features(const double *theta, const double *D, double field, int P, float *x) {
    hipLaunchKernelGGL(pc_features_kernel, dim3((unsigned)((P + 256) / 256)), dim3(256), 0, qgtl::stream(), theta, D,
                       field, P, x);
    QGTL_LAUNCH_CHECK("pc_features");
    return 0;
}

int qgtl_hip_pc_transform(const float *y, const double *D, int P, int r, double *d, double *U) {
    hipLaunchKernelGGL(pc_transform_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, qgtl::stream(), y, D, P,
                       r, d, U);
    QGTL_LAUNCH_CHECK("pc_transform");
    return 0;
}

int qgtl_hip_reciprocal(const double *D, int P, double *d) {
    hipLaunchKernelGGL(reciprocal_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, qgtl::stream(), D, P, d);
    QGTL_LAUNCH_CHECK("reciprocal");
    return 0;
}

int qgtl_hip_pinv_apply_c128(const void *V, const double *ev, const void *g, void *x, int P, double cutoff) {
    if (P > 1024) return qgtl::set_error_msg(-118, "pinv: P > 1024");
    hipLaunchKernelGGL(pinv_apply_c128_kernel, dim3(1), dim3(kT), 0, qgtl::stream(), (const double2 *)V, ev,
                       (const double2 *)g, (double2 *)x, P, cutoff);
    QGTL_LAUNCH_CHECK("pinv_apply_c128");
    return 0;
}

int qgtl_hip_lu_solve_c128(void *A, const void *b, void *x, int *piv, int P, int batch, int *status) {
    hipLaunchKernelGGL(lu_solve_c128_kernel, dim3(batch), dim3(kT), 0, qgtl::stream(), (double2 *)A,
                       (const double2 *)b, (double2 *)x, piv, P, status);
    QGTL_LAUNCH_CHECK("lu_solve_c128");
    return 0;
}

int qgtl_hip_herm_eig_c128(void *A, void *V, double *ev, int P, int batch, int max_sweeps, int *sweeps) {
    if (P > 1024) return qgtl::set_error_msg(-118, "herm_eig: P > 1024");
    hipLaunchKernelGGL(herm_eig_c128_kernel, dim3(batch), dim3(kT), 0, qgtl::stream(), (double2 *)A, (double2 *)V,
                       ev, P, max_sweeps, sweeps);
    QGTL_LAUNCH_CHECK("herm_eig_c128");
    return 0;
}

int qgtl_hip_cholesky_solve_f64(double *A, const double *b, double *x, int P, int batch, double lambda, int *status) {
    hipLaunchKernelGGL(cholesky_f64_kernel, dim3(batch), dim3(kT), 0, qgtl::stream(), A, b, x, P, lambda, status);
    QGTL_LAUNCH_CHECK("cholesky_f64");
    return 0;
}

int qgtl_hip_pcg_f64(const double *A, const double *b, double *x, int P, int batch, double lambda, const double *d,
                     const double *U, int r, double tol, int maxit, int *iters, double *relres) {
    if (P > 1024 || r > 64) return qgtl::set_error_msg(-118, "pcg: P > 1024 or rank > 64");
    hipLaunchKernelGGL(pcg_f64_kernel, dim3(batch), dim3(kT), 0, qgtl::stream(), A, b, x, P, lambda, d, U, r, tol,
                       maxit, iters, relres);
    QGTL_LAUNCH_CHECK("pcg_f64");
    return 0;
}

int qgtl_hip_gram_small_c128(const void *M, void *B, int P) {
    const unsigned grid = (unsigned)((P * P + 255) / 256);
    hipLaunchKernelGGL(gram_small_c128_kernel, dim3(grid), dim3(256), 0, qgtl::stream(), (const double2 *)M,
                       (double2 *)B, P);
    QGTL_LAUNCH_CHECK("gram_small_c128");
    return 0;
}

}  // extern "C"
