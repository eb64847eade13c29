// Complex Gram G = D^H D of the derivative states on MFMA with split-BF16
// operands and FP32 accumulation.
//
// D holds P complex64 states of N amplitudes (state s at D + s*N).  The real
// matrix Z has rows 2s = Re d_s and 2s+1 = Im d_s, so C = Z Z^T gives
//   Re G_kl = C[2k][2l] + C[2k+1][2l+1],  Im G_kl = C[2k][2l+1] - C[2k+1][2l].
// Each fp32 value x is split into bf16 pieces and a product keeps the terms
// above a cut (v_mfma_f32_32x32x16_bf16, fp32 accumulation):
//   TERMS = 3: x = h + l (l = rn(x - h)), keep hh + hl + lh;
//   TERMS = 4: the same split, keep hh + hl + lh + ll;
//   TERMS = 6: x = h + m + l exactly (3 x 8 = 24 significand bits), keep
//              hh + hm + mh + hl + lh + mm (fp32-level products);
//   TERMS = 33 / 34: fp16 pieces of 2^15 x (11 significand bits each, so
//              h + l carries 22 bits; |x| <= 1 keeps 2^15 x below the fp16
//              maximum), keep hh + hl + lh (+ ll), v_mfma_f32_32x32x16_f16.
//              The representation error is <= 2^-22 |x| per value, so a
//              product sum is off by <= 2^-21 sum |x y| even when the errors
//              are fully correlated; the 2^30 scale of C cancels in the
//              normalisation of the finalize step.
// Accumulation is two-level: the MFMA accumulators cover kInner amplitudes
// and are then added into an fp32 outer sum, so rounding does not grow with
// the chunk length.  K (the amplitude index) is split into chunks; each
// workgroup writes an fp32 partial 128x128 tile of C and a second kernel sums
// the chunks in FP64 in a fixed order.  Chunk c's tiles all run on XCD c % 8
// so the chunk is fetched into one L2 only.
//
// QGTL counterpart: the per-element overlap loop
// src/quantum_geometric/core/quantum_geometric_tensor_network.c:1127-1137 and
// its FP64 form src/quantum_geometric/distributed/differential_geometry.c:2838-2842.
#include "qgtl_hip_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr int kTile = 128;          // output tile (rows of Z)
constexpr int kKS = QGTL_GRAM_KSTEP;  // amplitudes per LDS stage
constexpr int kRow = kKS + 8;          // 16-bit pieces per LDS row (16-B aligned, spreads banks)
constexpr int kXcd = 8;
constexpr int kInner = 512;       // amplitudes per MFMA accumulation before the fp32 outer add

__device__ __forceinline__ uint32_t bf16_rn_bits(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ void split2(float a, float b, uint32_t &hi, uint32_t &lo) {
    const uint32_t ha = bf16_rn_bits(a), hb = bf16_rn_bits(b);
    const float ra = a - __uint_as_float(ha << 16), rb = b - __uint_as_float(hb << 16);
    hi = ha | (hb << 16);
    lo = bf16_rn_bits(ra) | (bf16_rn_bits(rb) << 16);
}

constexpr float kF16Scale = 32768.0f;

constexpr bool is_f16(int terms) { return terms >= 30; }
constexpr int product_terms(int terms) { return terms >= 30 ? terms - 30 : terms; }

template <bool F16> struct Frag;
template <> struct Frag<false> {
    typedef bf16x8 type;
    static __device__ __forceinline__ f32x16 mma(type a, type b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct Frag<true> {
    typedef f16x8 type;
    static __device__ __forceinline__ f32x16 mma(type a, type b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ void split3(float a, float b, uint32_t &hi, uint32_t &mid, uint32_t &lo) {
    const uint32_t ha = bf16_rn_bits(a), hb = bf16_rn_bits(b);
    const float ra = a - __uint_as_float(ha << 16), rb = b - __uint_as_float(hb << 16);
    const uint32_t ma = bf16_rn_bits(ra), mb = bf16_rn_bits(rb);
    const float sa = ra - __uint_as_float(ma << 16), sb = rb - __uint_as_float(mb << 16);
    hi = ha | (hb << 16);
    mid = ma | (mb << 16);
    lo = bf16_rn_bits(sa) | (bf16_rn_bits(sb) << 16);
}

__device__ __forceinline__ int upper_tiles(int nt) { return nt * (nt + 1) / 2; }

__device__ __forceinline__ void tile_coords(int t, int nt, int &ti, int &tj) {
    int i = 0, rem = t;
    while (rem >= nt - i) { rem -= nt - i; i++; }
    ti = i;
    tj = i + rem;
}

constexpr int kStageLoads = 64 * (kKS / 2) / kThreads;   // float4 loads per thread for 64 states

// Load 64 states x kKS amplitudes of D starting at state s0, amplitude k0 into
// registers.  Every load is issued unconditionally (rows past P and amplitudes
// past N read a clamped, valid address) and bit i of the returned mask says
// whether load i is real; the zeroing happens at the split, so the loads of a
// stage are all in flight together instead of each waiting in its own branch.
// Load i of a thread reads state s0 + 8 i + t / 32, amplitudes k0 + 2 (t % 32)
// + {0, 1}.
__device__ __forceinline__ uint32_t load_rows(const float *__restrict__ D, int P, size_t N, int s0, size_t k0,
                                              float4 (&v)[kStageLoads]) {
    constexpr int kPairs = kKS / 2;                 // amplitude pairs per row
    constexpr int kRowsPerLoad = kThreads / kPairs;   // 8 states per load instruction
    const int sl = threadIdx.x / kPairs, pair = threadIdx.x % kPairs;
    const bool k_ok = k0 + 2 * pair < N;
    const size_t kk = k_ok ? k0 + 2 * pair : k0;
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < kStageLoads; i++) {
        const int s = s0 + i * kRowsPerLoad + sl;
        const bool ok = s < P && k_ok;
        mask |= (ok ? 1u : 0u) << i;
        const int sc = s < P ? s : P - 1;
        v[i] = *reinterpret_cast<const float4 *>(D + 2 * ((size_t)sc * N + kk));
    }
    return mask;
}

// Split the loaded values into LDS rows (2s, 2s+1) = (Re, Im) as S 16-bit
// pieces (S = 2 or 3 arrays); loads outside the matrix count as zeros.
template <int S, bool F16>
__device__ __forceinline__ void store_rows(const float4 (&vin)[kStageLoads], uint32_t mask, uint32_t *const *rows) {
    constexpr int kPairs = kKS / 2;
#pragma unroll
    for (int i = 0; i < kStageLoads; i++) {
        const int L = i * kThreads + threadIdx.x;
        const int sl = L / kPairs, pair = L % kPairs;
        const int wr = (2 * sl) * (kRow / 2) + pair, wi = (2 * sl + 1) * (kRow / 2) + pair;
        const bool ok = (mask >> i) & 1u;
        if (S == 2 && F16) {
            // the 2^15 scale and the zeroing of loads outside the matrix in one multiply
            const float sc = ok ? kF16Scale : 0.f;
            uint32_t hr, lr, hm, lm;
            qgtl::split_f16_pair(vin[i].x * sc, vin[i].z * sc, hr, lr);  // Re of amplitudes 2pair, 2pair+1
            qgtl::split_f16_pair(vin[i].y * sc, vin[i].w * sc, hm, lm);  // Im
            rows[0][wr] = hr; rows[1][wr] = lr;
            rows[0][wi] = hm; rows[1][wi] = lm;
            continue;
        }
        const float4 v = ok ? vin[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (S == 2) {
            uint32_t hr, lr, hm, lm;
            split2(v.x, v.z, hr, lr);  // Re of amplitudes 2pair, 2pair+1
            split2(v.y, v.w, hm, lm);  // Im
            rows[0][wr] = hr; rows[1][wr] = lr;
            rows[0][wi] = hm; rows[1][wi] = lm;
        } else {
            uint32_t a0, a1, a2, b0, b1, b2;
            split3(v.x, v.z, a0, a1, a2);
            split3(v.y, v.w, b0, b1, b2);
            rows[0][wr] = a0; rows[1][wr] = a1; rows[2][wr] = a2;
            rows[0][wi] = b0; rows[1][wi] = b1; rows[2][wi] = b2;
        }
    }
}

template <int TERMS>
struct __attribute__((aligned(16))) GramLds {
    static constexpr int S = (TERMS == 6) ? 3 : 2;     // bf16 pieces per value
    uint32_t rows[2 * S][kTile * kRow / 2];            // A pieces, then B pieces
};

// Unit `bid` of one Gram: chunk c's tiles are the units with bid % 8 == c % 8.
// ONE_ROUND: a stage issues the A and B loads together (one memory latency per
// stage); otherwise A, then B (two latencies, 32 fewer registers).
template <int TERMS, bool ONE_ROUND>
__device__ __forceinline__ void gram_unit(const float *__restrict__ D, int P, size_t N, int nt, int nchunks,
                                          size_t chunk_len, float *__restrict__ partials, int bid,
                                          GramLds<TERMS> &sm) {
    constexpr int S = GramLds<TERMS>::S;
    constexpr bool F16 = is_f16(TERMS);
    constexpr int PT = product_terms(TERMS);
    typedef typename Frag<F16>::type frag_t;
    auto &lds = sm.rows;

    const int T = upper_tiles(nt);
    const int xcd = bid % kXcd, slot = bid / kXcd;
    const int tile = slot % T, chunk = (slot / T) * kXcd + xcd;
    if (chunk >= nchunks) return;
    int ti, tj;
    tile_coords(tile, nt, ti, tj);
    const bool diag = (ti == tj);

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // on a diagonal tile the quadrant below the diagonal is the transpose of
    // the one above it and nothing reads it: its wave only loads and syncs;
    // the waves on the diagonal skip their 32x32 block below it the same way
    // (the only below-diagonal entries read, C[2k+1][2k], sit in diagonal blocks)
    const bool mirror = diag && wm > wn;
    const bool lower_sub = diag && wm == wn;

    f32x16 acc[2][2], outer[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int e = 0; e < 16; e++) { acc[a][b][e] = 0.f; outer[a][b][e] = 0.f; }

    const size_t kbeg = (size_t)chunk * chunk_len;
    const size_t kend = (kbeg + chunk_len < N) ? kbeg + chunk_len : N;
    uint32_t *arows[S], *brows[S];
#pragma unroll
    for (int p = 0; p < S; p++) {
        arows[p] = lds[p];
        brows[p] = diag ? lds[p] : lds[S + p];
    }

    for (size_t kb = kbeg; kb < kend; kb += kInner) {
        const size_t ke = (kb + kInner < kend) ? kb + kInner : kend;
        for (size_t k0 = kb; k0 < ke; k0 += kKS) {
            if (ONE_ROUND) {
                float4 va[kStageLoads], vb[kStageLoads];
                const uint32_t ma = load_rows(D, P, N, ti * (kTile / 2), k0, va);
                const uint32_t mb = diag ? 0u : load_rows(D, P, N, tj * (kTile / 2), k0, vb);
                store_rows<S, F16>(va, ma, arows);
                if (!diag) store_rows<S, F16>(vb, mb, brows);
            } else {
                {
                    float4 va[kStageLoads];
                    const uint32_t ma = load_rows(D, P, N, ti * (kTile / 2), k0, va);
                    store_rows<S, F16>(va, ma, arows);
                }
                if (!diag) {
                    float4 vb[kStageLoads];
                    const uint32_t mb = load_rows(D, P, N, tj * (kTile / 2), k0, vb);
                    store_rows<S, F16>(vb, mb, brows);
                }
            }
            __syncthreads();
            // unrolled: the next step's fragment reads issue under this step's MFMAs
#pragma unroll
            for (int kk = 0; kk < kKS / 16; kk++) {
                if (mirror) break;
                const int koff = kk * 16 + 8 * (lane >> 5);
                frag_t fa[S][2], fb[S][2];
#pragma unroll
                for (int p = 0; p < S; p++)
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        const int ra = wm * 64 + b * 32 + (lane & 31);
                        const int rb = wn * 64 + b * 32 + (lane & 31);
                        fa[p][b] = *reinterpret_cast<const frag_t *>(reinterpret_cast<const uint16_t *>(arows[p]) + ra * kRow + koff);
                        fb[p][b] = *reinterpret_cast<const frag_t *>(reinterpret_cast<const uint16_t *>(brows[p]) + rb * kRow + koff);
                    }
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++) {
                        if (a > b && lower_sub) continue;
                        // smallest terms first
                        if (PT == 6) {
                            acc[a][b] = Frag<F16>::mma(fa[1][a], fb[1][b], acc[a][b]);
                            acc[a][b] = Frag<F16>::mma(fa[2 % S][a], fb[0][b], acc[a][b]);
                            acc[a][b] = Frag<F16>::mma(fa[0][a], fb[2 % S][b], acc[a][b]);
                        }
                        if (PT == 4)
                            acc[a][b] = Frag<F16>::mma(fa[1][a], fb[1][b], acc[a][b]);
                        acc[a][b] = Frag<F16>::mma(fa[1][a], fb[0][b], acc[a][b]);
                        acc[a][b] = Frag<F16>::mma(fa[0][a], fb[1][b], acc[a][b]);
                        acc[a][b] = Frag<F16>::mma(fa[0][a], fb[0][b], acc[a][b]);
                    }
            }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                outer[a][b] += acc[a][b];
#pragma unroll
                for (int e = 0; e < 16; e++) acc[a][b][e] = 0.f;
            }
    }

    float *out = partials + ((size_t)chunk * T + tile) * (kTile * kTile);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = wm * 64 + a * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
                const int col = wn * 64 + b * 32 + (lane & 31);
                out[row * kTile + col] = outer[a][b][e];
            }
        }
}

template <int TERMS>
__global__ __launch_bounds__(kThreads, (TERMS == 6) ? 1 : 2) void gram_split_kernel(const float *__restrict__ D, int P, size_t N,
                                                              int nt, int nchunks, size_t chunk_len,
                                                              float *__restrict__ partials) {
    __shared__ GramLds<TERMS> sm;
    gram_unit<TERMS, TERMS != 6>(D, P, N, nt, nchunks, chunk_len, partials, blockIdx.x, sm);
}

// Grouped: grid and every unit_begin are multiples of 8, so u % 8 == local % 8
// and a chunk's tiles still share one XCD.
template <int TERMS>
__global__ __launch_bounds__(kThreads, (TERMS == 6) ? 1 : 2) void gram_grouped_kernel(const qgtl_gram_item_t *__restrict__ items,
                                                                int nitems, uint32_t units) {
    __shared__ GramLds<TERMS> sm;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_gram_item_t it = items[qgtl::find_item(items, nitems, u)];
        __syncthreads();
        gram_unit<TERMS, false>(qgtl::global_ptr<const float>(qgtl::uniform(it.D)), qgtl::uniform(it.P),
                         (size_t)qgtl::uniform(it.N), qgtl::uniform(it.nt), qgtl::uniform(it.nchunks),
                         (size_t)qgtl::uniform(it.chunk_len), qgtl::global_ptr<float>(qgtl::uniform(it.partials)),
                         (int)(u - qgtl::uniform(it.unit_begin)), sm);
    }
}

// C[ti*128 + r][tj*128 + c] = sum over chunks (fixed order, FP64); C is R x R.
__device__ __forceinline__ void gram_reduce_elem(const float *__restrict__ partials, int nt, int nchunks,
                                                 double *__restrict__ C, size_t idx) {
    const int T = upper_tiles(nt);
    const int R = nt * kTile;
    const int tile = (int)(idx / (kTile * kTile));
    const int e = (int)(idx % (kTile * kTile));
    // nchunks is a multiple of 8: eight loads in flight, then the adds in chunk order
    const float *q = partials + (size_t)tile * (kTile * kTile) + e;
    const size_t cs = (size_t)T * (kTile * kTile);
    double s = 0.0;
    for (int c = 0; c < nchunks; c += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = q[(size_t)(c + k) * cs];
#pragma unroll
        for (int k = 0; k < 8; k++) s += (double)v[k];
    }
    int ti, tj;
    tile_coords(tile, nt, ti, tj);
    C[(size_t)(ti * kTile + e / kTile) * R + tj * kTile + (e % kTile)] = s;
}

__global__ __launch_bounds__(kThreads) void gram_reduce_kernel(const float *__restrict__ partials, int nt,
                                                               int nchunks, double *__restrict__ C) {
    const size_t total = (size_t)upper_tiles(nt) * kTile * kTile;
    for (size_t idx = blockIdx.x * (size_t)kThreads + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * kThreads)
        gram_reduce_elem(partials, nt, nchunks, C, idx);
}

__global__ __launch_bounds__(kThreads) void gram_reduce_grouped_kernel(const qgtl_gram_reduce_item_t *__restrict__ items,
                                                                       int nitems, uint32_t units) {
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_gram_reduce_item_t it = items[qgtl::find_item(items, nitems, u)];
        const size_t idx = (size_t)(u - it.unit_begin) * kThreads + threadIdx.x;
        gram_reduce_elem(reinterpret_cast<const float *>(it.partials), it.nt, it.nchunks,
                         reinterpret_cast<double *>(it.C), idx);
    }
}

// Q_kl = G_kl / (4 sqrt(G_kk G_ll)) - a_k conj(a_l), a_k = <d_k|psi> / (2 sqrt(|d_k|^2 |psi|^2)).
// The normalisation applies the exact identities ||d_k psi|| = 1/2 and
// ||psi|| = 1 to remove the norm drift of the complex64 sweeps.
__device__ __forceinline__ void qgt_finalize_elem(const double *__restrict__ C, int R,
                                                  const double *__restrict__ inj, int P,
                                                  double *__restrict__ Q, size_t idx) {
    {
        const int k = (int)(idx / P), l = (int)(idx % P);
        if (k > l) return;
        const double gre = C[(size_t)(2 * k) * R + 2 * l] + C[(size_t)(2 * k + 1) * R + 2 * l + 1];
        const double gim = C[(size_t)(2 * k) * R + 2 * l + 1] - C[(size_t)(2 * k + 1) * R + 2 * l];
        const double gkk = C[(size_t)(2 * k) * R + 2 * k] + C[(size_t)(2 * k + 1) * R + 2 * k + 1];
        const double gll = C[(size_t)(2 * l) * R + 2 * l] + C[(size_t)(2 * l + 1) * R + 2 * l + 1];
        const double sg = 0.25 / sqrt(gkk * gll);
        const double sk = 0.5 / sqrt(inj[4 * k + 2] * inj[4 * k + 3]);
        const double sl = 0.5 / sqrt(inj[4 * l + 2] * inj[4 * l + 3]);
        const double akr = inj[4 * k] * sk, aki = inj[4 * k + 1] * sk;
        const double alr = inj[4 * l] * sl, ali = inj[4 * l + 1] * sl;
        // a_k conj(a_l)
        const double pr = akr * alr + aki * ali, pi = aki * alr - akr * ali;
        const double qr = gre * sg - pr, qi = (k == l) ? 0.0 : gim * sg - pi;
        Q[2 * ((size_t)k * P + l)] = qr;
        Q[2 * ((size_t)k * P + l) + 1] = qi;
        if (k != l) {
            Q[2 * ((size_t)l * P + k)] = qr;
            Q[2 * ((size_t)l * P + k) + 1] = -qi;
        }
    }
}

// The same with the overlaps read from the Gram: the primal state at the
// states' meeting point is Gram state P (Z rows 2P, 2P+1), so
// <d_k|psi> = G_kP, |d_k|^2 = G_kk and |psi|^2 = G_PP.
__device__ __forceinline__ void qgt_finalize_psi_elem(const double *__restrict__ C, int R, int P,
                                                      double *__restrict__ Q, size_t idx) {
    const int k = (int)(idx / P), l = (int)(idx % P);
    if (k > l) return;
    const int m = P;   // the primal state
    const double gre = C[(size_t)(2 * k) * R + 2 * l] + C[(size_t)(2 * k + 1) * R + 2 * l + 1];
    const double gim = C[(size_t)(2 * k) * R + 2 * l + 1] - C[(size_t)(2 * k + 1) * R + 2 * l];
    const double gkk = C[(size_t)(2 * k) * R + 2 * k] + C[(size_t)(2 * k + 1) * R + 2 * k + 1];
    const double gll = C[(size_t)(2 * l) * R + 2 * l] + C[(size_t)(2 * l + 1) * R + 2 * l + 1];
    const double gmm = C[(size_t)(2 * m) * R + 2 * m] + C[(size_t)(2 * m + 1) * R + 2 * m + 1];
    const double akr0 = C[(size_t)(2 * k) * R + 2 * m] + C[(size_t)(2 * k + 1) * R + 2 * m + 1];
    const double aki0 = C[(size_t)(2 * k) * R + 2 * m + 1] - C[(size_t)(2 * k + 1) * R + 2 * m];
    const double alr0 = C[(size_t)(2 * l) * R + 2 * m] + C[(size_t)(2 * l + 1) * R + 2 * m + 1];
    const double ali0 = C[(size_t)(2 * l) * R + 2 * m + 1] - C[(size_t)(2 * l + 1) * R + 2 * m];
    const double sg = 0.25 / sqrt(gkk * gll);
    const double sk = 0.5 / sqrt(gkk * gmm), sl = 0.5 / sqrt(gll * gmm);
    const double akr = akr0 * sk, aki = aki0 * sk, alr = alr0 * sl, ali = ali0 * sl;
    const double pr = akr * alr + aki * ali, pi = aki * alr - akr * ali;
    const double qr = gre * sg - pr, qi = (k == l) ? 0.0 : gim * sg - pi;
    Q[2 * ((size_t)k * P + l)] = qr;
    Q[2 * ((size_t)k * P + l) + 1] = qi;
    if (k != l) {
        Q[2 * ((size_t)l * P + k)] = qr;
        Q[2 * ((size_t)l * P + k) + 1] = -qi;
    }
}

__global__ __launch_bounds__(kThreads) void qgt_finalize_psi_kernel(const double *__restrict__ C, int R, int P,
                                                                    double *__restrict__ Q) {
    for (size_t idx = blockIdx.x * (size_t)kThreads + threadIdx.x; idx < (size_t)P * P;
         idx += (size_t)gridDim.x * kThreads)
        qgt_finalize_psi_elem(C, R, P, Q, idx);
}

__global__ __launch_bounds__(kThreads) void qgt_finalize_kernel(const double *__restrict__ C, int R,
                                                                const double *__restrict__ inj, int P,
                                                                double *__restrict__ Q) {
    for (size_t idx = blockIdx.x * (size_t)kThreads + threadIdx.x; idx < (size_t)P * P;
         idx += (size_t)gridDim.x * kThreads)
        qgt_finalize_elem(C, R, inj, P, Q, idx);
}

__global__ __launch_bounds__(kThreads) void finalize_grouped_kernel(const qgtl_finalize_item_t *__restrict__ items,
                                                                    int nitems, uint32_t units) {
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_finalize_item_t it = items[qgtl::find_item(items, nitems, u)];
        const size_t idx = (size_t)(u - it.unit_begin) * kThreads + threadIdx.x;
        if (idx >= (size_t)it.P * it.P) continue;
        if (it.inj)             // overlaps and norms from FP64 injection sums
            qgt_finalize_elem(reinterpret_cast<const double *>(it.C), it.R, reinterpret_cast<const double *>(it.inj),
                              it.P, reinterpret_cast<double *>(it.Q), idx);
        else                    // ... from the Gram row of the primal state (state P)
            qgt_finalize_psi_elem(reinterpret_cast<const double *>(it.C), it.R, it.P, reinterpret_cast<double *>(it.Q),
                                  idx);
    }
}

// Lane layout probe of v_mfma_f32_32x32x16_bf16 with the operand layout the
// Gram assumes (lane l holds row/col l % 32, k = 8 (l / 32) + e):
// A[i][k] = (k == i % 16) ? 1 + i / 16 : 0 and B[k][j] = k + 16 (j % 8), so
// D[i][j] = (1 + i / 16) (i % 16 + 16 (j % 8)), all exact in bf16/fp32.
__global__ void mfma_layout_probe_kernel(float *out) {
    const int lane = threadIdx.x;
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int i = lane & 31, k = 8 * (lane >> 5) + e, j = lane & 31;
        a[e] = (__bf16)(float)((k == (i % 16)) ? 1 + i / 16 : 0);
        b[e] = (__bf16)(float)(k + 16 * (j % 8));
    }
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.f;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int row = (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3), col = lane & 31;
        out[row * 32 + col] = acc[e];
    }
}

}  // namespace

extern "C" {

int qgtl_hip_gram_tiles(int P) {
    const int nt = (2 * P + kTile - 1) / kTile;
    return nt * (nt + 1) / 2;
}

int qgtl_hip_gram_bf16x3(const void *D, int P, size_t N, float *partials, int nchunks, int terms) {
    if (P < 1 || nchunks < 1 || nchunks % kXcd != 0 || (N & 1))
        return qgtl::set_error_msg(-118, "qgtl_hip_gram_bf16x3: nchunks must be a multiple of 8 and N even");
    const int nt = (2 * P + kTile - 1) / kTile;
    const int T = nt * (nt + 1) / 2;
    const size_t chunk_len = ((N + nchunks - 1) / nchunks + kKS - 1) / kKS * kKS;
    const dim3 grid((unsigned)(T * nchunks)), block(kThreads);
    if (terms == 6)
        hipLaunchKernelGGL(gram_split_kernel<6>, grid, block, 0, qgtl::stream(), (const float *)D, P, N, nt, nchunks, chunk_len, partials);
    else if (terms == QGTL_GRAM_F16X3)
        hipLaunchKernelGGL(gram_split_kernel<QGTL_GRAM_F16X3>, grid, block, 0, qgtl::stream(), (const float *)D, P, N, nt, nchunks, chunk_len, partials);
    else if (terms == QGTL_GRAM_F16X4)
        hipLaunchKernelGGL(gram_split_kernel<QGTL_GRAM_F16X4>, grid, block, 0, qgtl::stream(), (const float *)D, P, N, nt, nchunks, chunk_len, partials);
    else if (terms == 4)
        hipLaunchKernelGGL(gram_split_kernel<4>, grid, block, 0, qgtl::stream(), (const float *)D, P, N, nt, nchunks, chunk_len, partials);
    else
        hipLaunchKernelGGL(gram_split_kernel<3>, grid, block, 0, qgtl::stream(), (const float *)D, P, N, nt, nchunks, chunk_len, partials);
    QGTL_LAUNCH_CHECK("gram_bf16x3");
    return 0;
}

int qgtl_hip_gram_reduce(const float *partials, int P, int nchunks, double *C) {
    if (P < 1 || nchunks < kXcd || nchunks % kXcd != 0)
        return qgtl::set_error_msg(-118, "qgtl_hip_gram_reduce: nchunks must be a positive multiple of 8");
    const int nt = (2 * P + kTile - 1) / kTile;
    const size_t total = (size_t)(nt * (nt + 1) / 2) * kTile * kTile;
    unsigned grid = (unsigned)((total + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(grid), dim3(kThreads), 0, qgtl::stream(), partials, nt, nchunks, C);
    QGTL_LAUNCH_CHECK("gram_reduce");
    return 0;
}

int qgtl_hip_qgt_finalize_psi(const double *C, int P, double *Q) {
    const int nt = (2 * (P + 1) + kTile - 1) / kTile;
    const size_t work = (size_t)P * P;
    unsigned grid = (unsigned)((work + kThreads - 1) / kThreads);
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(qgt_finalize_psi_kernel, dim3(grid), dim3(kThreads), 0, qgtl::stream(), C, nt * kTile, P, Q);
    QGTL_LAUNCH_CHECK("qgt_finalize_psi");
    return 0;
}

int qgtl_hip_qgt_finalize(const double *C, const double *inj, int P, double *Q) {
    const int nt = (2 * P + kTile - 1) / kTile;
    const size_t work = (size_t)P * P;
    unsigned grid = (unsigned)((work + kThreads - 1) / kThreads);
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(qgt_finalize_kernel, dim3(grid), dim3(kThreads), 0, qgtl::stream(), C, nt * kTile, inj, P, Q);
    QGTL_LAUNCH_CHECK("qgt_finalize");
    return 0;
}

static unsigned grouped_grid(uint32_t units, int grid) {
    if (grid <= 0 || (uint32_t)grid > units) return units;
    return (unsigned)grid;
}

int qgtl_hip_gram_grouped(const qgtl_gram_item_t *items, int nitems, uint32_t units, int grid, int terms) {
    if (!items || nitems < 1 || units < 1 || units % kXcd != 0)
        return qgtl::set_error_msg(-118, "qgtl_hip_gram_grouped: empty items or units not a multiple of 8");
    unsigned g = grouped_grid(units, grid);
    g = (g + kXcd - 1) / kXcd * kXcd;
    if (terms == 6)
        hipLaunchKernelGGL(gram_grouped_kernel<6>, dim3(g), dim3(kThreads), 0, qgtl::stream(), items, nitems, units);
    else if (terms == QGTL_GRAM_F16X3)
        hipLaunchKernelGGL(gram_grouped_kernel<QGTL_GRAM_F16X3>, dim3(g), dim3(kThreads), 0, qgtl::stream(), items, nitems, units);
    else if (terms == QGTL_GRAM_F16X4)
        hipLaunchKernelGGL(gram_grouped_kernel<QGTL_GRAM_F16X4>, dim3(g), dim3(kThreads), 0, qgtl::stream(), items, nitems, units);
    else if (terms == 4)
        hipLaunchKernelGGL(gram_grouped_kernel<4>, dim3(g), dim3(kThreads), 0, qgtl::stream(), items, nitems, units);
    else
        hipLaunchKernelGGL(gram_grouped_kernel<3>, dim3(g), dim3(kThreads), 0, qgtl::stream(), items, nitems, units);
    QGTL_LAUNCH_CHECK("gram_grouped");
    return 0;
}

int qgtl_hip_gram_reduce_grouped(const qgtl_gram_reduce_item_t *items, int nitems, uint32_t units, int grid) {
    if (!items || nitems < 1 || units < 1) return qgtl::set_error_msg(-118, "gram_reduce_grouped: empty item list");
    hipLaunchKernelGGL(gram_reduce_grouped_kernel, dim3(grouped_grid(units, grid)), dim3(kThreads), 0, qgtl::stream(),
                       items, nitems, units);
    QGTL_LAUNCH_CHECK("gram_reduce_grouped");
    return 0;
}

int qgtl_hip_finalize_grouped(const qgtl_finalize_item_t *items, int nitems, uint32_t units, int grid) {
    if (!items || nitems < 1 || units < 1) return qgtl::set_error_msg(-118, "finalize_grouped: empty item list");
    hipLaunchKernelGGL(finalize_grouped_kernel, dim3(grouped_grid(units, grid)), dim3(kThreads), 0, qgtl::stream(),
                       items, nitems, units);
    QGTL_LAUNCH_CHECK("finalize_grouped");
    return 0;
}

int qgtl_hip_debug_mfma_layout(float *out_dev) {
    hipLaunchKernelGGL(mfma_layout_probe_kernel, dim3(1), dim3(64), 0, qgtl::stream(), out_dev);
    QGTL_LAUNCH_CHECK("mfma_layout_probe");
    return 0;
}

}  // extern "C"
