// Complex Gram G = D^H D of derivative states in split-plane form on MFMA
// (v_mfma_f32_32x32x16_f16: fp16 pieces, fp32 accumulation), in complex form.
//
// A state's split words read as a row of fp16 values X_s = (Re a_0, Im a_0,
// Re a_1, Im a_1, ...) (one row for the hi pieces, one for the lo pieces), and
// Y_s = the same row with each pair turned into (Im a, -Re a), give
//   Re G_kl = sum_a Re_k Re_l + Im_k Im_l = X_k . X_l,
//   Im G_kl = sum_a Re_k Im_l - Im_k Re_l = X_k . Y_l,
// so one pass over a tile pair of 128 x 128 states yields both parts of G, and
// Y is formed in registers from the B operand (a half swap and a sign flip per
// 32-bit word): four products per pair of fragments read from LDS, where the
// real form C = Z Z^T (Z rows Re d_s, Im d_s) gets two.  Each product keeps
// hh + hl + lh of the fp16 pieces (PT = 3) or also ll (PT = 4).
//
// Work split: tile pairs (ti <= tj) of 128-state tiles times split-K chunks of
// amplitudes, one workgroup of 8 waves (two per SIMD, the whole LDS budget) per
// (tile pair, chunk); chunk c's tile pairs run on XCD c % 8, so they share one
// L2.  A wave owns a 64 x 32 strip (2 MFMA blocks of 32 states against one,
// Re and Im: 4 accumulators, which with the outer sums leaves the registers for
// two waves per SIMD); on a diagonal tile pair only the 10 block pairs on or
// above the diagonal are computed, spread so that no SIMD gets more than three
// of them.  Stages of 16 amplitudes go straight from global memory into LDS
// (global_load_lds_dwordx4) through four buffers: three stages' loads are in
// flight while the MFMAs of the fourth run (a stage's MFMAs are shorter than a
// load's latency), one barrier per stage and no fence on it.  The MFMA
// accumulators are added into an fp32 outer sum every 256 amplitudes, so
// rounding does not grow with the chunk length; the chunks' fp32 partials are
// summed in FP64 in a fixed order by the reduce kernel.  For 129 .. 256 states
// (two tiles) one workgroup per chunk computes all three tile pairs instead
// (gram_cplx256_kernel below), reading each row once.
//
// LDS row of a state and stage: 8 slots of 16 bytes, slots 0..3 the hi words of
// the 16 amplitudes (4 per slot), 4..7 the lo words, stored swizzled (swz).
//
// QGTL counterpart: the per-element overlap loop
// src/quantum_geometric/core/quantum_geometric_tensor_network.c:1127-1137 and
// its FP64 form src/quantum_geometric/distributed/differential_geometry.c:2838-2842.
#include <type_traits>

#include "qgtl_hip_common.h"

namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void gbl_void_t;

constexpr int kThreads = 512;                          // 8 waves
constexpr int kTileStates = 128;
constexpr int kTileElems = kTileStates * kTileStates;
constexpr int kLdsBytes = 128 * 1024;                  // stage buffers of one workgroup
constexpr int kInnerAmps = 256;                        // amplitudes per MFMA accumulation
constexpr int kXcd = 8;

// Stage geometry for SA amplitudes per stage: a state's row holds the hi words
// then the lo words of the SA amplitudes, S = SA / 2 slots of 16 bytes; the
// 128 KB of stage buffers hold BUFS stages of both tiles (16 amplitudes: 4
// buffers, three stages in flight; 32 amplitudes: 2, one in flight).
template <int SA>
struct Geo {
    static constexpr int kRowBytes = SA * 8;
    static constexpr int kSlots = SA / 2;
    static constexpr int kTileBytes = kTileStates * kRowBytes;
    static constexpr int kSteps = SA / 8;                      // MFMA K-steps (8 amplitudes = 16 fp16)
    static constexpr int kBufs = kLdsBytes / (2 * kTileBytes);
    static constexpr int kInner = kInnerAmps / SA;             // stages per MFMA accumulation
    static constexpr int kLoads = kTileBytes / 16 / kThreads;  // 16-byte loads per thread, tile and stage
    static constexpr int kRowsPerLoad = 1024 / kRowBytes;      // rows one wave load fills
    // stored position of logical slot c in row r is c ^ swz(r): the 8 rows a
    // 128-byte bank cycle serves (and the 16 of a 256-byte one) land in distinct
    // 16-byte bank groups
    static __device__ __forceinline__ uint32_t swz(uint32_t r) {
        return kSlots == 16 ? (r & 15) : ((r & 7) ^ ((r >> 3) & 1));
    }
};

__device__ __forceinline__ void tile_coords(int t, int nt, int &ti, int &tj) {
    int i = 0, rem = t;
    while (rem >= nt - i) { rem -= nt - i; i++; }
    ti = i;
    tj = i + rem;
}

// Y word of an X word (Re piece, Im piece) -> (Im piece, -Re piece)
__device__ __forceinline__ uint32_t swap_neg(uint32_t x) {
    return __builtin_amdgcn_alignbit(x, x, 16) ^ 0x80000000u;
}

__device__ __forceinline__ f16x8 as_f16(u32x4 v) { return __builtin_bit_cast(f16x8, v); }

__device__ __forceinline__ f32x16 mma(u32x4 a, u32x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(as_f16(a), as_f16(b), c, 0, 0, 0);
}

// The blocks (of 32 states, within the tile) a wave multiplies: A blocks a0, a1
// against B block b; which of the two pairs (a0, b), (a1, b) it computes (bits 0,
// 1) is the template argument MASK of the code the wave runs.  Waves w and w + 4
// share SIMD w % 4.  Off-diagonal tile pairs: A blocks 2 (w / 4), 2 (w / 4) + 1
// against B block w % 4 (four pairs per SIMD).  Diagonal tile pairs: the 10
// pairs (a, b), a <= b, as 2 + 1, 2 + 1, 2 + 0, 2 + 0 per SIMD.
struct Role {
    int a0, a1, b;
};

__device__ __forceinline__ Role wave_role(bool diag, int w) {
    if (!diag) return Role{2 * (w >> 2), 2 * (w >> 2) + 1, w & 3};
    switch (w) {
    case 0: return Role{0, 1, 3};        // MASK 3: (0,3) (1,3)
    case 1: return Role{2, 3, 3};        // MASK 3: (2,3) (3,3)
    case 2: return Role{0, 1, 2};        // MASK 3: (0,2) (1,2)
    case 3: return Role{0, 1, 1};        // MASK 3: (0,1) (1,1)
    case 4: return Role{0, 0, 0};        // MASK 1: (0,0)
    case 5: return Role{2, 2, 2};        // MASK 1: (2,2)
    default: return Role{0, 0, 0};       // MASK 0: loads and barriers only
    }
}

// fragment of block blk, piece p (0 hi, 1 lo), K-step t: lane l reads row
// 32 blk + l % 32, logical slot S/2 p + 2 t + l / 32
template <int SA>
__device__ __forceinline__ u32x4 frag(const uint8_t *tile, int blk, int p, int t, uint32_t lane_row, uint32_t lane_x,
                                      uint32_t half) {
    const uint32_t slot = ((uint32_t)(Geo<SA>::kSlots / 2 * p + 2 * t) + half) ^ lane_x;
    return *reinterpret_cast<const u32x4 *>(tile + blk * 32 * Geo<SA>::kRowBytes + lane_row + slot * 16);
}

struct Acc {
    f32x16 re[2], im[2];
};

// The K-steps of one stage.  Not unrolled: unrolling hoists the fragment reads
// of later steps and spills at the 256 registers of two waves per SIMD.
template <int SA, int PT, int MASK>
__device__ __forceinline__ void stage_mfma(const uint8_t *A, const uint8_t *B, const Role &ro, Acc &acc,
                                           uint32_t lane_row, uint32_t lane_x, uint32_t half) {
    if constexpr (MASK == 0) return;
#pragma unroll 1
    for (int t = 0; t < Geo<SA>::kSteps; t++) {
        u32x4 ah[2], al[2], bh, bl, yh, yl;
        ah[0] = frag<SA>(A, ro.a0, 0, t, lane_row, lane_x, half);
        al[0] = frag<SA>(A, ro.a0, 1, t, lane_row, lane_x, half);
        if constexpr ((MASK & 2) != 0) {
            ah[1] = frag<SA>(A, ro.a1, 0, t, lane_row, lane_x, half);
            al[1] = frag<SA>(A, ro.a1, 1, t, lane_row, lane_x, half);
        } else {
            ah[1] = ah[0];
            al[1] = al[0];
        }
        bh = frag<SA>(B, ro.b, 0, t, lane_row, lane_x, half);
        bl = frag<SA>(B, ro.b, 1, t, lane_row, lane_x, half);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            yh[e] = swap_neg(bh[e]);
            yl[e] = swap_neg(bl[e]);
        }
        // smallest terms first; the accumulators interleaved
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (!((MASK >> i) & 1)) continue;
            if (PT == 4) {
                acc.re[i] = mma(al[i], bl, acc.re[i]);
                acc.im[i] = mma(al[i], yl, acc.im[i]);
            }
            acc.re[i] = mma(al[i], bh, acc.re[i]);
            acc.im[i] = mma(al[i], yh, acc.im[i]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (!((MASK >> i) & 1)) continue;
            acc.re[i] = mma(ah[i], bl, acc.re[i]);
            acc.im[i] = mma(ah[i], yl, acc.im[i]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (!((MASK >> i) & 1)) continue;
            acc.re[i] = mma(ah[i], bh, acc.re[i]);
            acc.im[i] = mma(ah[i], yh, acc.im[i]);
        }
    }
}

__device__ __forceinline__ void flush(Acc &acc, Acc &outer) {
#pragma unroll
    for (int i = 0; i < 2; i++) {
        outer.re[i] += acc.re[i];
        outer.im[i] += acc.im[i];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            acc.re[i][e] = 0.f;
            acc.im[i][e] = 0.f;
        }
    }
}

// Issue the direct-to-LDS loads of one stage (amplitudes k0 .. k0 + SA - 1) of
// the 128 states from s0 into one LDS tile: load i of wave w fills R = 1024 /
// row-bytes rows, R (8 i + w) .. R (8 i + w) + R - 1 (1 KB); lane l takes row
// 8 R i + rr, rr = R w + l / S, and stored slot l % S, i.e. logical slot
// (l % S) ^ swz(rr).  States past P read state P - 1: their rows only feed G rows
// and columns >= P, which nothing reads.  The row index passes through an empty
// asm so that the addresses are formed at every stage instead of being hoisted
// out of the loop (where they would hold registers the accumulators need).
template <int SA>
__device__ __forceinline__ void issue_tile(const uint8_t *__restrict__ D, size_t SB, int P, int s0, size_t k0,
                                           uint8_t *lds, int w, int rr, uint32_t in_blk) {
    typedef Geo<SA> G;
    const size_t ko = (k0 >> 6) * 512 + (k0 & 63) * 4 + in_blk;
#pragma unroll
    for (int i = 0; i < G::kLoads; i++) {
        int st = s0 + rr + 8 * G::kRowsPerLoad * i;
        __asm__ volatile("" : "+v"(st));
        st = st < P ? st : P - 1;
        const uint8_t *src = D + (size_t)st * SB + ko;
        const uint32_t dst = (uint32_t)(uintptr_t)(lds_void_t *)(lds + (i * 8 + w) * 1024);
        // written out so that the compiler, which cannot tell which buffer a load
        // fills, inserts no wait for it before the other buffers' fragment reads
        // (the waits are wait_stage's); the instruction writes no register, and m0
        // is set here and read only by this instruction in the kernel
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                     :
                     : "s"(dst), "v"(src)
                     : "memory");
    }
}

// Wait until at most `ahead` (<= 2) stages' loads of this wave are in flight,
// then for every wave (the barrier carries no memory fence: nothing else needs
// one).
template <int LW>
__device__ __forceinline__ void wait_stage(int ahead) {
    if (ahead >= 2) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(2 * LW) : "memory");
    else if (ahead == 1) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(LW) : "memory");
    else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// One (tile pair, chunk) unit for a wave that computes the block pairs MASK;
// DIAG: A and B are the same tile, loaded once.  Stage s uses LDS buffer
// s % BUFS; its loads were issued BUFS - 1 stages earlier.
template <int SA, int PT, bool DIAG, int MASK>
__device__ __forceinline__ void gram_cplx_unit(const uint8_t *__restrict__ D, int P, size_t N, int T, int tile,
                                               int chunk, int ti, int tj, size_t chunk_len, int w,
                                               float *__restrict__ partials, uint8_t *lds) {
    typedef Geo<SA> G;
    constexpr int BUFS = G::kBufs;
    constexpr int LW = DIAG ? G::kLoads : 2 * G::kLoads;   // loads per wave and stage
    const uint32_t lane = threadIdx.x & 63;
    const Role ro = wave_role(DIAG, w);
    // load constants: row rr of each load's row group and the logical slot its stored slot holds
    const int rr = G::kRowsPerLoad * w + (int)(lane / G::kSlots);
    const uint32_t lgc = (lane % G::kSlots) ^ G::swz((uint32_t)rr);
    const uint32_t in_blk = (lgc / (G::kSlots / 2)) * 256 + (lgc % (G::kSlots / 2)) * 16;
    // fragment constants
    const uint32_t lane_row = (lane & 31) * G::kRowBytes, lane_x = G::swz(lane & 31), half = lane >> 5;

    const size_t SB = N * 8;
    const size_t kbeg = (size_t)chunk * chunk_len;
    const size_t kend = (kbeg + chunk_len < N) ? kbeg + chunk_len : N;
    const int nst = kend > kbeg ? (int)((kend - kbeg) / SA) : 0;
    const int sa = ti * kTileStates, sb = tj * kTileStates;

    Acc acc, outer;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int e = 0; e < 16; e++) {
            acc.re[i][e] = 0.f; acc.im[i][e] = 0.f;
            outer.re[i][e] = 0.f; outer.im[i][e] = 0.f;
        }

    auto bufA = [&](int b) { return lds + (size_t)b * 2 * G::kTileBytes; };
    auto bufB = [&](int b) { return lds + (size_t)b * 2 * G::kTileBytes + G::kTileBytes; };
    auto issue = [&](int st, int b) {
        const size_t k0 = kbeg + (size_t)st * SA;
        issue_tile<SA>(D, SB, P, sa, k0, bufA(b), w, rr, in_blk);
        if (!DIAG) issue_tile<SA>(D, SB, P, sb, k0, bufB(b), w, rr, in_blk);
    };
    // stage st in buffer B: wait for it, release the buffer of stage st - 1 for
    // stage st + BUFS - 1, multiply
    auto stage = [&](auto Bc, int st) {
        constexpr int B = decltype(Bc)::value;
        const int left = nst - 1 - st;
        wait_stage<LW>(left < BUFS - 2 ? left : BUFS - 2);
        if (st + BUFS - 1 < nst) issue(st + BUFS - 1, (B + BUFS - 1) % BUFS);
        stage_mfma<SA, PT, MASK>(bufA(B), DIAG ? bufA(B) : bufB(B), ro, acc, lane_row, lane_x, half);
        if (((st + 1) % G::kInner) == 0) flush(acc, outer);
    };
    for (int st = 0; st < BUFS - 1 && st < nst; st++) issue(st, st);
    for (int s = 0; s < nst; s += BUFS) {
        stage(std::integral_constant<int, 0>{}, s);
        if (s + 1 >= nst) break;
        stage(std::integral_constant<int, 1>{}, s + 1);
        if constexpr (BUFS == 4) {
            if (s + 2 >= nst) break;
            stage(std::integral_constant<int, 2 % BUFS>{}, s + 2);
            if (s + 3 >= nst) break;
            stage(std::integral_constant<int, 3 % BUFS>{}, s + 3);
        }
    }
    flush(acc, outer);

    // fp32 partials of Re G and Im G of this (chunk, tile pair); pairs no wave
    // computes stay unwritten (the reduce skips them)
    float *out = partials + ((size_t)chunk * T + tile) * 2 * kTileElems;
    const int ab[2] = {ro.a0, ro.a1};
#pragma unroll
    for (int i = 0; i < 2; i++) {
        if (!((MASK >> i) & 1)) continue;
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int row = ab[i] * 32 + (e >> 2) * 8 + (int)(lane >> 5) * 4 + (e & 3);
            const int col = ro.b * 32 + (int)(lane & 31);
            out[row * kTileStates + col] = outer.re[i][e];
            out[kTileElems + row * kTileStates + col] = outer.im[i][e];
        }
    }
}

template <int SA, int PT>
__global__ __launch_bounds__(kThreads, 1) void gram_cplx_kernel(const uint8_t *__restrict__ D, int P, size_t N,
                                                               int nt, int nchunks, size_t chunk_len,
                                                               float *__restrict__ partials) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[kLdsBytes];
    const int T = nt * (nt + 1) / 2;
    const int bid = blockIdx.x;
    const int xcd = bid % kXcd, slot = bid / kXcd;
    const int tile = slot % T, chunk = (slot / T) * kXcd + xcd;
    int ti, tj;
    tile_coords(tile, nt, ti, tj);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (ti != tj)
        gram_cplx_unit<SA, PT, false, 3>(D, P, N, T, tile, chunk, ti, tj, chunk_len, w, partials, lds);
    else if (w < 4)
        gram_cplx_unit<SA, PT, true, 3>(D, P, N, T, tile, chunk, ti, tj, chunk_len, w, partials, lds);
    else if (w < 6)
        gram_cplx_unit<SA, PT, true, 1>(D, P, N, T, tile, chunk, ti, tj, chunk_len, w, partials, lds);
    else
        gram_cplx_unit<SA, PT, true, 0>(D, P, N, T, tile, chunk, ti, tj, chunk_len, w, partials, lds);
}

// ---------------------------------------------------------------------------
// Whole-chunk form for 129 .. 256 states (two 128-state tiles): one workgroup
// of 12 waves per split-K chunk computes all three tile pairs, i.e. the 36
// block pairs (a, b) of 32-state blocks a <= b of the 256-state upper triangle,
// from one stage buffer holding both tiles -- every row of the chunk is read
// from HBM once (the per-tile-pair form reads each row tile twice, once per
// tile pair that holds it).  Wave w multiplies three A blocks against one B
// block (the B fragment and its Y form are shared); a pair with a > b yields
// G_ab = conj(G_ba) and is stored transposed with Im negated.  The assignment
// orients the 28 block pairs a != b so that blocks 0..3 serve as B for 2 of
// them and blocks 4..7 for 5 (with its diagonal pair: 3 and 6 = one and two
// waves), so every wave has three pairs and every SIMD nine.  A diagonal pair
// (always the wave's first, a0 == b) skips its 16 x 16 sub-block below the
// diagonal, which no consumer reads; waves w and w + 4 k share a SIMD (the
// round-robin placement of a workgroup's waves), and the eight waves with a
// diagonal pair are spread two per SIMD, so each SIMD issues 34 sub-block
// products per K-step instead of 36.  The 96
// accumulator registers of three pairs leave no room for the outer fp32 sums
// at three waves per SIMD: the MFMA accumulators hold the whole chunk.  Their
// rounding grows with the chunk length (each MFMA adds its products into an
// accumulator that grows over the chunk, whatever the instruction's K), so the
// chunks are kept at 1024 amplitudes: normwise error 3.4e-7 at n = 20, L = 6
// against 1.4e-7 with the 256-amplitude outer sums of the per-tile-pair form.
constexpr int kThreads256 = 768;
constexpr int kWaves256 = kThreads256 / 64;

struct Role3 {
    int a0, a1, a2, b;
};

__device__ __forceinline__ Role3 wave_role3(int w) {
    switch (w) {
    case 0: return Role3{0, 1, 2, 0};
    case 1: return Role3{1, 2, 3, 1};
    case 2: return Role3{2, 3, 4, 2};
    case 3: return Role3{3, 0, 4, 3};
    case 4: return Role3{4, 0, 1, 4};
    case 5: return Role3{5, 6, 7, 4};
    case 6: return Role3{5, 0, 1, 5};
    case 7: return Role3{2, 3, 6, 5};
    case 8: return Role3{2, 3, 7, 6};
    case 9: return Role3{6, 0, 1, 6};
    case 10: return Role3{2, 3, 5, 7};
    default: return Role3{7, 0, 1, 7};
    }
}

// Direct-to-LDS loads of one stage (amplitudes k0 .. k0 + 31) of states 0..255:
// 64 wave-loads of 1 KB (four 256-byte rows each), wave-load g = 12 i + w.
// States past P read state P - 1 (their rows feed only G entries >= P).
__device__ __forceinline__ void issue_stage256(const uint8_t *__restrict__ D, size_t SB, int P, size_t k0,
                                               uint8_t *buf, int w, uint32_t lane) {
    typedef Geo<32> G;
    const size_t ko = (k0 >> 6) * 512 + (k0 & 63) * 4;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int g = i * kWaves256 + w;
        if (g >= 64) break;                                  // uniform: waves 4..11 issue five
        int row = 4 * g + (int)(lane / G::kSlots);
        __asm__ volatile("" : "+v"(row));
        const uint32_t lgc = (lane % G::kSlots) ^ G::swz((uint32_t)row);
        const uint32_t in_blk = (lgc / (G::kSlots / 2)) * 256 + (lgc % (G::kSlots / 2)) * 16;
        const int st = row < P ? row : P - 1;
        const uint8_t *src = D + (size_t)st * SB + ko + in_blk;
        const uint32_t dst =
            __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lds_void_t *)(buf + g * 1024));  // uniform
        asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                     :
                     : "s"(dst), "v"(src)
                     : "memory");
    }
}

// 16 x 16 x 32 form of the same stage (v_mfma_f32_16x16x32_f16: each
// instruction adds 32 products, half as many accumulator roundings per entry
// as 32 x 32 x 16 for the same K): a block pair is 2 x 2 sub-blocks of 16
// states.  Lane l holds A[row l % 16][k = 8 (l / 16) .. + 7] (one 16-byte
// slot: four amplitudes), B likewise by column, and accumulator register e
// is D[4 (l / 16) + e][l % 16] (checked on the device by
// qgtl_debug_mfma_layout16).  K-step t of a 32-amplitude stage covers
// amplitudes 16 t .. 16 t + 15: logical slot 8 p + 4 t + l / 16.
__device__ __forceinline__ f32x4 mma16(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(as_f16(a), as_f16(b), c, 0, 0, 0);
}

__device__ __forceinline__ u32x4 frag16(const uint8_t *buf, int blk, int sub, int p, int t, uint32_t lane) {
    const uint32_t row = (uint32_t)(32 * blk + 16 * sub) + (lane & 15);
    const uint32_t slot = ((uint32_t)(8 * p + 4 * t) + (lane >> 4)) ^ (lane & 15);   // swz(row) = row & 15
    return *reinterpret_cast<const u32x4 *>(buf + row * Geo<32>::kRowBytes + slot * 16);
}

// acc[i][2 sa + sb]: pair i, sub-block (sa, sb); Re and Im.  DIAG: pair 0 is a
// diagonal pair and its sub-block (1, 0), below the diagonal, is not computed.
template <int PT, bool DIAG>
__device__ __forceinline__ void stage_mfma3_16(const uint8_t *buf, const Role3 &ro, f32x4 (&re)[3][4],
                                               f32x4 (&im)[3][4], uint32_t lane) {
    const int ab[3] = {ro.a0, ro.a1, ro.a2};
#pragma unroll 1
    for (int t = 0; t < 2; t++) {
        u32x4 bh[2], bl[2], yh[2], yl[2];
#pragma unroll
        for (int sb = 0; sb < 2; sb++) {
            bh[sb] = frag16(buf, ro.b, sb, 0, t, lane);
            bl[sb] = frag16(buf, ro.b, sb, 1, t, lane);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                yh[sb][e] = swap_neg(bh[sb][e]);
                yl[sb][e] = swap_neg(bl[sb][e]);
            }
        }
#pragma unroll
        for (int sa = 0; sa < 2; sa++) {
            u32x4 ah[3], al[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                ah[i] = frag16(buf, ab[i], sa, 0, t, lane);
                al[i] = frag16(buf, ab[i], sa, 1, t, lane);
            }
            // smallest terms first; the accumulators interleaved
            if constexpr (PT == 4) {
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int sb = 0; sb < 2; sb++) {
                        if (DIAG && i == 0 && sa == 1 && sb == 0) continue;
                        re[i][2 * sa + sb] = mma16(al[i], bl[sb], re[i][2 * sa + sb]);
                        im[i][2 * sa + sb] = mma16(al[i], yl[sb], im[i][2 * sa + sb]);
                    }
            }
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int sb = 0; sb < 2; sb++) {
                    if (DIAG && i == 0 && sa == 1 && sb == 0) continue;
                    re[i][2 * sa + sb] = mma16(al[i], bh[sb], re[i][2 * sa + sb]);
                    im[i][2 * sa + sb] = mma16(al[i], yh[sb], im[i][2 * sa + sb]);
                }
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int sb = 0; sb < 2; sb++) {
                    if (DIAG && i == 0 && sa == 1 && sb == 0) continue;
                    re[i][2 * sa + sb] = mma16(ah[i], bl[sb], re[i][2 * sa + sb]);
                    im[i][2 * sa + sb] = mma16(ah[i], yl[sb], im[i][2 * sa + sb]);
                }
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int sb = 0; sb < 2; sb++) {
                    if (DIAG && i == 0 && sa == 1 && sb == 0) continue;
                    re[i][2 * sa + sb] = mma16(ah[i], bh[sb], re[i][2 * sa + sb]);
                    im[i][2 * sa + sb] = mma16(ah[i], yh[sb], im[i][2 * sa + sb]);
                }
        }
    }
}

// one accumulator entry of pair (a, b) into the tile-pair partials (transposed and
// conjugated when a > b)
__device__ __forceinline__ void store_entry256(float *__restrict__ out, int a, int b, int r, int c, float vr, float vi) {
    const bool tr = a > b;
    const int R = tr ? c : r, C = tr ? r : c;
    const int tile = (R >> 7) + (C >> 7);                    // (0,0) 0, (0,1) 1, (1,1) 2
    float *t = out + (size_t)tile * 2 * kTileElems;
    t[(R & 127) * kTileStates + (C & 127)] = vr;
    t[kTileElems + (R & 127) * kTileStates + (C & 127)] = tr ? -vi : vi;
}

__device__ __forceinline__ void store_pair256_16(float *__restrict__ out, int a, int b, const f32x4 (&re)[4],
                                                 const f32x4 (&im)[4], uint32_t lane) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int r = a * 32 + 16 * (q >> 1) + 4 * (int)(lane >> 4) + e;
            const int c = b * 32 + 16 * (q & 1) + (int)(lane & 15);
            store_entry256(out, a, b, r, c, re[q][e], im[q][e]);
        }
}

// One split-K chunk of a two-tile Gram: stages of 32 amplitudes through the two
// LDS buffers, the chunk's fp32 partials of the three tile pairs at the end.
template <int PT, bool DIAG>
__device__ __forceinline__ void gram_cplx256_chunk(const uint8_t *__restrict__ D, int P, size_t N, size_t chunk_len,
                                                   int chunk, float *__restrict__ partials, uint8_t *lds,
                                                   const Role3 &ro, int w, uint32_t lane) {
    typedef Geo<32> G;
    constexpr int kBuf = 2 * G::kTileBytes;                  // 256 rows of one stage: 64 KB
    static_assert(2 * kBuf <= kLdsBytes, "two stage buffers");
    const size_t SB = N * 8;
    const size_t kbeg = (size_t)chunk * chunk_len;
    const size_t kend = (kbeg + chunk_len < N) ? kbeg + chunk_len : N;
    const int nst = kend > kbeg ? (int)((kend - kbeg) / 32) : 0;

    f32x4 re4[3][4], im4[3][4];
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                re4[i][q][e] = 0.f;
                im4[i][q][e] = 0.f;
            }
    }
    if (nst > 0) issue_stage256(D, SB, P, kbeg, lds, w, lane);
    for (int st = 0; st < nst; st++) {
        // this wave's loads of stage st have landed; after the barrier every wave's
        // have, and every wave is done with stage st - 1's buffer, which stage st + 1 refills
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (st + 1 < nst) issue_stage256(D, SB, P, kbeg + (size_t)(st + 1) * 32, lds + ((st + 1) & 1) * kBuf, w, lane);
        stage_mfma3_16<PT, DIAG>(lds + (st & 1) * kBuf, ro, re4, im4, lane);
    }
    float *out = partials + (size_t)chunk * 3 * 2 * kTileElems;
    store_pair256_16(out, ro.a0, ro.b, re4[0], im4[0], lane);
    store_pair256_16(out, ro.a1, ro.b, re4[1], im4[1], lane);
    store_pair256_16(out, ro.a2, ro.b, re4[2], im4[2], lane);
}

// The wave's role decides the form once per chunk (a diagonal pair's waves
// skip one sub-block), so each form keeps its own accumulator allocation.
template <int PT>
__device__ __forceinline__ void gram_cplx256_unit(const uint8_t *__restrict__ D, int P, size_t N, size_t chunk_len,
                                                  int chunk, float *__restrict__ partials, uint8_t *lds) {
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & 63;
    const Role3 ro = wave_role3(w);
    if (ro.a0 == ro.b) gram_cplx256_chunk<PT, true>(D, P, N, chunk_len, chunk, partials, lds, ro, w, lane);
    else gram_cplx256_chunk<PT, false>(D, P, N, chunk_len, chunk, partials, lds, ro, w, lane);
}

template <int PT>
__global__ __launch_bounds__(kThreads256, 1) void gram_cplx256_kernel(const uint8_t *__restrict__ D, int P, size_t N,
                                                                     size_t chunk_len, float *__restrict__ partials) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[kLdsBytes];
    gram_cplx256_unit<PT>(D, P, N, chunk_len, (int)blockIdx.x, partials, lds);
}

// Grouped (many instances, 129..256 Gram states each): unit u = chunk u - unit_begin of
// its item; the workgroup loops over units with a grid stride.
template <int PT>
__global__ __launch_bounds__(kThreads256, 1) void gram_cplx256_grouped_kernel(const qgtl_gram_item_t *__restrict__ items,
                                                                             int nitems, uint32_t units) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[kLdsBytes];
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_gram_item_t it = items[qgtl::find_item(items, nitems, u)];
        // every wave is done with the previous unit's stage buffers
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        gram_cplx256_unit<PT>(qgtl::global_ptr<const uint8_t>(qgtl::uniform(it.D)), qgtl::uniform(it.P),
                              (size_t)qgtl::uniform(it.N), (size_t)qgtl::uniform(it.chunk_len),
                              (int)(u - qgtl::uniform(it.unit_begin)), qgtl::global_ptr<float>(qgtl::uniform(it.partials)),
                              lds);
    }
}

// v_mfma_f32_16x16x32_f16 lane-layout probe: A[i][k] = 1 for k = i + 16 (i & 1) (else 0),
// B[k][j] = k + 32 j, so D[i][j] = i + 16 (i & 1) + 32 j, written out[16 i + j] with the layout
// the 16 x 16 Gram form assumes.
__global__ void mfma16_layout_probe_kernel(float *out) {
    const uint32_t lane = threadIdx.x;
    f16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int i = (int)(lane & 15), k = 8 * (int)(lane >> 4) + e, j = (int)(lane & 15);
        a[e] = (_Float16)((k == i + 16 * (i & 1)) ? 1.f : 0.f);
        b[e] = (_Float16)(float)(k + 32 * j);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < 4; e++) out[(4 * (lane >> 4) + e) * 16 + (lane & 15)] = acc[e];
}

// G[ti 128 + r][tj 128 + c] = sum over chunks (FP64, fixed order) of the
// partials, for the blocks the Gram computed (on diagonal tile pairs the
// 32-state blocks on or above the diagonal).  A workgroup takes 64 adjacent
// entries; its eight waves sum the chunks c = w (mod 8), each in increasing
// order with eight loads in flight, and wave 0 adds the eight sums in wave
// order.
// Group grp (64 adjacent entries of the tile-pair partials) of one reduction.
__device__ __forceinline__ void gram_cplx_reduce_group(const float *__restrict__ partials, int nt, int nchunks,
                                                       double2 *__restrict__ G, size_t grp, double2 (*part)[64]) {
    constexpr int kGroups = kThreads / 64;                   // waves per entry, chunk stride
    const int T = nt * (nt + 1) / 2;
    const int R = nt * kTileStates;
    const size_t cs = (size_t)T * 2 * kTileElems;
    const int lane = (int)(threadIdx.x & 63), g = (int)(threadIdx.x >> 6);
    const size_t idx = grp * 64 + lane;
    const int tile = (int)(idx / kTileElems), e = (int)(idx % kTileElems);
    const int r = e / kTileStates, c = e % kTileStates;
    int ti, tj;
    tile_coords(tile, nt, ti, tj);
    const bool skip = ti == tj && (r >> 5) > (c >> 5);       // never written
    double sr = 0.0, si = 0.0;
    if (!skip) {
        const float *q = partials + (size_t)tile * 2 * kTileElems + e;
        int k = g;
        for (; k + 7 * kGroups < nchunks; k += 8 * kGroups) {
            float vr[8], vi[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                // streamed: every partial is read once
                vr[u] = __builtin_nontemporal_load(q + (size_t)(k + u * kGroups) * cs);
                vi[u] = __builtin_nontemporal_load(q + (size_t)(k + u * kGroups) * cs + kTileElems);
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                sr += (double)vr[u];
                si += (double)vi[u];
            }
        }
        for (; k < nchunks; k += kGroups) {
            sr += (double)q[(size_t)k * cs];
            si += (double)q[(size_t)k * cs + kTileElems];
        }
    }
    part[g][lane] = make_double2(sr, si);
    __syncthreads();
    if (g == 0 && !skip) {
        double2 t = part[0][lane];
#pragma unroll
        for (int w = 1; w < kGroups; w++) {
            t.x += part[w][lane].x;
            t.y += part[w][lane].y;
        }
        G[(size_t)(ti * kTileStates + r) * R + tj * kTileStates + c] = t;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void gram_cplx_reduce_kernel(const float *__restrict__ partials, int nt,
                                                                    int nchunks, double2 *__restrict__ G) {
    __shared__ double2 part[kThreads / 64][64];
    const size_t ngrp = (size_t)(nt * (nt + 1) / 2) * kTileElems / 64;
    for (size_t grp = blockIdx.x; grp < ngrp; grp += gridDim.x) gram_cplx_reduce_group(partials, nt, nchunks, G, grp, part);
}

// grouped: unit = one 64-entry group of an item (units per item: T * 256)
__global__ __launch_bounds__(kThreads) void gram_cplx_reduce_grouped_kernel(
    const qgtl_gram_reduce_item_t *__restrict__ items, int nitems, uint32_t units) {
    __shared__ double2 part[kThreads / 64][64];
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_gram_reduce_item_t it = items[qgtl::find_item(items, nitems, u)];
        gram_cplx_reduce_group(qgtl::global_ptr<const float>(qgtl::uniform(it.partials)), qgtl::uniform(it.nt),
                               qgtl::uniform(it.nchunks), qgtl::global_ptr<double2>(qgtl::uniform(it.C)),
                               (size_t)(u - qgtl::uniform(it.unit_begin)), part);
    }
}

// Q_kl = G_kl / (4 sqrt(G_kk G_ll)) - a_k conj(a_l), a_k = G_kP / (2 sqrt(G_kk G_PP)):
// state P is the primal state, and the normalisation applies the exact
// identities ||d_k psi|| = 1/2 and ||psi|| = 1 (the 2^30 scale of G cancels).
__device__ __forceinline__ void qgt_finalize_cplx_elem(const double2 *__restrict__ G, int R, int P,
                                                       double *__restrict__ Q, size_t idx) {
    const int k = (int)(idx / P), l = (int)(idx % P);
    if (k > l) return;
    const int m = P;
    const double2 g = G[(size_t)k * R + l];
    const double gkk = G[(size_t)k * R + k].x, gll = G[(size_t)l * R + l].x, gmm = G[(size_t)m * R + m].x;
    const double2 ak0 = G[(size_t)k * R + m], al0 = G[(size_t)l * R + m];
    const double sg = 0.25 / sqrt(gkk * gll);
    const double sk = 0.5 / sqrt(gkk * gmm), sl = 0.5 / sqrt(gll * gmm);
    const double akr = ak0.x * sk, aki = ak0.y * sk, alr = al0.x * sl, ali = al0.y * sl;
    const double pr = akr * alr + aki * ali, pi = aki * alr - akr * ali;
    const double qr = g.x * sg - pr, qi = (k == l) ? 0.0 : g.y * sg - pi;
    Q[2 * ((size_t)k * P + l)] = qr;
    Q[2 * ((size_t)k * P + l) + 1] = qi;
    if (k != l) {
        Q[2 * ((size_t)l * P + k)] = qr;
        Q[2 * ((size_t)l * P + k) + 1] = -qi;
    }
}

__global__ __launch_bounds__(kThreads) void qgt_finalize_cplx_kernel(const double2 *__restrict__ G, int R, int P,
                                                                     double *__restrict__ Q) {
    for (size_t idx = blockIdx.x * (size_t)kThreads + threadIdx.x; idx < (size_t)P * P;
         idx += (size_t)gridDim.x * kThreads)
        qgt_finalize_cplx_elem(G, R, P, Q, idx);
}

// grouped: unit = kThreads entries of an item's P x P (units per item: ceil(P^2 / kThreads))
__global__ __launch_bounds__(kThreads) void finalize_cplx_grouped_kernel(const qgtl_finalize_item_t *__restrict__ items,
                                                                         int nitems, uint32_t units) {
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_finalize_item_t it = items[qgtl::find_item(items, nitems, u)];
        const size_t idx = (size_t)(u - it.unit_begin) * kThreads + threadIdx.x;
        if (idx < (size_t)it.P * it.P)
            qgt_finalize_cplx_elem(reinterpret_cast<const double2 *>(it.C), it.R, it.P, reinterpret_cast<double *>(it.Q),
                                   idx);
    }
}


}  // namespace

extern "C" {

size_t qgtl_hip_gram_cplx_partial_floats(int P) {
    const int nt = (P + kTileStates - 1) / kTileStates;
    return (size_t)(nt * (nt + 1) / 2) * 2 * kTileElems;
}

int qgtl_hip_gram_cplx(const void *D, int P, size_t N, float *partials, int nchunks, int products, int stage_amps) {
    if (P < 1 || nchunks < kXcd || nchunks % kXcd != 0 || N % 64 != 0 || (products != 3 && products != 4) ||
        (stage_amps != 16 && stage_amps != 32))
        return qgtl::set_error_msg(-118, "qgtl_hip_gram_cplx: needs nchunks a multiple of 8, N % 64 == 0, 3 or 4 "
                                         "products, stages of 16 or 32 amplitudes");
    const int nt = (P + kTileStates - 1) / kTileStates;
    const int T = nt * (nt + 1) / 2;
    const size_t chunk_len = ((N + nchunks - 1) / nchunks + 31) / 32 * 32;
    const dim3 grid((unsigned)(T * nchunks)), block(kThreads);
    const uint8_t *d = (const uint8_t *)D;
    if (nt == 2) {
        // every chunk's three tile pairs in one workgroup
        // (16 x 16 x 32 MFMA form: 0.75 ms at n = 20, P = 240, against 0.83 ms for a 32 x 32 x 16 form of the
        // same kernel)
        const dim3 g((unsigned)nchunks), b(kThreads256);
        if (products == 4)
            hipLaunchKernelGGL((gram_cplx256_kernel<4>), g, b, 0, qgtl::stream(), d, P, N, chunk_len, partials);
        else
            hipLaunchKernelGGL((gram_cplx256_kernel<3>), g, b, 0, qgtl::stream(), d, P, N, chunk_len, partials);
        QGTL_LAUNCH_CHECK("gram_cplx256");
        return 0;
    }
    if (stage_amps == 16 && products == 4)
        hipLaunchKernelGGL((gram_cplx_kernel<16, 4>), grid, block, 0, qgtl::stream(), d, P, N, nt, nchunks, chunk_len, partials);
    else if (stage_amps == 16)
        hipLaunchKernelGGL((gram_cplx_kernel<16, 3>), grid, block, 0, qgtl::stream(), d, P, N, nt, nchunks, chunk_len, partials);
    else if (products == 4)
        hipLaunchKernelGGL((gram_cplx_kernel<32, 4>), grid, block, 0, qgtl::stream(), d, P, N, nt, nchunks, chunk_len, partials);
    else
        hipLaunchKernelGGL((gram_cplx_kernel<32, 3>), grid, block, 0, qgtl::stream(), d, P, N, nt, nchunks, chunk_len, partials);
    QGTL_LAUNCH_CHECK("gram_cplx");
    return 0;
}

int qgtl_hip_gram_cplx_reduce(const float *partials, int P, int nchunks, double *G) {
    if (P < 1 || nchunks < kXcd || nchunks % kXcd != 0)
        return qgtl::set_error_msg(-118, "qgtl_hip_gram_cplx_reduce: nchunks must be a positive multiple of 8");
    const int nt = (P + kTileStates - 1) / kTileStates;
    const size_t total = (size_t)(nt * (nt + 1) / 2) * kTileElems;
    const unsigned grid = (unsigned)(total / 64);            // one 64-entry group per workgroup
    hipLaunchKernelGGL(gram_cplx_reduce_kernel, dim3(grid), dim3(kThreads), 0, qgtl::stream(), partials, nt, nchunks,
                       reinterpret_cast<double2 *>(G));
    QGTL_LAUNCH_CHECK("gram_cplx_reduce");
    return 0;
}

static unsigned grouped_grid(uint32_t units, int grid) {
    if (grid <= 0 || (uint32_t)grid > units) return units;
    return (unsigned)grid;
}

int qgtl_hip_gram_cplx_grouped(const qgtl_gram_item_t *items, int nitems, uint32_t units, int grid, int products) {
    if (!items || nitems < 1 || units < 1 || (products != 3 && products != 4))
        return qgtl::set_error_msg(-118, "gram_cplx_grouped: empty item list or bad product count");
    const dim3 g(grouped_grid(units, grid)), b(kThreads256);
    if (products == 4)
        hipLaunchKernelGGL((gram_cplx256_grouped_kernel<4>), g, b, 0, qgtl::stream(), items, nitems, units);
    else
        hipLaunchKernelGGL((gram_cplx256_grouped_kernel<3>), g, b, 0, qgtl::stream(), items, nitems, units);
    QGTL_LAUNCH_CHECK("gram_cplx_grouped");
    return 0;
}

int qgtl_hip_gram_cplx_reduce_grouped(const qgtl_gram_reduce_item_t *items, int nitems, uint32_t units, int grid) {
    if (!items || nitems < 1 || units < 1) return qgtl::set_error_msg(-118, "gram_cplx_reduce_grouped: empty item list");
    hipLaunchKernelGGL(gram_cplx_reduce_grouped_kernel, dim3(grouped_grid(units, grid)), dim3(kThreads), 0,
                       qgtl::stream(), items, nitems, units);
    QGTL_LAUNCH_CHECK("gram_cplx_reduce_grouped");
    return 0;
}

int qgtl_hip_finalize_cplx_grouped(const qgtl_finalize_item_t *items, int nitems, uint32_t units, int grid) {
    if (!items || nitems < 1 || units < 1) return qgtl::set_error_msg(-118, "finalize_cplx_grouped: empty item list");
    hipLaunchKernelGGL(finalize_cplx_grouped_kernel, dim3(grouped_grid(units, grid)), dim3(kThreads), 0, qgtl::stream(),
                       items, nitems, units);
    QGTL_LAUNCH_CHECK("finalize_cplx_grouped");
    return 0;
}

int qgtl_hip_debug_mfma16_layout(float *out_dev) {
    hipLaunchKernelGGL(mfma16_layout_probe_kernel, dim3(1), dim3(64), 0, qgtl::stream(), out_dev);
    QGTL_LAUNCH_CHECK("mfma16_layout_probe");
    return 0;
}

int qgtl_hip_qgt_finalize_cplx(const double *G, int P, double *Q) {
    const int nt = (P + 1 + kTileStates - 1) / kTileStates;
    unsigned grid = (unsigned)(((size_t)P * P + kThreads - 1) / kThreads);
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(qgt_finalize_cplx_kernel, dim3(grid), dim3(kThreads), 0, qgtl::stream(),
                       reinterpret_cast<const double2 *>(G), nt * kTileStates, P, Q);
    QGTL_LAUNCH_CHECK("qgt_finalize_cplx");
    return 0;
}

}  // extern "C"
