// Complex64 state-vector sweeps for the mixed-precision QGT path.
//
// A "pass" loads one tile of 2^m amplitudes (the tile bits vary, every other
// bit is fixed by the tile index) into LDS, applies a set of commuting 2x2
// matrices on tile bits (three at a time from registers, so one LDS round
// trip serves three gates), and stores the tile, optionally through an index
// permutation: the CNOT ring of the hardware-efficient ansatz
// (CNOT(i -> i+1 mod n), i ascending) or, when the whole state is one tile,
// any CNOT/SWAP list.  Counterparts in QGTL: the per-gate loops
// apply_rotation_y/z and apply_controlled_x
// (src/quantum_geometric/core/quantum_circuit_operations.c:171-265) and
// apply_single_qubit_gate / apply_cnot_gate
// (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:174-210).
//
// Ring store, derived from the ring map y_j = x_0 ^ ... ^ x_j (j >= 1),
// y_0 = x_1 ^ ... ^ x_{n-1}: for a tile over bits 0..m-1 with tile index r
// (bits m..n-1), the low output bits depend on the tile only, and the output
// high bits are prefix_xor(r), complemented when y_{m-1} = 1.  The tile thus
// lands in two contiguous half tiles: every store is coalesced.
#include "qgtl_hip_common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kHiTableBits = 9;     // tile bits outside the contiguous low run

// LDS index of tile element j: bits 0..3 XORed with bits 4..7 (a bijection on
// any 2^m range), so strided gate-group and store patterns spread over the
// banks without padding and the tile stays 64 KB (two workgroups per CU even
// in the grouped kernel, which also holds a descriptor).
__device__ __forceinline__ int lds_pad(int j) { return j ^ ((j >> 4) & 15); }

__device__ __forceinline__ float2 cmadd(float2 acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
    return acc;
}

__device__ __forceinline__ void apply2(float2 &a, float2 &b, const float2 *m) {
    float2 na = cmadd(cmadd(make_float2(0.f, 0.f), m[0], a), m[1], b);
    float2 nb = cmadd(cmadd(make_float2(0.f, 0.f), m[2], a), m[3], b);
    a = na;
    b = nb;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// Real rotation [[c, -s], [s, c]] on a complex pair: 4 real FMA per amplitude.
__device__ __forceinline__ void rot2(float2 &a, float2 &b, float c, float s) {
    const float2 na = make_float2(fmaf(c, a.x, -s * b.x), fmaf(c, a.y, -s * b.y));
    const float2 nb = make_float2(fmaf(s, a.x, c * b.x), fmaf(s, a.y, c * b.y));
    a = na;
    b = nb;
}

// Phase tables of a rotation-form pass: entry i < 128 is the product of the
// post (pre) phases of the gates on tile bits 0..6 at bit pattern i, entry
// 128 + i the same for tile bits 7..12, so amplitude j gets
// tab[j & 127] * tab[128 + (j >> 7)].
constexpr int kPhaseTab = QGTL_PHASE_TAB;

__device__ __forceinline__ void build_phase_tables(const qgtl_pass_desc_t &d, float2 *post_tab, float2 *pre_tab) {
    for (int t = threadIdx.x; t < kPhaseTab; t += kThreads) {
        const bool lo = t < 128;
        const int i = lo ? t : t - 128;
        float2 a = make_float2(1.f, 0.f), b = make_float2(1.f, 0.f);
        for (int g = 0; g < d.ngates; g++) {
            const int bit = d.gate_bit[g];
            if (lo ? bit >= 7 : bit < 7) continue;
            const bool v = (i >> (lo ? bit : bit - 7)) & 1;
            // uniform loads of both entries, then a per-lane select
            const float2 p0 = make_float2(d.post[g][0].re, d.post[g][0].im);
            const float2 p1 = make_float2(d.post[g][1].re, d.post[g][1].im);
            const float2 q0 = make_float2(d.pre[g][0].re, d.pre[g][0].im);
            const float2 q1 = make_float2(d.pre[g][1].re, d.pre[g][1].im);
            a = cmul(a, v ? p1 : p0);
            b = cmul(b, v ? q1 : q0);
        }
        post_tab[t] = a;
        pre_tab[t] = b;
    }
}

// The same tables built by the host (post then pre, kPhaseTab entries each),
// copied in instead of rebuilt by every workgroup of the launch.
__device__ __forceinline__ void load_phase_tables(const float2 *__restrict__ g, float2 *post_tab, float2 *pre_tab) {
    for (int t = threadIdx.x; t < 2 * kPhaseTab; t += kThreads) {
        if (t < kPhaseTab) post_tab[t] = g[t];
        else pre_tab[t - kPhaseTab] = g[t];
    }
}

__device__ __forceinline__ float2 phase_at(const float2 *tab, uint32_t j) {
    return cmul(tab[j & 127u], tab[128u + (j >> 7)]);
}

__device__ __forceinline__ uint32_t insert_zero(uint32_t x, int b) {
    uint32_t lo = x & ((1u << b) - 1u);
    return ((x >> b) << (b + 1)) | lo;
}

// Three (or fewer) gates on tile bits bits[0] < bits[1] < bits[2]: each
// thread owns 2^G amplitudes that differ only in those bits.
template <int G>
__device__ __forceinline__ void apply_group(float2 *tile, int m, const int *bits, const float2 (*mats)[4]) {
    const int count = 1 << (m - G);
    for (int o = threadIdx.x; o < count; o += kThreads) {
        uint32_t base = (uint32_t)o;
#pragma unroll
        for (int i = 0; i < G; i++) base = insert_zero(base, bits[i]);
        float2 v[1 << G];
        int idx[1 << G];
#pragma unroll
        for (int c = 0; c < (1 << G); c++) {
            uint32_t id = base;
#pragma unroll
            for (int i = 0; i < G; i++)
                if ((c >> i) & 1) id |= 1u << bits[i];
            idx[c] = lds_pad((int)id);
            v[c] = tile[idx[c]];
        }
#pragma unroll
        for (int i = 0; i < G; i++) {
#pragma unroll
            for (int c = 0; c < (1 << G); c++) {
                if ((c >> i) & 1) continue;
                apply2(v[c], v[c | (1 << i)], mats[i]);
            }
        }
#pragma unroll
        for (int c = 0; c < (1 << G); c++) tile[idx[c]] = v[c];
    }
}

template <int G>
__device__ __forceinline__ void apply_group_rot(float2 *tile, int m, const int *bits, const float (*cs)[2]) {
    const int count = 1 << (m - G);
    for (int o = threadIdx.x; o < count; o += kThreads) {
        uint32_t base = (uint32_t)o;
#pragma unroll
        for (int i = 0; i < G; i++) base = insert_zero(base, bits[i]);
        float2 v[1 << G];
        int idx[1 << G];
#pragma unroll
        for (int c = 0; c < (1 << G); c++) {
            uint32_t id = base;
#pragma unroll
            for (int i = 0; i < G; i++)
                if ((c >> i) & 1) id |= 1u << bits[i];
            idx[c] = lds_pad((int)id);
            v[c] = tile[idx[c]];
        }
#pragma unroll
        for (int i = 0; i < G; i++) {
#pragma unroll
            for (int c = 0; c < (1 << G); c++) {
                if ((c >> i) & 1) continue;
                rot2(v[c], v[c | (1 << i)], cs[i][0], cs[i][1]);
            }
        }
#pragma unroll
        for (int c = 0; c < (1 << G); c++) tile[idx[c]] = v[c];
    }
}

// Inverse of the local part of the ring map for the low tile (see header).
__device__ __forceinline__ uint32_t ring_local_inverse(uint32_t y, int m, uint32_t par_r) {
    uint32_t x0 = (y & 1u) ^ ((y >> (m - 1)) & 1u) ^ par_r;
    uint32_t x1 = ((y >> 1) & 1u) ^ x0;
    uint32_t rest = (y ^ (y << 1)) & ~3u & ((1u << m) - 1u);
    return rest | (x1 << 1) | x0;
}

constexpr int kLdsTileElems = 1 << QGTL_TILE_BITS_MAX;

// One tile r of one state: load, gate groups through LDS, store.
__device__ __forceinline__ void pass_lds_body(const qgtl_pass_desc_t &d, const float2 *__restrict__ src,
                                              float2 *dst, uint32_t r, float2 *tile, uint32_t *hi_off,
                                              float2 *post_tab, float2 *pre_tab, const float2 *gtabs = nullptr) {
    const int m = d.m, nlow = d.nlow, nhi = m - nlow;
    const bool rotf = d.form == 1, pre = rotf && d.has_pre;
    if (rotf) {
        if (gtabs) load_phase_tables(gtabs, post_tab, pre_tab);
        else build_phase_tables(d, post_tab, pre_tab);
    }

    uint32_t base = 0;
    for (int i = 0; i < d.nrest; i++) base |= ((r >> i) & 1u) << d.rest_pos[i];
    for (int h = threadIdx.x; h < (1 << nhi); h += kThreads) {
        uint32_t off = 0;
        for (int i = 0; i < nhi; i++) off |= ((uint32_t)(h >> i) & 1u) << d.hi_pos[i];
        hi_off[h] = off;
    }
    __syncthreads();

    const int T = 1 << m;
    const uint32_t lowmask = (1u << nlow) - 1u;
    for (int p = threadIdx.x; p < (T >> 1); p += kThreads) {
        const int j = 2 * p;
        const uint32_t g = base + hi_off[j >> nlow] + ((uint32_t)j & lowmask);
        const float4 v = *reinterpret_cast<const float4 *>(src + g);
        float2 a = make_float2(v.x, v.y), b = make_float2(v.z, v.w);
        if (pre) {
            a = cmul(a, phase_at(pre_tab, (uint32_t)j));
            b = cmul(b, phase_at(pre_tab, (uint32_t)j + 1u));
        }
        tile[lds_pad(j)] = a;
        tile[lds_pad(j + 1)] = b;
    }
    __syncthreads();

    for (int g0 = 0; g0 < d.ngates; g0 += 3) {
        const int G = min(3, d.ngates - g0);
        // apply_group inserts the gate bits in ascending order; a pass planned
        // for the register kernel lists its gates in phase order, so sort the
        // group's gate indices by tile bit first
        int gidx[3];
#pragma unroll
        for (int i = 0; i < 3; i++) gidx[i] = min(g0 + i, d.ngates - 1);
        if (G >= 2 && d.gate_bit[gidx[1]] < d.gate_bit[gidx[0]]) { const int t = gidx[0]; gidx[0] = gidx[1]; gidx[1] = t; }
        if (G == 3) {
            if (d.gate_bit[gidx[2]] < d.gate_bit[gidx[1]]) { const int t = gidx[1]; gidx[1] = gidx[2]; gidx[2] = t; }
            if (d.gate_bit[gidx[1]] < d.gate_bit[gidx[0]]) { const int t = gidx[0]; gidx[0] = gidx[1]; gidx[1] = t; }
        }
        int bits[3];
#pragma unroll
        for (int i = 0; i < 3; i++) bits[i] = d.gate_bit[gidx[i]];
        if (rotf) {
            float cs[3][2];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                cs[i][0] = d.rot[gidx[i]].re;
                cs[i][1] = d.rot[gidx[i]].im;
            }
            if (G == 3) apply_group_rot<3>(tile, m, bits, cs);
            else if (G == 2) apply_group_rot<2>(tile, m, bits, cs);
            else apply_group_rot<1>(tile, m, bits, cs);
        } else {
            float2 mats[3][4];
#pragma unroll
            for (int i = 0; i < 3; i++) {
#pragma unroll
                for (int e = 0; e < 4; e++) mats[i][e] = make_float2(d.mat[gidx[i]][e].re, d.mat[gidx[i]][e].im);
            }
            if (G == 3) apply_group<3>(tile, m, bits, mats);
            else if (G == 2) apply_group<2>(tile, m, bits, mats);
            else apply_group<1>(tile, m, bits, mats);
        }
        __syncthreads();
    }

    if (d.store == QGTL_STORE_INPLACE) {
        for (int p = threadIdx.x; p < (T >> 1); p += kThreads) {
            const int j = 2 * p;
            const uint32_t g = base + hi_off[j >> nlow] + ((uint32_t)j & lowmask);
            float2 a = tile[lds_pad(j)], b = tile[lds_pad(j + 1)];
            if (rotf) {
                a = cmul(a, phase_at(post_tab, (uint32_t)j));
                b = cmul(b, phase_at(post_tab, (uint32_t)j + 1u));
            }
            *reinterpret_cast<float4 *>(dst + g) = make_float4(a.x, a.y, b.x, b.y);
        }
    } else if (d.store == QGTL_STORE_RING) {
        // tile = bits 0..m-1, tile index r = bits m..n-1
        const int nh = d.n - m;
        uint32_t pre = r;
        pre ^= pre << 1; pre ^= pre << 2; pre ^= pre << 4; pre ^= pre << 8; pre ^= pre << 16;
        const uint32_t ones = (nh >= 32) ? 0xffffffffu : ((1u << nh) - 1u);
        pre &= ones;
        const uint32_t par_r = (uint32_t)__popc(r) & 1u;
        for (int p = threadIdx.x; p < (T >> 1); p += kThreads) {
            const uint32_t y = 2u * (uint32_t)p;
            const uint32_t ja = ring_local_inverse(y, m, par_r), jb = ring_local_inverse(y + 1u, m, par_r);
            float2 a = tile[lds_pad((int)ja)];
            float2 b = tile[lds_pad((int)jb)];
            if (rotf) {
                a = cmul(a, phase_at(post_tab, ja));
                b = cmul(b, phase_at(post_tab, jb));
            }
            const uint32_t yhi = ((y >> (m - 1)) & 1u) ? (pre ^ ones) : pre;
            const size_t g = ((size_t)yhi << m) | y;
            *reinterpret_cast<float4 *>(dst + g) = make_float4(a.x, a.y, b.x, b.y);
        }
    } else {  // QGTL_STORE_PERM: whole state in this tile, undo the gate list per output index
        for (int p = threadIdx.x; p < (T >> 1); p += kThreads) {
            float2 o[2];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                uint32_t x = 2u * (uint32_t)p + (uint32_t)e;
                for (int i = d.nperm - 1; i >= 0; i--) {
                    const uint32_t ab = (x >> d.perm_a[i]) & 1u, bb = (x >> d.perm_b[i]) & 1u;
                    if (d.perm_type[i] == 0) x ^= ab << d.perm_b[i];              // CNOT a -> b
                    else x ^= (ab ^ bb) * ((1u << d.perm_a[i]) | (1u << d.perm_b[i]));  // SWAP
                }
                o[e] = tile[lds_pad((int)x)];
                if (rotf) o[e] = cmul(o[e], phase_at(post_tab, x));
            }
            *reinterpret_cast<float4 *>(dst + 2 * p) = make_float4(o[0].x, o[0].y, o[1].x, o[1].y);
        }
    }
}

__global__ __launch_bounds__(kThreads) void pass_c64_kernel(const qgtl_pass_desc_t d,
                                                            const float2 *__restrict__ src_all,
                                                            float2 *dst_all, size_t stride,
                                                            const float2 *__restrict__ gtabs) {
    __shared__ float2 tile[kLdsTileElems];
    __shared__ uint32_t hi_off[1 << kHiTableBits];
    __shared__ float2 tabs[2 * kPhaseTab];
    pass_lds_body(d, src_all + (size_t)blockIdx.y * stride, dst_all + (size_t)blockIdx.y * stride, blockIdx.x,
                  tile, hi_off, tabs, tabs + kPhaseTab, gtabs);
}

// ---------------------------------------------------------------------------
// Register-phase pass for full 2^13 tiles.  Each of the 256 threads owns 32
// amplitudes: in a layout, 5 tile bits select the register and the other 8
// tile bits (ascending) are the thread index.  Gates on register bits run
// with compile-time register pairing; a layout change is one LDS transpose
// (write all, barrier, read all, barrier).  The load/store layout {0, 7..10}
// gives each 64-lane load instruction 1 KB of contiguous amplitudes when the
// tile's low 7 bits are contiguous.
constexpr int kRT = 13;
constexpr int kRegs = 32;

// XOR swizzle of the register kernel's LDS tile: bits 0..3 of j are XORed with
// bits 4..7, 8..11 and 12, so a quarter-wave whose four lane bits have
// independent bank vectors (the host picks such thread-bit orders) is
// conflict-free for 8-byte accesses.
__device__ __forceinline__ int rpad(uint32_t j) { return (int)(j ^ (((j >> 4) ^ (j >> 8) ^ (j >> 12)) & 15u)); }

// Tile index of register r for a thread whose other-bit deposit is tdep: tdep
// | (the register bits of r).  rpad is linear over GF(2) and the thread and
// register parts are disjoint, so the LDS word of register r is lbase ^ (xor
// of lbit over the bits of r): one XOR with a uniform value per access.
// lbase and lbit hold byte offsets (8 bytes per word), so the XOR is the
// whole address computation.
struct RegLayout {
    uint32_t tdep, lbase;
    uint32_t rb[5], lbit[5];
};

__device__ __forceinline__ RegLayout make_layout(const int32_t *rbits, const int32_t *tbits, uint32_t t) {
    RegLayout L;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        L.rb[i] = qgtl::uniform(1u << rbits[i]);
        L.lbit[i] = (uint32_t)rpad(L.rb[i]) << 3;
    }
    uint32_t d = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) d |= ((t >> i) & 1u) << qgtl::uniform(tbits[i]);
    L.tdep = d;
    L.lbase = (uint32_t)rpad(d) << 3;
    return L;
}

// the complex64 word at byte offset b of the LDS tile
__device__ __forceinline__ float2 &lds_word(float2 *tile, uint32_t b) {
    return *reinterpret_cast<float2 *>(reinterpret_cast<char *>(tile) + b);
}
__device__ __forceinline__ const float2 &lds_word(const float2 *tile, uint32_t b) {
    return *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(tile) + b);
}

template <int R>
__device__ __forceinline__ uint32_t reg_index(const RegLayout &L) {
    uint32_t j = L.tdep;
#pragma unroll
    for (int i = 0; i < 5; i++)
        if ((R >> i) & 1) j |= L.rb[i];
    return j;
}

template <int R>
__device__ __forceinline__ uint32_t lds_index(const RegLayout &L) {
    uint32_t u = 0;
#pragma unroll
    for (int i = 0; i < 5; i++)
        if ((R >> i) & 1) u ^= L.lbit[i];
    return L.lbase ^ u;
}

// Linear part of the in-tile ring map (ring_local_inverse's inverse without
// the parity-of-r constant): prefix xor, bit 12 folded into bit 0.
__device__ __forceinline__ uint32_t ring_lin(uint32_t x) {
    uint32_t p = x;
    p ^= p << 1; p ^= p << 2; p ^= p << 4; p ^= p << 8;
    p &= (1u << kRT) - 1u;
    return p ^ ((p >> (kRT - 1)) & 1u);
}

// LDS words of the ring positions of a layout's elements: element x sits at
// rpad(ring_lin(x) ^ par_r) = base ^ (xor of bit over its register bits), as
// byte offsets.
struct RingLds {
    uint32_t base;
    uint32_t bit[5];
};

__device__ __forceinline__ RingLds ring_lds(const RegLayout &L, uint32_t par_r, uint32_t xm = 0) {
    RingLds q;
    q.base = (uint32_t)rpad(ring_lin(L.tdep ^ xm) ^ par_r) << 3;
#pragma unroll
    for (int i = 0; i < 5; i++) q.bit[i] = (uint32_t)rpad(ring_lin(L.rb[i])) << 3;
    return q;
}

template <int R>
__device__ __forceinline__ uint32_t ring_index(const RingLds &q) {
    uint32_t u = 0;
#pragma unroll
    for (int i = 0; i < 5; i++)
        if ((R >> i) & 1) u ^= q.bit[i];
    return q.base ^ u;
}

// Global element offsets of a layout's register pairs (register bit 0 = tile
// bit 0): lane + the sum of bit[i] over register bits i >= 1 of the pair; the
// parts are disjoint bit sets of the state index.
struct GlobalMap {
    size_t lane;
    uint32_t bit[5];
};

template <int R>
__device__ __forceinline__ size_t global_index(const GlobalMap &g) {
    uint32_t u = 0;
#pragma unroll
    for (int i = 1; i < 5; i++)
        if ((R >> i) & 1) u += g.bit[i];
    return g.lane + u;
}

// State offset of tile bits nlow.. given as h (bit i of h is tile bit nlow + i,
// at state bit hi_pos[i]).  Fixed trip count with unconditional reads of the
// (zero-filled) table and masked bits, so the scalar loads issue together
// instead of one load and one wait per bit ahead of the tile's loads.
__device__ __forceinline__ uint32_t deposit_hi(const qgtl_pass_desc_t &d, uint32_t h) {
    const int nhi = kRT - d.nlow;
    uint32_t off = 0;
#pragma unroll
    for (int i = 0; i < kHiTableBits; i++) {
        const uint32_t bit = i < nhi ? (h >> i) & 1u : 0u;
        off |= bit << ((uint32_t)d.hi_pos[i] & 31u);
    }
    return off;
}

// natural placement: tile bits below nlow are state bits 0..nlow-1, the others
// go to hi_pos, the tile index r to base
__device__ __forceinline__ GlobalMap natural_map(const RegLayout &L, uint32_t base, const qgtl_pass_desc_t &d) {
    GlobalMap g;
    const int nlow = d.nlow;
    const uint32_t lowmask = (1u << nlow) - 1u;
    g.lane = (size_t)(base + deposit_hi(d, L.tdep >> nlow) + (L.tdep & lowmask));
#pragma unroll
    for (int i = 0; i < 5; i++)
        g.bit[i] = qgtl::uniform((L.rb[i] & lowmask) ? L.rb[i] : deposit_hi(d, L.rb[i] >> nlow));
    return g;
}

// ring placement of tile r (bits 0..12 of the output index y): the high part
// of y is pre, or pre ^ ones when y_12 = 1 (a thread bit of the layout)
__device__ __forceinline__ GlobalMap ring_map(const RegLayout &L, uint32_t pre, uint32_t ones) {
    GlobalMap g;
    const uint32_t yhi = ((L.tdep >> (kRT - 1)) & 1u) ? (pre ^ ones) : pre;
    g.lane = ((size_t)yhi << kRT) | L.tdep;
#pragma unroll
    for (int i = 0; i < 5; i++) g.bit[i] = L.rb[i];
    return g;
}

template <int P>
__device__ __forceinline__ void apply_reg(float2 (&v)[kRegs], const float2 *m) {
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        if (!((r >> P) & 1)) apply2(v[r], v[r | (1 << P)], m);
}

// Rotation with its factor taken out, K(t): (a, b) -> (a - t b, b + t a),
// 2 FMA per amplitude (the host folds the rest of the rotation into the
// phase layers and the store index).
template <int P>
__device__ __forceinline__ void apply_reg_scaled(float2 (&v)[kRegs], float t) {
#pragma unroll
    for (int r = 0; r < kRegs; r++) {
        if ((r >> P) & 1) continue;
        const float2 a = v[r], b = v[r | (1 << P)];
        v[r] = make_float2(fmaf(-t, b.x, a.x), fmaf(-t, b.y, a.y));
        v[r | (1 << P)] = make_float2(fmaf(t, a.x, b.x), fmaf(t, a.y, b.y));
    }
}

// The descriptor fields one phase of a register pass reads (layout, gate
// mask, rotation-form coefficients), loaded one phase ahead so that their
// scalar loads are in flight during the previous phase's work.
struct PhaseParams {
    int32_t rbits[5], tbits[8];
    int32_t mask;
    float tan[5];
};

__device__ __forceinline__ PhaseParams load_phase(const qgtl_pass_desc_t &d, int ph) {
    PhaseParams q;
#pragma unroll
    for (int i = 0; i < 5; i++) q.rbits[i] = d.ph_rbits[ph][i];
#pragma unroll
    for (int i = 0; i < 8; i++) q.tbits[i] = d.ph_tbits[ph][i];
    q.mask = d.ph_mask[ph];
#pragma unroll
    for (int i = 0; i < 5; i++) q.tan[i] = d.ph_tan[ph][i];
    return q;
}

// The gates of a phase (compile-time pairing), one per register bit: t = 0
// where the phase has no gate (an exact identity), so the phase is straight-line
// code and no register copies arise where branches would join.
__device__ __forceinline__ void apply_phase_rots(float2 (&v)[kRegs], const PhaseParams &q) {
    apply_reg_scaled<0>(v, q.tan[0]);
    apply_reg_scaled<1>(v, q.tan[1]);
    apply_reg_scaled<2>(v, q.tan[2]);
    apply_reg_scaled<3>(v, q.tan[3]);
    apply_reg_scaled<4>(v, q.tan[4]);
}

// The same with the phase's gate mask known at compile time: the register
// bits without a gate get no instructions.
template <uint32_t MASK>
__device__ __forceinline__ void apply_phase_rots_m(float2 (&v)[kRegs], const PhaseParams &q) {
    if constexpr ((MASK & 1u) != 0) apply_reg_scaled<0>(v, q.tan[0]);
    if constexpr ((MASK & 2u) != 0) apply_reg_scaled<1>(v, q.tan[1]);
    if constexpr ((MASK & 4u) != 0) apply_reg_scaled<2>(v, q.tan[2]);
    if constexpr ((MASK & 8u) != 0) apply_reg_scaled<3>(v, q.tan[3]);
    if constexpr ((MASK & 16u) != 0) apply_reg_scaled<4>(v, q.tan[4]);
}

// Phase shapes of rotation-form register passes with their gate masks fixed
// at compile time: SHAPE = 0 is the general form (run-time phase count, every
// register bit rotated, t = 0 where there is no gate); otherwise bits 0..1
// hold the phase count and bits 2 + 5 p .. 6 + 5 p the mask of phase p.
constexpr uint32_t shape_code(int nph, uint32_t m0, uint32_t m1, uint32_t m2) {
    return (uint32_t)nph | (m0 << 2) | (m1 << 7) | (m2 << 12);
}
constexpr uint32_t kShape3Full = shape_code(3, 0x1f, 0x1f, 0x07);   // 13 gates over three phases
constexpr uint32_t kShape2Hi7 = shape_code(2, 0x1e, 0x07, 0);       // 7 gates, register bit 0 without one
constexpr int shape_phases(uint32_t shape) { return (int)(shape & 3u); }
constexpr uint32_t shape_mask(uint32_t shape, int ph) { return (shape >> (2 + 5 * ph)) & 31u; }

template <int J>
__device__ __forceinline__ void apply_phase_mat(float2 (&v)[kRegs], const qgtl_pass_desc_t &d, int ph, int mask) {
    if ((mask >> J) & 1) {
        float2 m[4];
#pragma unroll
        for (int e = 0; e < 4; e++) m[e] = make_float2(d.ph_mat[ph][J][e].re, d.ph_mat[ph][J][e].im);
        apply_reg<J>(v, m);
    }
}

// Amplitudes g (even) and g + 1 of a state in split-plane form: per block of 64
// amplitudes 128 words, the hi words (Re hi, Im hi as an fp16 pair) of the 64
// amplitudes, then their lo words (Re lo, Im lo) -- the pieces the Gram would
// cut from the complex64 values, interleaved per amplitude as the Gram's
// complex products read them.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

// 16-byte global load / store of two amplitudes; NT: non-temporal (streamed
// through the caches: each state is read and written once per pass)
template <bool NT>
__device__ __forceinline__ float4 gload4(const float2 *p) {
    if constexpr (NT) {
        const f32x4_t q = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t *>(p));
        return make_float4(q.x, q.y, q.z, q.w);
    } else {
        return *reinterpret_cast<const float4 *>(p);
    }
}

template <bool NT>
__device__ __forceinline__ void gstore4(float2 *p, float2 a, float2 b) {
    if constexpr (NT) {
        const f32x4_t q = {a.x, a.y, b.x, b.y};
        __builtin_nontemporal_store(q, reinterpret_cast<f32x4_t *>(p));
    } else {
        *reinterpret_cast<float4 *>(p) = make_float4(a.x, a.y, b.x, b.y);
    }
}

template <bool NT = false>
__device__ __forceinline__ void write_split_pair(uint32_t *planes, size_t g, float2 a, float2 b) {
    uint32_t h0, l0, h1, l1;
    qgtl::split_f16_pair(a.x * qgtl::kSplitScale, a.y * qgtl::kSplitScale, h0, l0);
    qgtl::split_f16_pair(b.x * qgtl::kSplitScale, b.y * qgtl::kSplitScale, h1, l1);
    uint32_t *blk = planes + (g >> 6) * 128 + (g & 63);
    if constexpr (NT) {
        const u32x2_t hi = {h0, h1}, lo = {l0, l1};
        __builtin_nontemporal_store(hi, reinterpret_cast<u32x2_t *>(blk));
        __builtin_nontemporal_store(lo, reinterpret_cast<u32x2_t *>(blk + 64));
    } else {
        *reinterpret_cast<uint2 *>(blk) = make_uint2(h0, h1);
        *reinterpret_cast<uint2 *>(blk + 64) = make_uint2(l0, l1);
    }
}


template <int R>
struct RegLoop {
    __device__ __forceinline__ static void store_lds(float2 *tile, const RegLayout &L, const float2 (&v)[kRegs]) {
        RegLoop<R - 1>::store_lds(tile, L, v);
        lds_word(tile, lds_index<R - 1>(L)) = v[R - 1];
    }
    __device__ __forceinline__ static void load_lds(const float2 *tile, const RegLayout &L, float2 (&v)[kRegs]) {
        RegLoop<R - 1>::load_lds(tile, L, v);
        v[R - 1] = lds_word(tile, lds_index<R - 1>(L));
    }
    // element x of the layout to / from its ring position
    __device__ __forceinline__ static void store_lds_ring(float2 *tile, const RingLds &q, const float2 (&v)[kRegs]) {
        RegLoop<R - 1>::store_lds_ring(tile, q, v);
        lds_word(tile, ring_index<R - 1>(q)) = v[R - 1];
    }
    __device__ __forceinline__ static void load_lds_ring(const float2 *tile, const RingLds &q, float2 (&v)[kRegs]) {
        RegLoop<R - 1>::load_lds_ring(tile, q, v);
        v[R - 1] = lds_word(tile, ring_index<R - 1>(q));
    }
    // phase layer as (thread part pt) x (register part from the descriptor)
    template <bool POST>
    __device__ __forceinline__ static void phase_split(float2 pt, const qgtl_pass_desc_t &d, float2 (&v)[kRegs]) {
        RegLoop<R - 1>::template phase_split<POST>(pt, d, v);
        const qgtl_c64_t q = POST ? d.reg_post[R - 1] : d.reg_pre[R - 1];
        v[R - 1] = cmul(v[R - 1], cmul(pt, make_float2(q.re, q.im)));
    }
    // global access in pairs (register bit 0 = tile bit 0)
    template <bool NT = false>
    __device__ __forceinline__ static void load_global(const float2 *src, const GlobalMap &g, float2 (&v)[kRegs]) {
        RegLoop<R - 2>::template load_global<NT>(src, g, v);
        const float4 q = gload4<NT>(src + global_index<R - 2>(g));
        v[R - 2] = make_float2(q.x, q.y);
        v[R - 1] = make_float2(q.z, q.w);
    }
    template <bool NT = false>
    __device__ __forceinline__ static void store_global(float2 *dst, const GlobalMap &g, const float2 (&v)[kRegs]) {
        RegLoop<R - 2>::template store_global<NT>(dst, g, v);
        gstore4<NT>(dst + global_index<R - 2>(g), v[R - 2], v[R - 1]);
    }
    // derivative state formed while loading: element y of the state is
    // M[s][s] psi[y] + M[s][1-s] psi[y ^ v], s = parity(y & w), the same
    // arithmetic as the injection kernel (psi is shared by every state of the
    // launch and read through the caches)
    __device__ __forceinline__ static void load_global_inject(const float2 *psi, const GlobalMap &g,
                                                              const qgtl_inject_slot_t &q, float2 (&v)[kRegs]) {
        RegLoop<R - 2>::load_global_inject(psi, g, q, v);
        const size_t a = global_index<R - 2>(g);
        const float4 qa = *reinterpret_cast<const float4 *>(psi + a);
        const float4 qb = *reinterpret_cast<const float4 *>(psi + ((a ^ (size_t)q.v) & ~(size_t)1));
        const bool odd = (q.v & 1u) != 0;
        const float2 b0 = odd ? make_float2(qb.z, qb.w) : make_float2(qb.x, qb.y);
        const float2 b1 = odd ? make_float2(qb.x, qb.y) : make_float2(qb.z, qb.w);
        const bool s0 = (__popc((uint32_t)a & q.w) & 1) != 0, s1 = s0 != ((q.w & 1u) != 0);
        const float2 m00 = make_float2(q.M[0].re, q.M[0].im), m01 = make_float2(q.M[1].re, q.M[1].im);
        const float2 m10 = make_float2(q.M[2].re, q.M[2].im), m11 = make_float2(q.M[3].re, q.M[3].im);
        v[R - 2] = cmadd(cmadd(make_float2(0.f, 0.f), s0 ? m11 : m00, make_float2(qa.x, qa.y)), s0 ? m10 : m01, b0);
        v[R - 1] = cmadd(cmadd(make_float2(0.f, 0.f), s1 ? m11 : m00, make_float2(qa.z, qa.w)), s1 ? m10 : m01, b1);
    }
    // the same for a diagonal matrix (q.v == 0): element y is M[s][s] psi[y], the
    // partner term being an exact zero -- no partner load, one multiply-add
    __device__ __forceinline__ static void load_global_inject_diag(const float2 *psi, const GlobalMap &g,
                                                                   const qgtl_inject_slot_t &q, float2 (&v)[kRegs]) {
        RegLoop<R - 2>::load_global_inject_diag(psi, g, q, v);
        const size_t a = global_index<R - 2>(g);
        const float4 qa = *reinterpret_cast<const float4 *>(psi + a);
        const bool s0 = (__popc((uint32_t)a & q.w) & 1) != 0, s1 = s0 != ((q.w & 1u) != 0);
        const float2 m00 = make_float2(q.M[0].re, q.M[0].im), m11 = make_float2(q.M[3].re, q.M[3].im);
        v[R - 2] = cmadd(make_float2(0.f, 0.f), s0 ? m11 : m00, make_float2(qa.x, qa.y));
        v[R - 1] = cmadd(make_float2(0.f, 0.f), s1 ? m11 : m00, make_float2(qa.z, qa.w));
    }
    // the same for an anti-diagonal matrix (M[0][0] = M[1][1] = 0): element y is
    // M[s][1-s] psi[y ^ v], the partner load only
    __device__ __forceinline__ static void load_global_inject_anti(const float2 *psi, const GlobalMap &g,
                                                                   const qgtl_inject_slot_t &q, float2 (&v)[kRegs]) {
        RegLoop<R - 2>::load_global_inject_anti(psi, g, q, v);
        const size_t a = global_index<R - 2>(g);
        const float4 qb = *reinterpret_cast<const float4 *>(psi + ((a ^ (size_t)q.v) & ~(size_t)1));
        const bool odd = (q.v & 1u) != 0;
        const float2 b0 = odd ? make_float2(qb.z, qb.w) : make_float2(qb.x, qb.y);
        const float2 b1 = odd ? make_float2(qb.x, qb.y) : make_float2(qb.z, qb.w);
        const bool s0 = (__popc((uint32_t)a & q.w) & 1) != 0, s1 = s0 != ((q.w & 1u) != 0);
        const float2 m01 = make_float2(q.M[1].re, q.M[1].im), m10 = make_float2(q.M[2].re, q.M[2].im);
        v[R - 2] = cmadd(make_float2(0.f, 0.f), s0 ? m10 : m01, b0);
        v[R - 1] = cmadd(make_float2(0.f, 0.f), s1 ? m10 : m01, b1);
    }
    // the same stores writing the split-plane form
    template <bool NT = false>
    __device__ __forceinline__ static void store_global_split(uint32_t *planes, const GlobalMap &g,
                                                              const float2 (&v)[kRegs]) {
        RegLoop<R - 2>::template store_global_split<NT>(planes, g, v);
        write_split_pair<NT>(planes, global_index<R - 2>(g), v[R - 2], v[R - 1]);
    }
};

template <>
struct RegLoop<0> {
    __device__ __forceinline__ static void store_lds(float2 *, const RegLayout &, const float2 (&)[kRegs]) {}
    __device__ __forceinline__ static void load_lds(const float2 *, const RegLayout &, float2 (&)[kRegs]) {}
    __device__ __forceinline__ static void store_lds_ring(float2 *, const RingLds &, const float2 (&)[kRegs]) {}
    __device__ __forceinline__ static void load_lds_ring(const float2 *, const RingLds &, float2 (&)[kRegs]) {}
    template <bool POST>
    __device__ __forceinline__ static void phase_split(float2, const qgtl_pass_desc_t &, float2 (&)[kRegs]) {}
    template <bool NT = false>
    __device__ __forceinline__ static void load_global(const float2 *, const GlobalMap &, float2 (&)[kRegs]) {}
    __device__ __forceinline__ static void load_global_inject(const float2 *, const GlobalMap &,
                                                              const qgtl_inject_slot_t &, float2 (&)[kRegs]) {}
    __device__ __forceinline__ static void load_global_inject_diag(const float2 *, const GlobalMap &,
                                                                   const qgtl_inject_slot_t &, float2 (&)[kRegs]) {}
    __device__ __forceinline__ static void load_global_inject_anti(const float2 *, const GlobalMap &,
                                                                   const qgtl_inject_slot_t &, float2 (&)[kRegs]) {}
    template <bool NT = false>
    __device__ __forceinline__ static void store_global(float2 *, const GlobalMap &, const float2 (&)[kRegs]) {}
    template <bool NT = false>
    __device__ __forceinline__ static void store_global_split(uint32_t *, const GlobalMap &, const float2 (&)[kRegs]) {}
};

constexpr int kRegTileElems = 1 << kRT;

// Phase layer of a rotation-form pass in layout L: the tables give the phase
// of the thread's register-0 element (register bits 0), the descriptor the
// register part, and scale the factors the scaled-out rotations dropped.
// The swapped pairs of d.xmask: the post layer reads the tables at j ^ xmask,
// and the layer of d.xfold takes the sign (-1)^popcount(j & xmask) (its
// register part is in reg_post / reg_pre).
// f: a factor common to the whole tile (the layer moved in from the partner
// pass, or 1).
template <bool POST>
__device__ __forceinline__ void phase_layer(const float2 *tab, const RegLayout &L, const qgtl_pass_desc_t &d,
                                            float scale, float2 f, float2 (&v)[kRegs]) {
    const uint32_t xm = (uint32_t)d.xmask;
    const bool signed_layer = POST ? d.xfold == 0 : d.xfold == 1;
    float2 pt = cmul(phase_at(tab, POST ? (L.tdep ^ xm) : L.tdep), f);
    const float sc = (signed_layer && (__popc(L.tdep & xm) & 1)) ? -scale : scale;
    pt.x *= sc;
    pt.y *= sc;
    RegLoop<kRegs>::template phase_split<POST>(pt, d, v);
}

// A slot whose matrix has a zero diagonal: an X/Y-plane generator, e.g. the
// derivative of an RY whose later gates on its qubit are Z rotations.  The
// test reads the slot's fields, the same for every thread of the workgroup.
__device__ __forceinline__ bool inject_anti(const qgtl_inject_slot_t &q) {
    return q.M[0].re == 0.f && q.M[0].im == 0.f && q.M[3].re == 0.f && q.M[3].im == 0.f;
}

// The partner pass's layer as one factor for tile r: the product over the
// moved gates of tf[i][bit tf_bit[i] of r] (1 when nothing was moved in).
// Fixed trip count: the bit positions load together, then the selected
// factors (a computed index, never a select between two loaded values); the
// entries past tf_n are zero-filled and multiplied in as 1.  It runs after
// the tile's loads are issued, so its scalar loads overlap them.
__device__ __forceinline__ float2 tile_factor(const qgtl_pass_desc_t &d, uint32_t r) {
    float2 f = make_float2(1.f, 0.f);
    if (d.form != 1 || d.nolayer != 0) return f;
    for (int i = 0; i < d.tf_n; i++) {
        const qgtl_c64_t q = d.tf[i][(r >> d.tf_bit[i]) & 1u];
        f = cmul(f, make_float2(q.re, q.im));
    }
    return f;
}

// Output high bits of a ring store from tile r (bits 0..12 of an n-qubit
// state): prefix_xor(r), complemented (xor ones) when output bit 12 is set;
// par_r = parity(r) enters output bit 0.
struct RingFrame {
    uint32_t pre, ones, par_r;
};

__device__ __forceinline__ RingFrame ring_frame(int n, uint32_t r) {
    const int nh = n - kRT;
    uint32_t pre = r;
    pre ^= pre << 1; pre ^= pre << 2; pre ^= pre << 4; pre ^= pre << 8; pre ^= pre << 16;
    const uint32_t ones = (nh >= 32) ? 0xffffffffu : ((1u << nh) - 1u);
    return RingFrame{pre & ones, ones, (uint32_t)__popc(r) & 1u};
}

// Per-workgroup set-up of a register pass: the phase tables.  The host-built
// tables (2 kPhaseTab complex64, post then pre) are fetched as one 16-byte
// word per thread before the tile's global loads are issued and written to
// LDS after them, so the fetch and the barrier overlap the tile's latency.
__device__ __forceinline__ float4 fetch_tables(const qgtl_pass_desc_t &d, const float2 *gtabs) {
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (d.form == 1 && d.nolayer == 0 && gtabs && threadIdx.x < (unsigned)kPhaseTab)
        q = reinterpret_cast<const float4 *>(gtabs)[threadIdx.x];
    return q;
}

__device__ __forceinline__ void pass_setup(const qgtl_pass_desc_t &d, float2 *post_tab, float2 *pre_tab,
                                           const float2 *gtabs, float4 q) {
    static_assert(kPhaseTab % 2 == 0 && kPhaseTab <= kThreads, "one 16-byte table word per thread");
    if (d.form == 1 && d.nolayer == 0) {
        if (gtabs) {
            const uint32_t e = 2u * threadIdx.x;
            if (threadIdx.x < (unsigned)kPhaseTab) {
                float2 *w = e < (uint32_t)kPhaseTab ? post_tab + e : pre_tab + (e - kPhaseTab);
                w[0] = make_float2(q.x, q.y);
                w[1] = make_float2(q.z, q.w);
            }
        } else {
            build_phase_tables(d, post_tab, pre_tab);
        }
    }
    __syncthreads();
}

// Where tile r of a state is read from and written to, in the load/store
// layout L0: gin for the load (the ring map for a gathering load), gout for
// the store (the ring map for a ring store).
struct TileMaps {
    RingFrame f;
    GlobalMap gin, gout;
};

__device__ __forceinline__ TileMaps tile_maps(const qgtl_pass_desc_t &d, uint32_t r, const RegLayout &L0) {
    // state offset of the tile index (fixed trip count, as deposit_hi)
    uint32_t base = 0;
    const int nrest = d.nrest;
#pragma unroll
    for (int i = 0; i < QGTL_MAX_QUBITS - kRT; i++) {
        const uint32_t bit = i < nrest ? (r >> i) & 1u : 0u;
        base |= bit << ((uint32_t)d.rest_pos[i] & 31u);
    }
    const bool ring_load = d.load == QGTL_LOAD_RING, ring_store = d.store == QGTL_STORE_RING;
    TileMaps m;
    m.f = RingFrame{0u, 0u, 0u};
    if (ring_load || ring_store) m.f = ring_frame(d.n, r);
    const GlobalMap gnat = natural_map(L0, base, d);
    const GlobalMap gring = (ring_load || ring_store) ? ring_map(L0, m.f.pre, m.f.ones) : gnat;
    m.gin = ring_load ? gring : gnat;
    m.gout = ring_store ? gring : gnat;
    return m;
}

// Everything after the tile's amplitudes sit in v (layout L0, as loaded
// through m.gin): the inverse ring's in-tile part, the phase layers and gate
// phases, the ring or natural store.  SPLIT: dst receives the split-plane
// form of the result (a final pass).
// The phases of a register pass: phase p > 0 first moves the amplitudes into
// its layout through the LDS tile; ROT: rotation form, else general matrices.
// A rotation-form pass of compile-time shape SHAPE: phase PH (and on).
template <uint32_t SHAPE, int PH>
__device__ __forceinline__ void run_phases_shaped(const qgtl_pass_desc_t &d, float2 (&v)[kRegs], RegLayout &Lc,
                                                  float2 *tile, const PhaseParams &cur) {
    if constexpr (PH < shape_phases(SHAPE)) {
        const uint32_t t = threadIdx.x;
        PhaseParams nxt = cur;
        if constexpr (PH + 1 < shape_phases(SHAPE)) nxt = load_phase(d, PH + 1);
        if constexpr (PH > 0) {
            const RegLayout Ln = make_layout(cur.rbits, cur.tbits, t);
            RegLoop<kRegs>::store_lds(tile, Lc, v);
            __syncthreads();
            RegLoop<kRegs>::load_lds(tile, Ln, v);
            __syncthreads();
            Lc = Ln;
        }
        apply_phase_rots_m<shape_mask(SHAPE, PH)>(v, cur);
        run_phases_shaped<SHAPE, PH + 1>(d, v, Lc, tile, nxt);
    }
}

template <bool ROT>
__device__ __forceinline__ void run_phases(const qgtl_pass_desc_t &d, float2 (&v)[kRegs], RegLayout &Lc,
                                           float2 *tile) {
    const uint32_t t = threadIdx.x;
    PhaseParams cur = load_phase(d, 0);
    for (int ph = 0; ph < d.nphase; ph++) {
        PhaseParams nxt = cur;
        if (ph + 1 < d.nphase) nxt = load_phase(d, ph + 1);
        if (ph > 0) {
            const RegLayout Ln = make_layout(cur.rbits, cur.tbits, t);
            RegLoop<kRegs>::store_lds(tile, Lc, v);
            __syncthreads();
            RegLoop<kRegs>::load_lds(tile, Ln, v);
            __syncthreads();
            Lc = Ln;
        }
        if constexpr (ROT) {
            apply_phase_rots(v, cur);
        } else {
            const int mask = cur.mask;
            apply_phase_mat<0>(v, d, ph, mask);
            apply_phase_mat<1>(v, d, ph, mask);
            apply_phase_mat<2>(v, d, ph, mask);
            apply_phase_mat<3>(v, d, ph, mask);
            apply_phase_mat<4>(v, d, ph, mask);
        }
        cur = nxt;
    }
}

template <bool SPLIT, bool NT = false, uint32_t SHAPE = 0>
__device__ __forceinline__ void pass_compute_store(const qgtl_pass_desc_t &d, float2 (&v)[kRegs], const TileMaps &m,
                                                   const RegLayout &L0, float2 tfac, float2 *dst, float2 *tile,
                                                   const float2 *post_tab, const float2 *pre_tab) {
    const uint32_t t = threadIdx.x;
    const bool rotf = d.form == 1;
    // phase layers: none when the partner pass applies this pass's layer; the
    // scale and the partner's moved-in layer go with the post layer if there
    // is one, else with the pre layer
    const bool layers = rotf && d.nolayer == 0;
    const bool post_on = layers && (d.has_post || !d.has_pre), pre_on = layers && d.has_pre;
    const bool ring_load = d.load == QGTL_LOAD_RING, ring_store = d.store == QGTL_STORE_RING;
    RegLayout Lc = L0;
    if (ring_load) {
        // inverse ring: tile r of the output was gathered from the two half
        // tiles the ring store writes tile r to (coalesced); undo the in-tile
        // part of the map through LDS into a layout with phase 0's register bits
        RegLoop<kRegs>::store_lds(tile, L0, v);
        __syncthreads();
        Lc = make_layout(d.ph_rbits[0], d.gather_tbits, t);
        RegLoop<kRegs>::load_lds_ring(tile, ring_lds(Lc, m.f.par_r), v);
        __syncthreads();
    }
    if (pre_on)
        phase_layer<false>(pre_tab, Lc, d, post_on ? 1.f : d.scale, post_on ? make_float2(1.f, 0.f) : tfac, v);

    // one phase loop per form, so the rotation form's straight-line gates share
    // no joins with the general matrices' branches
    if constexpr (SHAPE != 0) run_phases_shaped<SHAPE, 0>(d, v, Lc, tile, load_phase(d, 0));
    else if (rotf) run_phases<true>(d, v, Lc, tile);
    else run_phases<false>(d, v, Lc, tile);
    if (post_on) phase_layer<true>(post_tab, Lc, d, d.scale, tfac, v);

    // the last transpose stores element j at tile index j ^ xmask (the pair
    // swaps of the rotation form); without one (a single natural phase) a
    // swap adds it
    const uint32_t xm = rotf ? (uint32_t)d.xmask : 0u;
    if (ring_store) {
        RegLoop<kRegs>::store_lds_ring(tile, ring_lds(Lc, m.f.par_r, xm), v);
        __syncthreads();
        RegLoop<kRegs>::load_lds(tile, L0, v);
    } else if (d.nphase > 1 || ring_load || xm) {
        RegLayout Ls = Lc;
        Ls.lbase ^= (uint32_t)rpad(xm) << 3;
        RegLoop<kRegs>::store_lds(tile, Ls, v);
        __syncthreads();
        RegLoop<kRegs>::load_lds(tile, L0, v);
    }
    if constexpr (SPLIT)
        RegLoop<kRegs>::template store_global_split<NT>(reinterpret_cast<uint32_t *>(dst), m.gout, v);
    else
        RegLoop<kRegs>::template store_global<NT>(dst, m.gout, v);
}

// One tile r of one state.  inj != nullptr: src is the primal state the
// derivative state *inj is injected from while loading.  The global loads go
// out first; the phase tables are set up while they are in flight.
template <bool SPLIT = false, bool NT = false, uint32_t SHAPE = 0>
__device__ __forceinline__ void pass_regs_body(const qgtl_pass_desc_t &d, const float2 *__restrict__ src,
                                               float2 *dst, uint32_t r, float2 *tile, float2 *post_tab,
                                               float2 *pre_tab, const float2 *gtabs = nullptr,
                                               const qgtl_inject_slot_t *inj = nullptr) {
    const float4 tq = fetch_tables(d, gtabs);
    const RegLayout L0 = make_layout(d.ph_rbits[0], d.ph_tbits[0], threadIdx.x);
    const TileMaps m = tile_maps(d, r, L0);
    float2 v[kRegs];
    if (inj && inj->v == 0) RegLoop<kRegs>::load_global_inject_diag(src, m.gin, *inj, v);
    else if (inj && inject_anti(*inj)) RegLoop<kRegs>::load_global_inject_anti(src, m.gin, *inj, v);
    else if (inj) RegLoop<kRegs>::load_global_inject(src, m.gin, *inj, v);
    else RegLoop<kRegs>::template load_global<NT>(src, m.gin, v);
    const float2 tfac = tile_factor(d, r);
    pass_setup(d, post_tab, pre_tab, gtabs, tq);
    pass_compute_store<SPLIT, NT, SHAPE>(d, v, m, L0, tfac, dst, tile, post_tab, pre_tab);
}

template <bool SPLIT, bool NT = false, uint32_t SHAPE = 0>
__global__ __launch_bounds__(kThreads) void pass_c64_regs_kernel(const qgtl_pass_desc_t d,
                                                                 const float2 *__restrict__ src_all,
                                                                 float2 *dst_all, size_t stride,
                                                                 const float2 *__restrict__ gtabs) {
    __shared__ float2 tile[kRegTileElems];
    __shared__ float2 tabs[2 * kPhaseTab];
    pass_regs_body<SPLIT, NT, SHAPE>(d, src_all + (size_t)blockIdx.y * stride, dst_all + (size_t)blockIdx.y * stride,
                              blockIdx.x, tile, tabs, tabs + kPhaseTab, gtabs);
}

// The first pass of a derivative-state chunk with the injection folded into
// its load: state s of the launch is slot inj[s] applied to psi (one state,
// shared), so the injected states are never written to or read from HBM.
constexpr uint32_t kXcd = 8;   // XCDs; workgroups are dispatched to them round-robin

template <bool SPLIT, uint32_t SHAPE = 0>
__global__ __launch_bounds__(kThreads) void pass_c64_inject_kernel(const qgtl_pass_desc_t d,
                                                                   const float2 *__restrict__ psi,
                                                                   const qgtl_inject_slot_t *__restrict__ inj,
                                                                   float2 *dst_all, size_t stride,
                                                                   const float2 *__restrict__ gtabs,
                                                                   uint32_t nstates, uint32_t ntiles) {
    __shared__ float2 tile[kRegTileElems];
    __shared__ float2 tabs[2 * kPhaseTab];
    // one-dimensional grid of nstates x ntiles workgroups.  Workgroup b runs on
    // XCD b % 8 (round-robin dispatch); with ntiles a multiple of 8, XCD x takes
    // the tiles t = x (mod 8) in order, all states of a tile back to back, so
    // each XCD's L2 fetches a primal tile (and its partner tile) once instead of
    // every XCD fetching every tile
    const uint32_t b = blockIdx.x;
    uint32_t s, t;
    if ((ntiles & (kXcd - 1)) == 0) {
        const uint32_t slot = b / kXcd;
        s = slot % nstates;
        t = (slot / nstates) * kXcd + b % kXcd;
    } else {
        s = b % nstates;
        t = b / nstates;
    }
    const qgtl_inject_slot_t q = inj[s];
    pass_regs_body<SPLIT, true, SHAPE>(d, psi, dst_all + (size_t)s * stride, t, tile, tabs, tabs + kPhaseTab, gtabs,
                                       &q);
}

// Derivative states in the frame after a block (see qgtl_inject_desc_t), with
// FP64 per-tile sums of <d|psi>, |d|^2 and |psi|^2.
constexpr int kInjectPerThread = 8;
constexpr int kInjectTile = kThreads * kInjectPerThread;
// single-plan injections: 32 amplitudes per thread, so the loads of a
// workgroup are in flight together and its FP64 reduction is amortised
constexpr int kInjectPerThreadPlan = 32;
constexpr int kInjectTilePlan = kThreads * kInjectPerThreadPlan;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Tile `tile` of derivative state k; its four sums go to out[0..3].  SPLIT:
// dst receives the split-plane form (four fp16 planes of N values).
template <bool SPLIT = false, bool STORE = true, int PER = kInjectPerThread, bool SUMS = true>
__device__ __forceinline__ void inject_body(const qgtl_inject_desc_t &d, const float2 *__restrict__ psi,
                                            float2 *dst, int k, uint32_t tile, double *out,
                                            double (*red)[4]) {
    const uint32_t N = 1u << d.n;
    const uint32_t v = d.v[k], w = d.w[k];
    const float2 a00 = make_float2(d.A[k][0].re, d.A[k][0].im), a01 = make_float2(d.A[k][1].re, d.A[k][1].im);
    const float2 a10 = make_float2(d.A[k][2].re, d.A[k][2].im), a11 = make_float2(d.A[k][3].re, d.A[k][3].im);
    // two adjacent amplitudes per thread: one 16-byte load per operand, 16-byte
    // (complex64) or 4-byte-pair (split-plane) stores; the FP64 sums run over the
    // thread's amplitudes in a fixed order
    static_assert(PER % 2 == 0, "amplitude pairs");
    const bool odd = (v & 1u) != 0, wodd = (w & 1u) != 0;
    double sr = 0.0, si = 0.0, sd = 0.0, sp = 0.0;
#pragma unroll
    for (int i = 0; i < PER / 2; i++) {
        const uint32_t y = tile * (PER * kThreads) + 2 * (i * kThreads + threadIdx.x);
        if (y < N) {                                          // N even: y + 1 < N too
            const float4 qa = *reinterpret_cast<const float4 *>(psi + y);
            const float4 qb = *reinterpret_cast<const float4 *>(psi + ((y ^ v) & ~1u));
            const float2 a0 = make_float2(qa.x, qa.y), a1 = make_float2(qa.z, qa.w);
            const float2 b0 = odd ? make_float2(qb.z, qb.w) : make_float2(qb.x, qb.y);
            const float2 b1 = odd ? make_float2(qb.x, qb.y) : make_float2(qb.z, qb.w);
            const bool s0 = (__popc(y & w) & 1) != 0, s1 = s0 != wodd;
            const float2 o0 = cmadd(cmadd(make_float2(0.f, 0.f), s0 ? a11 : a00, a0), s0 ? a10 : a01, b0);
            const float2 o1 = cmadd(cmadd(make_float2(0.f, 0.f), s1 ? a11 : a00, a1), s1 ? a10 : a01, b1);
            if constexpr (!STORE) {
                // sums only: the state itself is formed by the first pass that reads it
            } else if constexpr (SPLIT) {
                write_split_pair<true>(reinterpret_cast<uint32_t *>(dst), y, o0, o1);
            } else {
                gstore4<true>(dst + y, o0, o1);
            }
            if constexpr (SUMS) {
                const double dr0 = o0.x, di0 = o0.y, pr0 = a0.x, pi0 = a0.y;
                const double dr1 = o1.x, di1 = o1.y, pr1 = a1.x, pi1 = a1.y;
                sr += dr0 * pr0 + di0 * pi0;
                si += dr0 * pi0 - di0 * pr0;
                sd += dr0 * dr0 + di0 * di0;
                sp += pr0 * pr0 + pi0 * pi0;
                sr += dr1 * pr1 + di1 * pi1;
                si += dr1 * pi1 - di1 * pr1;
                sd += dr1 * dr1 + di1 * di1;
                sp += pr1 * pr1 + pi1 * pi1;
            }
        }
    }
    if constexpr (!SUMS) return;
    sr = wave_sum(sr); si = wave_sum(si); sd = wave_sum(sd); sp = wave_sum(sp);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = sr; red[wave][1] = si; red[wave][2] = sd; red[wave][3] = sp;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int wv = 0; wv < kThreads / 64; wv++) s += red[wv][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

template <bool SPLIT, bool STORE = true, bool SUMS = true>
__global__ __launch_bounds__(kThreads) void inject_c64_kernel(const qgtl_inject_desc_t d,
                                                              const float2 *__restrict__ psi,
                                                              float2 *dst_all, size_t stride,
                                                              double *partials) {
    __shared__ double red[kThreads / 64][4];
    const int k = blockIdx.y;
    inject_body<SPLIT, STORE, kInjectPerThreadPlan, SUMS>(
        d, psi, STORE ? dst_all + (size_t)k * stride : nullptr, k, blockIdx.x,
        SUMS ? partials + ((size_t)k * gridDim.x + blockIdx.x) * 4 : nullptr, red);
}

// One complex64 state in split-plane form (the Gram's operand form).
__global__ __launch_bounds__(kThreads) void c64_to_split_kernel(const float2 *__restrict__ src, uint32_t *planes,
                                                                size_t N) {
    for (size_t g = 2 * ((size_t)blockIdx.x * kThreads + threadIdx.x); g < N; g += 2 * (size_t)gridDim.x * kThreads) {
        const float4 q = *reinterpret_cast<const float4 *>(src + g);
        write_split_pair(planes, g, make_float2(q.x, q.y), make_float2(q.z, q.w));
    }
}

// out[c] = sum over t of pk[t * 4 + c] (fixed tree order, FP64).
__device__ __forceinline__ void reduce_body(const double *__restrict__ pk, int tiles, double *out,
                                            double (*red)[4]) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < tiles; t += kThreads)
#pragma unroll
        for (int c = 0; c < 4; c++) s[c] += pk[(size_t)t * 4 + c];
#pragma unroll
    for (int c = 0; c < 4; c++) red[threadIdx.x][c] = s[c];
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
#pragma unroll
            for (int c = 0; c < 4; c++) red[threadIdx.x][c] += red[threadIdx.x + h][c];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = red[0][threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void reduce_partials_kernel(const double *__restrict__ partials,
                                                                   int tiles, double *out) {
    __shared__ double red[kThreads][4];
    const int k = blockIdx.x;
    reduce_body(partials + (size_t)k * tiles * 4, tiles, out + (size_t)k * 4, red);
}

// H psi for H = sum_t c_t P_t, P_t = i^{ny} X^x Z^z (ny = popcount(x & z)):
// (P psi)[y] = i^{ny} (-1)^{|(y ^ x) & z|} psi[y ^ x].  QGTL's Pauli-string
// Hamiltonians (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:320-380).
__global__ __launch_bounds__(kThreads) void pauli_apply_c64_kernel(const float2 *__restrict__ psi, float2 *out,
                                                                   uint32_t N, int nterms,
                                                                   const qgtl_pauli_term_t *__restrict__ terms) {
    for (uint32_t y = blockIdx.x * kThreads + threadIdx.x; y < N; y += gridDim.x * kThreads) {
        double ar = 0.0, ai = 0.0;
        for (int t = 0; t < nterms; t++) {
            const uint32_t x = terms[t].xmask, z = terms[t].zmask;
            const uint32_t src = y ^ x;
            const float2 v = psi[src];
            const int ny = __popc(x & z) & 3;
            const int neg = __popc(src & z) & 1;
            // phase = i^ny * (-1)^neg applied to c_t * v
            double pr = terms[t].coeff_re * v.x - terms[t].coeff_im * v.y;
            double pi = terms[t].coeff_re * v.y + terms[t].coeff_im * v.x;
            double qr, qi;
            switch (ny) {
            case 0: qr = pr; qi = pi; break;
            case 1: qr = -pi; qi = pr; break;
            case 2: qr = -pr; qi = -pi; break;
            default: qr = pi; qi = -pr; break;
            }
            if (neg) { qr = -qr; qi = -qi; }
            ar += qr;
            ai += qi;
        }
        out[y] = make_float2((float)ar, (float)ai);
    }
}

// FP64 sums per (state k, tile): <D_k|v> (re, im), |D_k|^2, |v|^2.
// PLANES: D holds split-plane states, read back as (hi + lo) 2^-15 (exact in fp32).
template <bool PLANES>
__global__ __launch_bounds__(kThreads) void dot_states_c64_kernel(const float2 *__restrict__ D, size_t stride,
                                                                  const float2 *__restrict__ v, uint32_t N,
                                                                  double *partials) {
    const int k = blockIdx.y;
    const float2 *d = D + (size_t)k * stride;
    double sr = 0.0, si = 0.0, sd = 0.0, sv = 0.0;
#pragma unroll
    for (int i = 0; i < kInjectPerThread; i++) {
        const uint32_t y = blockIdx.x * kInjectTile + i * kThreads + threadIdx.x;
        if (y < N) {
            float2 a;
            if constexpr (PLANES) {
                // hi word (Re hi, Im hi) and lo word of amplitude y
                const _Float16 *pl = reinterpret_cast<const _Float16 *>(d) + (size_t)(y >> 6) * 256 + 2 * (y & 63);
                const float inv = 1.0f / qgtl::kSplitScale;
                a = make_float2(((float)pl[0] + (float)pl[128]) * inv, ((float)pl[1] + (float)pl[129]) * inv);
            } else {
                a = d[y];
            }
            const float2 b = v[y];
            const double ar = a.x, aim = a.y, br = b.x, bi = b.y;
            sr += ar * br + aim * bi;
            si += ar * bi - aim * br;
            sd += ar * ar + aim * aim;
            sv += br * br + bi * bi;
        }
    }
    __shared__ double red[kThreads / 64][4];
    sr = wave_sum(sr); si = wave_sum(si); sd = wave_sum(sd); sv = wave_sum(sv);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = sr; red[wave][1] = si; red[wave][2] = sd; red[wave][3] = sv;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int wv = 0; wv < kThreads / 64; wv++) s += red[wv][threadIdx.x];
        partials[((size_t)k * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = s;
    }
}

// energy = <psi|H psi> / |psi|^2; grad_k = 2 Re <d_k|H psi> / (2 |d_k| |psi|)
// (normalised with the exact |d_k| = 1/2, |psi| = 1 as the QGT is).
__global__ void energy_gradient_finalize_kernel(const double *__restrict__ dk_h, const double *__restrict__ psi_h,
                                                int P, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const double npsi = psi_h[2];
    if (k == 0) out[0] = psi_h[0] / npsi;
    if (k < P) out[1 + k] = 2.0 * dk_h[4 * k] * 0.5 / sqrt(dk_h[4 * k + 2] * npsi);
}

__global__ void fill_zero_state_c64_kernel(float2 *psi, uint32_t N) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        psi[i] = make_float2(i == 0 ? 1.f : 0.f, 0.f);
}

__global__ void cnot_c64_kernel(float2 *base, size_t stride, uint32_t quarter, int c, int t) {
    float2 *psi = base + (size_t)blockIdx.y * stride;
    const int lo = min(c, t), hi = max(c, t);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < quarter; x += gridDim.x * blockDim.x) {
        const uint32_t b0 = insert_zero(insert_zero(x, lo), hi);
        const uint32_t i = b0 | (1u << c), j = i | (1u << t);
        const float2 a = psi[i];
        psi[i] = psi[j];
        psi[j] = a;
    }
}

__global__ void swap_c64_kernel(float2 *base, size_t stride, uint32_t quarter, int qa, int qb) {
    float2 *psi = base + (size_t)blockIdx.y * stride;
    const int lo = min(qa, qb), hi = max(qa, qb);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < quarter; x += gridDim.x * blockDim.x) {
        const uint32_t b0 = insert_zero(insert_zero(x, lo), hi);
        const uint32_t i = b0 | (1u << qa), j = b0 | (1u << qb);
        const float2 a = psi[i];
        psi[i] = psi[j];
        psi[j] = a;
    }
}

__global__ void c64_to_c128_kernel(const float2 *src, double2 *dst, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = make_double2(src[i].x, src[i].y);
}


// ---------------------------------------------------------------------------
// Grouped launches: the same work for many instances in one launch, with a
// grid-stride loop over units (see qgtl_pass_item_t and friends).
__global__ __launch_bounds__(kThreads) void fill_grouped_kernel(const qgtl_fill_item_t *__restrict__ items,
                                                                int nitems, uint32_t units) {
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_fill_item_t it = items[qgtl::find_item(items, nitems, u)];
        float2 *psi = reinterpret_cast<float2 *>(it.psi);
        const uint32_t base = (u - it.unit_begin) * kInjectTile;
        for (uint32_t j = threadIdx.x; j < (uint32_t)kInjectTile; j += kThreads) {
            const uint32_t y = base + j;
            if (y < it.N) psi[y] = make_float2(y == 0 ? 1.f : 0.f, 0.f);
        }
    }
}

// VARIANT 1: every item runs the register-phase body; 0: the LDS body
// (separate instantiations keep each body's register allocation).
template <int VARIANT>
__global__ __launch_bounds__(kThreads, 2) void pass_grouped_kernel(const qgtl_pass_item_t *__restrict__ items,
                                                                int nitems, uint32_t units) {
    __shared__ float2 tile[kLdsTileElems];
    __shared__ uint32_t hi_off[1 << kHiTableBits];
    __shared__ float2 tabs[2 * kPhaseTab];
    static_assert(kLdsTileElems >= kRegTileElems, "LDS tile too small for the register pass");
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_pass_item_t it = items[qgtl::find_item(items, nitems, u)];
        const uint32_t local = u - it.unit_begin;
        const uint32_t r = local % it.ntiles, s = local / it.ntiles;
        __syncthreads();   // the previous unit is done with LDS
        // the descriptor table is written before the launch: read it through
        // the constant address space (scalar loads, no copy)
        const qgtl_pass_desc_t &sd = *qgtl::const_ptr<qgtl_pass_desc_t>(qgtl::uniform(it.desc));
        const float2 *src = qgtl::global_ptr<const float2>(qgtl::uniform(it.src)) + (size_t)s * it.stride;
        float2 *dst = qgtl::global_ptr<float2>(qgtl::uniform(it.dst)) + (size_t)s * it.stride;
        const uint64_t ta = qgtl::uniform(it.tabs);
        const float2 *gt = ta ? qgtl::global_ptr<const float2>(ta) : nullptr;
        if (VARIANT == 2) pass_regs_body<true, true>(sd, src, dst, r, tile, tabs, tabs + kPhaseTab, gt);
        else if (VARIANT == 1) pass_regs_body<false, true>(sd, src, dst, r, tile, tabs, tabs + kPhaseTab, gt);
        else pass_lds_body(sd, src, dst, r, tile, hi_off, tabs, tabs + kPhaseTab, gt);
    }
}

// Phase tables of descriptor blockIdx.x (grid-stride over descriptors), the
// same entries build_phase_tables puts in LDS, written to global memory.
__global__ __launch_bounds__(kThreads) void phase_tables_grouped_kernel(const qgtl_pass_desc_t *__restrict__ descs,
                                                                       int ndesc, float2 *tabs) {
    __shared__ float2 post_tab[kPhaseTab], pre_tab[kPhaseTab];
    for (int i = blockIdx.x; i < ndesc; i += gridDim.x) {
        const qgtl_pass_desc_t &d = *qgtl::const_ptr<qgtl_pass_desc_t>((uint64_t)(uintptr_t)(descs + i));
        if (d.form != 1) continue;                           // uniform: the whole workgroup skips
        __syncthreads();                                     // the previous descriptor's copy-out is done
        build_phase_tables(d, post_tab, pre_tab);
        __syncthreads();
        float2 *out = tabs + (size_t)i * 2 * kPhaseTab;
        for (int t = threadIdx.x; t < 2 * kPhaseTab; t += kThreads)
            out[t] = t < kPhaseTab ? post_tab[t] : pre_tab[t - kPhaseTab];
    }
}

__global__ __launch_bounds__(kThreads) void inject_grouped_kernel(const qgtl_inject_item_t *__restrict__ items,
                                                                  int nitems, uint32_t units) {
    __shared__ double red[kThreads / 64][4];
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_inject_item_t it = items[qgtl::find_item(items, nitems, u)];
        const uint32_t local = u - it.unit_begin;
        const uint32_t t = local % it.tiles;
        const int k = (int)(local / it.tiles);
        __syncthreads();
        const qgtl_inject_desc_t &d = *reinterpret_cast<const qgtl_inject_desc_t *>(it.desc);
        float2 *dst = reinterpret_cast<float2 *>(it.dst) + (size_t)k * it.stride;
        if (it.partials)        // FP64 sums of <d|psi>, |d|^2, |psi|^2 per tile as well
            inject_body(d, reinterpret_cast<const float2 *>(it.psi), dst, k, t,
                        reinterpret_cast<double *>(it.partials) + ((size_t)k * it.tiles + t) * 4, red);
        else if (it.pad & QGTL_INJECT_ITEM_SPLIT)   // the state in split-plane form
            inject_body<true, true, kInjectPerThread, false>(d, reinterpret_cast<const float2 *>(it.psi), dst, k, t,
                                                             nullptr, red);
        else
            inject_body<false, true, kInjectPerThread, false>(d, reinterpret_cast<const float2 *>(it.psi), dst, k, t,
                                                              nullptr, red);
    }
}

__global__ __launch_bounds__(kThreads) void reduce_grouped_kernel(const qgtl_reduce_item_t *__restrict__ items,
                                                                  int nitems, uint32_t units) {
    __shared__ double red[kThreads][4];
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const qgtl_reduce_item_t it = items[qgtl::find_item(items, nitems, u)];
        const uint32_t k = u - it.unit_begin;
        __syncthreads();
        reduce_body(reinterpret_cast<const double *>(it.partials) + (size_t)k * it.tiles * 4, (int)it.tiles,
                    reinterpret_cast<double *>(it.out) + (size_t)k * 4, red);
    }
}

unsigned grouped_grid(uint32_t units, int grid) {
    if (grid <= 0 || (uint32_t)grid > units) return units;
    return (unsigned)grid;
}

__global__ void fill_f32_kernel(float *dst, size_t count, float value) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

__global__ void copy_f32x4_kernel(float4 *dst, const float4 *__restrict__ src, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

unsigned grid_for(size_t work, unsigned threads = 256, unsigned cap = 8192) {
    size_t g = (work + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace

extern "C" {

#define QGTL_GROUPED_ARGS_OK(items, nitems, units)                                       \
    do {                                                                                  \
        if (!(items) || (nitems) < 1 || (units) < 1)                                      \
            return qgtl::set_error_msg(-118, "grouped launch: empty item list");          \
    } while (0)

int qgtl_hip_fill_grouped(const qgtl_fill_item_t *items, int nitems, uint32_t units, int grid) {
    QGTL_GROUPED_ARGS_OK(items, nitems, units);
    hipLaunchKernelGGL(fill_grouped_kernel, dim3(grouped_grid(units, grid)), dim3(kThreads), 0, qgtl::stream(),
                       items, nitems, units);
    QGTL_LAUNCH_CHECK("fill_grouped");
    return 0;
}

int qgtl_hip_phase_tables_grouped(const qgtl_pass_desc_t *descs, int ndesc, void *tabs) {
    if (!descs || !tabs || ndesc < 0) return qgtl::set_error_msg(-118, "qgtl_hip_phase_tables_grouped: bad arguments");
    if (ndesc == 0) return 0;
    const int grid = ndesc < 4096 ? ndesc : 4096;
    hipLaunchKernelGGL(phase_tables_grouped_kernel, dim3(grid), dim3(kThreads), 0, qgtl::stream(), descs, ndesc,
                       (float2 *)tabs);
    QGTL_LAUNCH_CHECK("phase_tables_grouped");
    return 0;
}

int qgtl_hip_pass_grouped(const qgtl_pass_item_t *items, int nitems, uint32_t units, int grid, int variant) {
    QGTL_GROUPED_ARGS_OK(items, nitems, units);
    if (variant == 2)
        hipLaunchKernelGGL(pass_grouped_kernel<2>, dim3(grouped_grid(units, grid)), dim3(kThreads), 0,
                           qgtl::stream(), items, nitems, units);
    else if (variant)
        hipLaunchKernelGGL(pass_grouped_kernel<1>, dim3(grouped_grid(units, grid)), dim3(kThreads), 0,
                           qgtl::stream(), items, nitems, units);
    else
        hipLaunchKernelGGL(pass_grouped_kernel<0>, dim3(grouped_grid(units, grid)), dim3(kThreads), 0,
                           qgtl::stream(), items, nitems, units);
    QGTL_LAUNCH_CHECK("pass_grouped");
    return 0;
}

int qgtl_hip_inject_grouped(const qgtl_inject_item_t *items, int nitems, uint32_t units, int grid) {
    QGTL_GROUPED_ARGS_OK(items, nitems, units);
    hipLaunchKernelGGL(inject_grouped_kernel, dim3(grouped_grid(units, grid)), dim3(kThreads), 0, qgtl::stream(),
                       items, nitems, units);
    QGTL_LAUNCH_CHECK("inject_grouped");
    return 0;
}

int qgtl_hip_reduce_grouped(const qgtl_reduce_item_t *items, int nitems, uint32_t units, int grid) {
    QGTL_GROUPED_ARGS_OK(items, nitems, units);
    hipLaunchKernelGGL(reduce_grouped_kernel, dim3(grouped_grid(units, grid)), dim3(kThreads), 0, qgtl::stream(),
                       items, nitems, units);
    QGTL_LAUNCH_CHECK("reduce_grouped");
    return 0;
}

int qgtl_hip_fill_zero_state_c64(void *psi, int n) {
    const uint32_t N = 1u << n;
    hipLaunchKernelGGL(fill_zero_state_c64_kernel, dim3(grid_for(N)), dim3(256), 0, qgtl::stream(),
                       (float2 *)psi, N);
    QGTL_LAUNCH_CHECK("fill_zero_state_c64");
    return 0;
}

// The compile-time phase shape matching a rotation-form register pass (0: the
// general form).
static uint32_t pass_shape(const qgtl_pass_desc_t *desc) {
    if (desc->form != 1 || desc->nphase < 1 || desc->nphase > 3) return 0;
    uint32_t m[3] = {0, 0, 0};
    for (int p = 0; p < desc->nphase; p++) m[p] = (uint32_t)desc->ph_mask[p] & 31u;
    const uint32_t code = shape_code(desc->nphase, m[0], m[1], m[2]);
    return (code == kShape3Full || code == kShape2Hi7) ? code : 0;
}

static bool regs_pass(const qgtl_pass_desc_t *desc, int tile_variant) {
    return desc->m == kRT && desc->nphase >= 1 && desc->nphase <= QGTL_MAX_PHASES && desc->store != QGTL_STORE_PERM &&
           tile_variant != 1;
}

static bool tbits_ok(const qgtl_pass_desc_t *desc, const int32_t *tb) {
    uint32_t used = 0;
    for (int i = 0; i < 5; i++) used |= 1u << desc->ph_rbits[0][i];
    for (int i = 0; i < 8; i++) {
        if (tb[i] < 0 || tb[i] >= kRT || ((used >> tb[i]) & 1u)) return false;
        used |= 1u << tb[i];
    }
    return used == (1u << kRT) - 1u;
}

int qgtl_hip_pass_split_ok(const qgtl_pass_desc_t *desc) { return regs_pass(desc, 0) ? 1 : 0; }

int qgtl_hip_pass_c64(const qgtl_pass_desc_t *desc, const void *src, void *dst, size_t stride,
                      int nstates, int tile_variant, int split_out, const void *tabs) {
    if (desc->m < 1 || desc->m > QGTL_TILE_BITS_MAX || desc->nlow < 1 || desc->nlow > desc->m ||
        desc->m - desc->nlow > kHiTableBits || desc->ngates > desc->m || desc->nperm > QGTL_PERM_MAX ||
        desc->m + desc->nrest != desc->n || nstates < 1)
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64: inconsistent pass descriptor");
    if (desc->store == QGTL_STORE_RING && (desc->nlow != desc->m || desc->m < 2 || desc->n <= desc->m))
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64: ring store needs the low tile of a wider state");
    if (desc->store == QGTL_STORE_PERM && desc->m != desc->n)
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64: permutation store needs the whole state in one tile");
    if ((desc->store == QGTL_STORE_RING || desc->store == QGTL_STORE_PERM) && src == dst && desc->store == QGTL_STORE_RING)
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64: ring store cannot run in place");
    const dim3 grid(1u << desc->nrest, (unsigned)nstates);
    const bool regs = regs_pass(desc, tile_variant);
    if (desc->load != QGTL_LOAD_NATURAL &&
        (desc->load != QGTL_LOAD_RING || !regs || desc->nlow != kRT || desc->n <= kRT ||
         desc->store != QGTL_STORE_INPLACE || src == dst || !tbits_ok(desc, desc->gather_tbits)))
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64: the ring gather needs a register pass over bits 0..12 "
                                         "of a wider state, a natural store and separate buffers");
    if (split_out && (!regs || src == dst || stride < ((size_t)1 << desc->n)))
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64: split-plane output needs a register pass, not in place");
    const float2 *gt = (const float2 *)tabs;
    // non-temporal loads and stores: every tile is read once and written once per pass
    const uint32_t shape = regs ? pass_shape(desc) : 0;
#define QGTL_REGS_LAUNCH(SPLIT, SHAPE)                                                                         \
    hipLaunchKernelGGL((pass_c64_regs_kernel<SPLIT, true, SHAPE>), grid, dim3(kThreads), 0, qgtl::stream(), *desc, \
                       (const float2 *)src, (float2 *)dst, stride, gt)
    if (regs && split_out && shape == kShape3Full) QGTL_REGS_LAUNCH(true, kShape3Full);
    else if (regs && split_out && shape == kShape2Hi7) QGTL_REGS_LAUNCH(true, kShape2Hi7);
    else if (regs && split_out) QGTL_REGS_LAUNCH(true, 0);
    else if (regs && shape == kShape3Full) QGTL_REGS_LAUNCH(false, kShape3Full);
    else if (regs && shape == kShape2Hi7) QGTL_REGS_LAUNCH(false, kShape2Hi7);
    else if (regs) QGTL_REGS_LAUNCH(false, 0);
#undef QGTL_REGS_LAUNCH
    else
        hipLaunchKernelGGL(pass_c64_kernel, grid, dim3(kThreads), 0, qgtl::stream(), *desc,
                           (const float2 *)src, (float2 *)dst, stride, gt);
    QGTL_LAUNCH_CHECK("pass_c64");
    return 0;
}

int qgtl_hip_c64_to_split(const void *src, void *dst, int n) {
    const size_t N = (size_t)1 << n;
    if (n < 6) return qgtl::set_error_msg(-118, "qgtl_hip_c64_to_split: needs at least 64 amplitudes");
    size_t blocks = (N / 2 + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(c64_to_split_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, qgtl::stream(),
                       (const float2 *)src, (uint32_t *)dst, N);
    QGTL_LAUNCH_CHECK("c64_to_split");
    return 0;
}

int qgtl_hip_inject_sums_c64(const qgtl_inject_desc_t *desc, const void *psi, double *partials, int *tiles_out) {
    if (desc->K < 1 || desc->K > QGTL_INJECT_MAX || desc->n < 1 || desc->n > QGTL_MAX_QUBITS)
        return qgtl::set_error_msg(-118, "qgtl_hip_inject_sums_c64: bad descriptor");
    const uint32_t N = 1u << desc->n;
    const unsigned tiles = (N + kInjectTilePlan - 1) / kInjectTilePlan;
    *tiles_out = (int)tiles;
    hipLaunchKernelGGL((inject_c64_kernel<false, false>), dim3(tiles, (unsigned)desc->K), dim3(kThreads), 0,
                       qgtl::stream(), *desc, (const float2 *)psi, (float2 *)nullptr, (size_t)0, partials);
    QGTL_LAUNCH_CHECK("inject_sums_c64");
    return 0;
}

int qgtl_hip_pass_c64_inject(const qgtl_pass_desc_t *desc, const void *psi, const void *slots, void *dst,
                             size_t stride, int nstates, int split_out, const void *tabs) {
    if (desc->m != kRT || desc->nlow < 1 || desc->nlow > kRT || kRT - desc->nlow > kHiTableBits ||
        desc->ngates > kRT || desc->m + desc->nrest != desc->n || nstates < 1 || !regs_pass(desc, 0) ||
        stride < ((size_t)1 << desc->n))
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64_inject: needs a register pass over separate states");
    if (desc->nrest > 15 || nstates > 65535)
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64_inject: at most 2^15 tiles per state");
    if (((size_t)nstates << desc->nrest) * kThreads >= ((size_t)1 << 32))   // one-dimensional grid of work-items
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64_inject: states x tiles must stay below 2^24");
    if (desc->store == QGTL_STORE_RING && (desc->nlow != desc->m || desc->n <= desc->m))
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64_inject: ring store needs the low tile of a wider state");
    if (desc->load != QGTL_LOAD_NATURAL &&
        (desc->load != QGTL_LOAD_RING || desc->nlow != kRT || desc->n <= kRT || desc->store != QGTL_STORE_INPLACE ||
         !tbits_ok(desc, desc->gather_tbits)))
        return qgtl::set_error_msg(-118, "qgtl_hip_pass_c64_inject: bad ring gather");
    const uint32_t ntiles = 1u << desc->nrest;
    const dim3 grid((unsigned)nstates * ntiles);   // <= 65535 * 2^15 workgroups
    const float2 *gt = (const float2 *)tabs;
    const qgtl_inject_slot_t *sl = (const qgtl_inject_slot_t *)slots;
    const uint32_t shape = pass_shape(desc);
#define QGTL_INJECT_LAUNCH(SPLIT, SHAPE)                                                                          \
    hipLaunchKernelGGL((pass_c64_inject_kernel<SPLIT, SHAPE>), grid, dim3(kThreads), 0, qgtl::stream(), *desc,    \
                       (const float2 *)psi, sl, (float2 *)dst, stride, gt, (uint32_t)nstates, ntiles)
    if (split_out && shape == kShape3Full) QGTL_INJECT_LAUNCH(true, kShape3Full);
    else if (split_out && shape == kShape2Hi7) QGTL_INJECT_LAUNCH(true, kShape2Hi7);
    else if (split_out) QGTL_INJECT_LAUNCH(true, 0);
    else if (shape == kShape3Full) QGTL_INJECT_LAUNCH(false, kShape3Full);
    else if (shape == kShape2Hi7) QGTL_INJECT_LAUNCH(false, kShape2Hi7);
    else QGTL_INJECT_LAUNCH(false, 0);
#undef QGTL_INJECT_LAUNCH
    QGTL_LAUNCH_CHECK("pass_c64_inject");
    return 0;
}

int qgtl_hip_inject_c64(const qgtl_inject_desc_t *desc, const void *psi, void *dst, size_t stride,
                        double *partials, int *tiles_out, int split_out) {
    if (desc->K < 1 || desc->K > QGTL_INJECT_MAX || desc->n < 1 || desc->n > QGTL_MAX_QUBITS)
        return qgtl::set_error_msg(-118, "qgtl_hip_inject_c64: bad descriptor");
    const uint32_t N = 1u << desc->n;
    if (split_out && stride < N) return qgtl::set_error_msg(-118, "qgtl_hip_inject_c64: states overlap");
    if (split_out && desc->n < 6)   /* the planes are laid out in blocks of 64 amplitudes */
        return qgtl::set_error_msg(-118, "qgtl_hip_inject_c64: split-plane output needs n >= 6");
    const unsigned tiles = (N + kInjectTilePlan - 1) / kInjectTilePlan;
    *tiles_out = (int)tiles;
    const dim3 grid(tiles, (unsigned)desc->K);
    if (split_out && partials)
        hipLaunchKernelGGL(inject_c64_kernel<true>, grid, dim3(kThreads), 0, qgtl::stream(), *desc,
                           (const float2 *)psi, (float2 *)dst, stride, partials);
    else if (partials)
        hipLaunchKernelGGL(inject_c64_kernel<false>, grid, dim3(kThreads), 0, qgtl::stream(), *desc,
                           (const float2 *)psi, (float2 *)dst, stride, partials);
    else if (split_out)   /* no sums: the overlaps come from elsewhere */
        hipLaunchKernelGGL((inject_c64_kernel<true, true, false>), grid, dim3(kThreads), 0, qgtl::stream(), *desc,
                           (const float2 *)psi, (float2 *)dst, stride, partials);
    else
        hipLaunchKernelGGL((inject_c64_kernel<false, true, false>), grid, dim3(kThreads), 0, qgtl::stream(), *desc,
                           (const float2 *)psi, (float2 *)dst, stride, partials);
    QGTL_LAUNCH_CHECK("inject_c64");
    return 0;
}

int qgtl_hip_reduce_partials_f64(const double *partials, int K, int tiles, double *out) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)K), dim3(kThreads), 0, qgtl::stream(),
                       partials, tiles, out);
    QGTL_LAUNCH_CHECK("reduce_partials");
    return 0;
}

int qgtl_hip_cnot_c64(void *base, size_t stride, int nstates, int n, int c, int t) {
    if (n < 2) return qgtl::set_error_msg(-118, "cnot needs two qubits");
    const uint32_t quarter = 1u << (n - 2);
    hipLaunchKernelGGL(cnot_c64_kernel, dim3(grid_for(quarter), (unsigned)nstates), dim3(256), 0,
                       qgtl::stream(), (float2 *)base, stride, quarter, c, t);
    QGTL_LAUNCH_CHECK("cnot_c64");
    return 0;
}

int qgtl_hip_swap_c64(void *base, size_t stride, int nstates, int n, int a, int b) {
    if (n < 2) return qgtl::set_error_msg(-118, "swap needs two qubits");
    const uint32_t quarter = 1u << (n - 2);
    hipLaunchKernelGGL(swap_c64_kernel, dim3(grid_for(quarter), (unsigned)nstates), dim3(256), 0,
                       qgtl::stream(), (float2 *)base, stride, quarter, a, b);
    QGTL_LAUNCH_CHECK("swap_c64");
    return 0;
}

int qgtl_hip_pauli_apply_c64(const void *psi, void *out, int n, int nterms, const void *terms_dev) {
    const uint32_t N = 1u << n;
    hipLaunchKernelGGL(pauli_apply_c64_kernel, dim3(grid_for(N)), dim3(kThreads), 0, qgtl::stream(),
                       (const float2 *)psi, (float2 *)out, N, nterms, (const qgtl_pauli_term_t *)terms_dev);
    QGTL_LAUNCH_CHECK("pauli_apply_c64");
    return 0;
}

int qgtl_hip_dot_states_c64(const void *D, size_t stride, int K, const void *v, int n, double *partials,
                            int *tiles_out, int planes) {
    const uint32_t N = 1u << n;
    const unsigned tiles = (N + kInjectTile - 1) / kInjectTile;
    *tiles_out = (int)tiles;
    if (planes)
        hipLaunchKernelGGL(dot_states_c64_kernel<true>, dim3(tiles, (unsigned)K), dim3(kThreads), 0, qgtl::stream(),
                           (const float2 *)D, stride, (const float2 *)v, N, partials);
    else
        hipLaunchKernelGGL(dot_states_c64_kernel<false>, dim3(tiles, (unsigned)K), dim3(kThreads), 0,
                           qgtl::stream(), (const float2 *)D, stride, (const float2 *)v, N, partials);
    QGTL_LAUNCH_CHECK("dot_states_c64");
    return 0;
}

int qgtl_hip_energy_gradient_finalize(const double *dk_h, const double *psi_h, int P, double *out) {
    hipLaunchKernelGGL(energy_gradient_finalize_kernel, dim3((unsigned)((P + 255) / 256 + 1)), dim3(256), 0,
                       qgtl::stream(), dk_h, psi_h, P, out);
    QGTL_LAUNCH_CHECK("energy_gradient_finalize");
    return 0;
}

int qgtl_hip_fill_f32(void *dst, size_t count, float value) {
    hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_for(count)), dim3(256), 0, qgtl::stream(), (float *)dst, count, value);
    QGTL_LAUNCH_CHECK("fill_f32");
    return 0;
}

int qgtl_hip_copy_f32x4(void *dst, const void *src, size_t count16) {
    hipLaunchKernelGGL(copy_f32x4_kernel, dim3(grid_for(count16, 256, 16384)), dim3(256), 0, qgtl::stream(),
                       (float4 *)dst, (const float4 *)src, count16);
    QGTL_LAUNCH_CHECK("copy_f32x4");
    return 0;
}

int qgtl_hip_c64_to_c128(const void *src, void *dst, size_t count) {
    hipLaunchKernelGGL(c64_to_c128_kernel, dim3(grid_for(count)), dim3(256), 0, qgtl::stream(),
                       (const float2 *)src, (double2 *)dst, count);
    QGTL_LAUNCH_CHECK("c64_to_c128");
    return 0;
}

}  // extern "C"
