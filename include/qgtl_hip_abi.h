/*
 * qgtl_hip_abi.h -- the thin C ABI between libqgtl_mi355x's C host code and
 * its HIP kernels (the .hip files of quantum_geometric_tensor_amd/csrc/hip).
 *
 * Internal to the library: no QGTL counterpart.  Plain C types only; device
 * pointers are void*; streams and events are opaque handles.  Every function
 * returns 0 or a negative qgt_error_t code (QGT_ERROR_GPU_*).
 */
#ifndef QGTL_HIP_ABI_H
#define QGTL_HIP_ABI_H

#include <stddef.h>
#include <stdint.h>

#include "qgtl_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

#define QGTL_TILE_BITS_MAX 13    /* LDS tile of a sweep pass: up to 2^13 complex64 amplitudes */
#define QGTL_MAX_QUBITS 30
#define QGTL_PERM_MAX 32
#define QGTL_INJECT_MAX 64
#define QGTL_MAX_PHASES 4

typedef struct { float re, im; } qgtl_c64_t;

enum {
    QGTL_STORE_INPLACE = 0,   /* write the tile back where it was read */
    QGTL_STORE_RING = 1,      /* tile = bits 0..m-1; apply CNOT(i->i+1 mod n), i ascending, while storing */
    QGTL_STORE_PERM = 2       /* whole state in one tile; apply perm_* (CNOT/SWAP list) while storing */
};

enum {
    QGTL_LOAD_NATURAL = 0,    /* read the tile from where its amplitudes are */
    QGTL_LOAD_RING = 1        /* tile = bits 0..m-1 of the output; read it from the two half tiles the ring
                                 store would have written it to (the inverse ring), register pass only,
                                 not in place, natural store */
};

/* One LDS-tiled sweep pass: a set of commuting 2x2 matrices on tile bits,
 * followed by an optional index permutation on store. */
typedef struct {
    int32_t n;                                 /* qubits of the state */
    int32_t m;                                 /* tile bits */
    int32_t nlow;                              /* tile bits 0..nlow-1 are physical bits 0..nlow-1 */
    int32_t nrest;                             /* non-tile bits (enumerated by the tile index) */
    int32_t store;                             /* QGTL_STORE_* */
    int32_t ngates;                            /* matrices in this pass (<= m) */
    int32_t nperm;                             /* QGTL_STORE_PERM: gates in perm_* */
    int32_t load;                              /* QGTL_LOAD_* */
    /* Index arrays are 32-bit so that kernels read them with scalar loads. */
    int32_t hi_pos[QGTL_TILE_BITS_MAX];        /* physical bit of tile bit nlow+i */
    int32_t rest_pos[QGTL_MAX_QUBITS];         /* physical bit of tile-index bit i */
    int32_t gate_bit[QGTL_TILE_BITS_MAX];      /* tile bit of matrix g */
    int32_t perm_type[QGTL_PERM_MAX];          /* 0 = CNOT(a -> b), 1 = SWAP(a, b) */
    int32_t perm_a[QGTL_PERM_MAX];
    int32_t perm_b[QGTL_PERM_MAX];
    qgtl_c64_t mat[QGTL_TILE_BITS_MAX][4];     /* row-major 2x2 */
    /* Register-phase plan (full 2^13 tiles): 32 amplitudes per thread; in
     * phase p the tile bits ph_rbits[p][0..4] index a thread's registers and
     * the gates ph_first[p] .. ph_first[p+1]-1 (in mat order) act on register
     * bit ph_greg[g].  Phase 0 is the coalesced layout {0, 7, 8, 9, 10}.
     * nphase = 0: no plan (the LDS-only kernel runs the pass). */
    int32_t nphase;
    int32_t ph_rbits[QGTL_MAX_PHASES][5];
    /* thread bit i of phase p's layout is tile bit ph_tbits[p][i]; the first four
     * (one quarter-wave) are chosen so that the transposes through the
     * XOR-swizzled LDS tile hit 16 distinct 8-byte banks */
    int32_t ph_tbits[QGTL_MAX_PHASES][8];
    int32_t ph_first[QGTL_MAX_PHASES + 1];
    int32_t ph_greg[QGTL_TILE_BITS_MAX];
    /* the same plan by register bit: in phase p register bit j carries a gate
     * iff bit j of ph_mask[p] is set, with rotation ph_rot[p][j] (form 1) or
     * matrix ph_mat[p][j] (form 0) */
    int32_t ph_mask[QGTL_MAX_PHASES];
    qgtl_c64_t ph_rot[QGTL_MAX_PHASES][5];
    qgtl_c64_t ph_mat[QGTL_MAX_PHASES][5][4];
    /* Rotation form (form = 1): matrix g = diag(post[g]) R_g diag(pre[g]) with the
     * real rotation R_g = [[c, -s], [s, c]], c = rot[g].re, s = rot[g].im;
     * pre/post[g][b] is the unit phase on tile-bit value b.  A pass then costs
     * 4 real FMA per amplitude per gate plus one complex product per phase
     * layer (the product over the pass's gates, from two tables in LDS).
     * has_pre = 0: every pre phase is 1.  form = 0: mat holds general matrices. */
    int32_t form;
    int32_t has_pre;
    qgtl_c64_t rot[QGTL_TILE_BITS_MAX];
    qgtl_c64_t pre[QGTL_TILE_BITS_MAX][2];
    qgtl_c64_t post[QGTL_TILE_BITS_MAX][2];
    /* QGTL_LOAD_RING: thread bits of the layout the gathered tile is read into
     * from LDS (register bits ph_rbits[0]; lanes 0..3 on independent banks
     * through the ring map) */
    int32_t gather_tbits[8];
    /* Rotation form in the register-phase kernel: every gate on register bit
     * j of phase p is K(t) = [[1, -t], [t, 1]], (a, b) -> (a - t b, b + t a),
     * t = ph_tan[p][j].  A rotation with |c| >= |s| is c K(s / c).  One with
     * |c| < |s| (its tile bit set in xmask) is X Z s K(-c / s) = X s K(c / s) Z,
     * X = [[0, 1], [1, 0]], Z = diag(1, -1): the swap X of its amplitude pair
     * is left to the store (element j of the tile is stored at tile index
     * j ^ xmask, an XOR on the LDS address of the last transpose), and the
     * sign Z goes into a phase layer:
     *   xfold = 0 (a post layer is applied): t = -c / s; the post layer reads
     *     its tables at j ^ xmask and multiplies by (-1)^popcount(j & xmask)
     *     (X diag(p0, p1) = diag(p1, p0) X);
     *   xfold = 1 (only a pre layer, has_pre = 1, has_post = 0): t = c / s;
     *     the pre layer multiplies by (-1)^popcount(j & xmask).
     * The product of the factors taken out (c or s of every gate) is scale.
     * A phase layer is (the tables at the layout's thread bits) x (its
     * register part): reg_post[r] / reg_pre[r] is the product, over the
     * register bits set in r, of the layer's phase ratio on that tile bit
     * (last / first layout): phase[1] / phase[0] of the gate, for a bit of
     * xmask conj(phase[1] / phase[0]) with the sign of the fold.  has_post =
     * 0: every post phase is 1. */
    float ph_tan[QGTL_MAX_PHASES][5];
    uint32_t xmask;
    int32_t xfold;
    int32_t has_post;
    float scale;
    qgtl_c64_t reg_post[32];
    qgtl_c64_t reg_pre[32];
    /* Phase layers moved between the two register passes of a block (the
     * register kernel only; the LDS kernel applies every pass's own tables).
     * nolayer = 1: the pass applies no phase layer and no scale; its layer
     * (phases, the factors taken out, the signs of its swapped gates) is a
     * product over its gate bits, which are tile-index bits of its partner
     * pass, and the partner applies it as one complex factor per tile.
     * tf_n > 0: the pass multiplies every amplitude of tile r by the product
     * over i < tf_n of tf[i][bit tf_bit[i] of r], in the layer that carries
     * its scale; the partner is the previous pass of the block (its post
     * layer, on the input bits of this pass's tile index) or the next pass
     * of the inverse block (its pre layer, on the output bits of this pass's
     * tile index, whose store is natural). */
    int32_t nolayer;
    int32_t tf_n;
    int32_t tf_bit[QGTL_TILE_BITS_MAX];
    qgtl_c64_t tf[QGTL_TILE_BITS_MAX][2];
} qgtl_pass_desc_t;

/* Derivative-state injection in the frame after a block:
 * d_k[y] = A_k[s][s] psi[y] + A_k[s][1-s] psi[y ^ v_k], s = parity(y & w_k). */
typedef struct {
    int32_t n;
    int32_t K;
    uint32_t v[QGTL_INJECT_MAX];
    uint32_t w[QGTL_INJECT_MAX];
    qgtl_c64_t A[QGTL_INJECT_MAX][4];
} qgtl_inject_desc_t;

/* One derivative state for a pass that forms it while loading (see
 * qgtl_hip_pass_c64_inject): element y is M[s][s] psi[y] + M[s][1-s]
 * psi[y ^ v], s = parity(y & w), M row-major. */
typedef struct {
    uint32_t v, w;
    qgtl_c64_t M[4];
} qgtl_inject_slot_t;

/* ---- grouped launches ----
 * One launch runs the same kind of work for many independent instances.
 * Work item i covers the units [unit_begin_i, unit_begin_{i+1}) of the
 * launch (the last item ends at the launch's unit count); a workgroup finds
 * its item by binary search over unit_begin and loops over units with a
 * grid stride, so the grid can be sized to the machine (persistent) rather
 * than to the work.  All addresses are device addresses. */
typedef struct {
    uint64_t desc;          /* qgtl_pass_desc_t */
    uint64_t src, dst;      /* state 0; state s at + s * stride */
    uint64_t stride;        /* complex64 elements between states */
    uint32_t ntiles;        /* tiles per state, 2^(n - m) */
    uint32_t nstates;
    uint32_t unit_begin;    /* units: ntiles * nstates, tile-major within a state */
    uint32_t variant;       /* 0: LDS gate groups; 1: register phases (m = 13, nphase >= 1, no PERM store);
                               2: register phases storing the split-plane form (a final pass) */
    uint64_t tabs;          /* the descriptor's phase tables (qgtl_hip_phase_tables_grouped), or 0:
                               built by every workgroup */
} qgtl_pass_item_t;

typedef struct {
    uint64_t desc;          /* qgtl_inject_desc_t */
    uint64_t psi, dst;      /* state k of the item written at dst + k * stride */
    uint64_t partials;      /* [K][tiles][4] FP64 sums; 0: no sums */
    uint64_t stride;
    uint32_t tiles;         /* ceil(2^n / 2048) */
    uint32_t K;
    uint32_t unit_begin;    /* units: tiles * K */
    uint32_t pad;           /* QGTL_INJECT_ITEM_SPLIT: write the states in split-plane form (no sums) */
} qgtl_inject_item_t;
#define QGTL_INJECT_ITEM_SPLIT 1u

typedef struct {
    uint64_t partials;      /* [K][tiles][4] */
    uint64_t out;           /* [K][4] */
    uint32_t tiles, K;
    uint32_t unit_begin;    /* units: K */
    uint32_t pad;
} qgtl_reduce_item_t;

typedef struct {
    uint64_t D;             /* P states of N complex64 */
    uint64_t partials;      /* [nchunks][T][128 * 128] fp32, T = nt (nt + 1) / 2 */
    uint64_t N;
    uint64_t chunk_len;     /* amplitudes per chunk, multiple of QGTL_GRAM_KSTEP */
    int32_t P, nt, nchunks; /* nchunks: multiple of 8 */
    uint32_t unit_begin;    /* multiple of 8; units: T * nchunks */
} qgtl_gram_item_t;

typedef struct {
    uint64_t partials;
    uint64_t C;             /* (nt * 128)^2 FP64 */
    int32_t nt, nchunks;
    uint32_t unit_begin;    /* units: T * 64 (256 elements each) */
    uint32_t pad;
} qgtl_gram_reduce_item_t;

typedef struct {
    uint64_t C, inj, Q;     /* Q: P x P complex FP64 in D-slot order; inj = 0: C holds P + 1 states, the
                               last the primal, whose row gives the overlaps and norms */
    int32_t R, P;
    uint32_t unit_begin;    /* units: ceil(P * P / 256) */
    uint32_t pad;
} qgtl_finalize_item_t;

typedef struct {
    uint64_t psi;           /* set to |0...0> */
    uint32_t N;
    uint32_t unit_begin;    /* units: ceil(N / 2048) */
} qgtl_fill_item_t;

/* grid <= 0: one workgroup per unit (capped at 2^31 - 1) */
int qgtl_hip_fill_grouped(const qgtl_fill_item_t *items, int nitems, uint32_t units, int grid);
/* variant: every item of the launch has this qgtl_pass_item_t.variant */
int qgtl_hip_pass_grouped(const qgtl_pass_item_t *items, int nitems, uint32_t units, int grid, int variant);
/* The phase tables of ndesc pass descriptors (rotation form; others are left
 * alone): descriptor i's post then pre tables, QGTL_PHASE_TAB complex64 each,
 * at tabs + 2 * QGTL_PHASE_TAB * i (one workgroup per descriptor). */
int qgtl_hip_phase_tables_grouped(const qgtl_pass_desc_t *descs, int ndesc, void *tabs);
int qgtl_hip_inject_grouped(const qgtl_inject_item_t *items, int nitems, uint32_t units, int grid);
int qgtl_hip_reduce_grouped(const qgtl_reduce_item_t *items, int nitems, uint32_t units, int grid);
/* grid: multiple of 8 (XCD affinity of a chunk's tiles) */
int qgtl_hip_gram_grouped(const qgtl_gram_item_t *items, int nitems, uint32_t units, int grid, int terms);
int qgtl_hip_gram_reduce_grouped(const qgtl_gram_reduce_item_t *items, int nitems, uint32_t units, int grid);
int qgtl_hip_finalize_grouped(const qgtl_finalize_item_t *items, int nitems, uint32_t units, int grid);
/* The complex-form Gram of split-plane states, grouped, for instances of 129..256 Gram
 * states (P = the item's state count, nt = 2): one unit per split-K chunk, the
 * three tile pairs' fp32 partials per chunk; its reduction into G (complex128,
 * nt 128 x nt 128; units per item: T 256, T = 3) and the FP64 correction with
 * the primal as the last state (units per item: ceil(P^2 / 512), P = parameters). */
int qgtl_hip_gram_cplx_grouped(const qgtl_gram_item_t *items, int nitems, uint32_t units, int grid, int products);
int qgtl_hip_gram_cplx_reduce_grouped(const qgtl_gram_reduce_item_t *items, int nitems, uint32_t units, int grid);
int qgtl_hip_finalize_cplx_grouped(const qgtl_finalize_item_t *items, int nitems, uint32_t units, int grid);

/* ---- graphs: capture the library stream's launches and replay them ---- */
typedef struct qgtl_hip_graph_s *qgtl_hip_graph_t;
int qgtl_hip_capture_begin(void);
int qgtl_hip_capture_end(qgtl_hip_graph_t *graph);       /* instantiates */
int qgtl_hip_capture_end_update(qgtl_hip_graph_t *graph);   /* update *graph in place, or (re)instantiate */
void qgtl_hip_capture_abort(void);
int qgtl_hip_graph_launch(qgtl_hip_graph_t graph);
void qgtl_hip_graph_destroy(qgtl_hip_graph_t graph);
int qgtl_hip_cu_count(void);

/* ---- runtime ---- */
int qgtl_hip_init(int device);
int qgtl_hip_device_count(void);
void *qgtl_hip_stream(void);                          /* the library's stream (default priority) */
int qgtl_hip_malloc(void **ptr, size_t bytes);
int qgtl_hip_free(void *ptr);
int qgtl_hip_mem_free(size_t *free_bytes);            /* free device memory on the current device */
int qgtl_hip_host_alloc(void **ptr, size_t bytes);     /* pinned host memory (hipHostMalloc) */
int qgtl_hip_host_free(void *ptr);
unsigned long long qgtl_hip_alloc_calls(void);        /* device/pinned allocations and frees made so far */
int qgtl_hip_h2d(void *dst, const void *src, size_t bytes);   /* async on the library stream */
int qgtl_hip_d2h(void *dst, const void *src, size_t bytes);   /* async on the library stream */
int qgtl_hip_d2d(void *dst, const void *src, size_t bytes);
int qgtl_hip_sync(void);
const char *qgtl_hip_last_error(void);

typedef struct qgtl_hip_event_s *qgtl_hip_event_t;
int qgtl_hip_event_create(qgtl_hip_event_t *ev);
int qgtl_hip_event_record(qgtl_hip_event_t ev);
int qgtl_hip_event_elapsed_ms(qgtl_hip_event_t start, qgtl_hip_event_t stop, float *ms);
int qgtl_hip_event_destroy(qgtl_hip_event_t ev);

/* ---- a second stream ----
 * qgtl_hip_set_lane(1) routes the kernel launches that follow to a second
 * stream (default priority) until qgtl_hip_set_lane(0); copies, syncs and
 * timing events stay on the library stream.  Order between the lanes goes
 * through events made by qgtl_hip_event_create_sync: _lane_signal records one
 * on the current lane, _lane_wait makes the current lane wait for it. */
int qgtl_hip_set_lane(int lane);
int qgtl_hip_event_create_sync(qgtl_hip_event_t *ev);
int qgtl_hip_lane_signal(qgtl_hip_event_t ev);
int qgtl_hip_lane_wait(qgtl_hip_event_t ev);

/* ---- complex64 sweeps (mixed mode) ---- */
int qgtl_hip_fill_zero_state_c64(void *psi, int n);                    /* |0...0> */
/* Split-plane form of a final derivative state (the Gram's operand form): the
 * 8 N bytes of state s hold, for each block of 64 amplitudes (512 bytes, the
 * footprint of the same block in complex64), 64 hi words then 64 lo words, one
 * per amplitude, each an fp16 pair (Re piece, Im piece), where for each
 * component x: hi = 2^15 x rounded to 11 significant bits (nearest, ties away)
 * and lo = fp16(2^15 x - hi).  Amplitude y: hi word 128 (y / 64) + y % 64, lo
 * word 64 further (32-bit word indices). */
/* nstates states at src + s*stride -> dst + s*stride (elements of complex64);
 * split_out = 1: dst receives the split-plane form (register-phase passes only,
 * see qgtl_hip_pass_split_ok); tabs: NULL, or the pass's rotation-form phase
 * tables in device memory (QGTL_PHASE_TAB post entries, then as many pre
 * entries, complex64; entry i < 128: product over the gates on tile bits 0..6
 * of their phase at bit pattern i; entry 128 + i: the same for tile bits 7..12),
 * which the workgroups then copy instead of building */
#define QGTL_PHASE_TAB 192
int qgtl_hip_pass_c64(const qgtl_pass_desc_t *desc, const void *src, void *dst, size_t stride,
                      int nstates, int tile_variant, int split_out, const void *tabs);
int qgtl_hip_pass_split_ok(const qgtl_pass_desc_t *desc);
/* split_out = 1: the derivative states are written in split-plane form; partials = NULL:
 * no FP64 sums (the overlaps are taken from the Gram instead) */
int qgtl_hip_inject_c64(const qgtl_inject_desc_t *desc, const void *psi, void *dst, size_t stride,
                        double *partials, int *tiles_out, int split_out);
/* the FP64 sums of qgtl_hip_inject_c64 only (the states are not written) */
int qgtl_hip_inject_sums_c64(const qgtl_inject_desc_t *desc, const void *psi, double *partials, int *tiles_out);
/* qgtl_hip_pass_c64 on nstates derivative states formed while loading from the
 * one state psi: state s uses slots[s] (device array); register passes only */
int qgtl_hip_pass_c64_inject(const qgtl_pass_desc_t *desc, const void *psi, const void *slots, void *dst,
                             size_t stride, int nstates, int split_out, const void *tabs);
/* out[k*4 + c] = sum over t of partials[(k*tiles + t)*4 + c], fixed order, FP64 */
int qgtl_hip_reduce_partials_f64(const double *partials, int K, int tiles, double *out);
int qgtl_hip_cnot_c64(void *base, size_t stride, int nstates, int n, int c, int t);
int qgtl_hip_swap_c64(void *base, size_t stride, int nstates, int n, int a, int b);
int qgtl_hip_c64_to_c128(const void *src, void *dst, size_t count);
int qgtl_hip_fill_f32(void *dst, size_t count, float value);
int qgtl_hip_copy_f32x4(void *dst, const void *src, size_t count16);    /* count16 units of 16 bytes */
/* out = H psi for a Pauli sum (terms in device memory) */
int qgtl_hip_pauli_apply_c64(const void *psi, void *out, int n, int nterms, const void *terms_dev);
/* per (state, tile) FP64 sums: <D_k|v>, |D_k|^2, |v|^2 (same layout as inject partials);
 * planes = 1: D holds split-plane states, read as (hi + lo) 2^-15 */
int qgtl_hip_dot_states_c64(const void *D, size_t stride, int K, const void *v, int n, double *partials,
                            int *tiles_out, int planes);
/* out[0] = energy, out[1 + k] = gradient (from reduced dot sums) */
int qgtl_hip_energy_gradient_finalize(const double *dk_h, const double *psi_h, int P, double *out);

/* ---- Gram (mixed): C = Z Z^T with Z rows (Re d_s, Im d_s), split products on MFMA ---- */
int qgtl_hip_gram_tiles(int P);                                          /* upper 128x128 tiles */
/* terms = 3 / 4: two bf16 pieces, 3 or 4 products; 6: exact three-piece bf16
 * split, six products; QGTL_GRAM_F16X3 / X4: two fp16 pieces of 2^15 x
 * (C is then scaled by 2^30), 3 or 4 products */
#define QGTL_GRAM_KSTEP 64   /* amplitudes per Gram LDS stage; chunk lengths are multiples of it */
#define QGTL_GRAM_F16X3 33
#define QGTL_GRAM_F16X4 34
/* Real-form Gram of complex64 states (split in the kernel): C = Z Z^T with Z
 * rows (Re d_s, Im d_s) */
int qgtl_hip_gram_bf16x3(const void *D, int P, size_t N, float *partials, int nchunks, int terms);
int qgtl_hip_gram_reduce(const float *partials, int P, int nchunks, double *C);   /* C: R x R tiles, FP64 */
/* Q = normalised G - a a^H from C and injection sums (P x 4 doubles). */
int qgtl_hip_qgt_finalize(const double *C, const double *inj, int P, double *Q);
/* the same with the overlaps taken from the Gram: C is the Gram of P + 1 states,
 * state P the primal state at the derivative states' meeting point */
int qgtl_hip_qgt_finalize_psi(const double *C, int P, double *Q);
/* one complex64 state of 2^n amplitudes (n >= 6) in split-plane form */
int qgtl_hip_c64_to_split(const void *src, void *dst, int n);

/* ---- complex-form Gram of split-plane states ----
 * G = D^H D of P states in split-plane form (N % 64 == 0), tiles of 128 states
 * (nt = ceil(P / 128), T = nt (nt + 1) / 2 upper tile pairs), split-K over
 * nchunks (a multiple of 8) chunks of amplitudes: fp32 partials of Re G and Im G
 * per chunk and tile pair (qgtl_hip_gram_cplx_partial_floats(P) floats per
 * chunk), FP64 sum over the chunks into G (complex128, R x R, R = 128 nt; only
 * entries (k, l) with k <= l are written), fp16 pieces with
 * products = 3 (hh + hl + lh) or 4 (+ ll); G is scaled by 2^30. */
size_t qgtl_hip_gram_cplx_partial_floats(int P);
/* stage_amps: amplitudes per LDS stage, 16 (four buffers, three stages in flight) or 32 (two) */
int qgtl_hip_gram_cplx(const void *D, int P, size_t N, float *partials, int nchunks, int products, int stage_amps);
int qgtl_hip_gram_cplx_reduce(const float *partials, int P, int nchunks, double *G);
/* Q = normalised G - a a^H from the complex Gram of P + 1 states (state P the
 * primal state at the derivative states' meeting point, so a_k = G_kP) */
int qgtl_hip_qgt_finalize_cplx(const double *G, int P, double *Q);
/* MFMA lane-layout probe (diagnostics): 32 x 32 floats to out_dev */
int qgtl_hip_debug_mfma_layout(float *out_dev);
int qgtl_hip_debug_mfma16_layout(float *out_dev);   /* v_mfma_f32_16x16x32_f16 probe, 16 x 16 floats */

/* ---- complex128 plain path ---- */
int qgtl_hip_fill_zero_state_c128(void *psi, int n);
int qgtl_hip_gate1q_c128(void *psi, int n, int q, const double *m8);
int qgtl_hip_cgate_c128(void *psi, int n, int c, int t, const double *m8);
int qgtl_hip_swap_c128(void *psi, int n, int a, int b);
/* G (P x P complex, upper triangle) = D^H D; partials sized by gram_f64_partials(P, N) doubles */
size_t qgtl_hip_gram_f64_partials(int P, size_t N);
int qgtl_hip_gram_f64(const void *D, int P, size_t N, double *partials, double *G);
/* a_k = <D_k|psi>, P complex */
int qgtl_hip_dotc_c128(const void *D, int P, size_t N, const void *psi, double *partials, double *a);
int qgtl_hip_qgt_finalize_plain(const double *G, const double *a, int P, double *Q);

/* ---- learned preconditioner: dense layers (bf16 MFMA, fp32 accumulate) ---- */
/* C (M x N fp32) = op(A) op(B) [+ bias] [+ C]; op(A)[m][k] = ta ? A[k][m] : A[m][k] (lda
 * columns), op(B)[k][n] = tb ? B[n][k] : B[k][n] (ldb columns); *_bf16: operand stored as bf16 */
int qgtl_hip_mlp_gemm(int M, int N, int K, const void *A, int a_bf16, int ta, int lda, const void *B, int b_bf16,
                      int tb, int ldb, float *C, int ldc, const float *bias, int accumulate);
int qgtl_hip_gelu_fwd(const float *z, float *a, size_t n);
int qgtl_hip_gelu_bwd(const float *z, float *da, size_t n);           /* da *= g'(z) */
int qgtl_hip_colsum(const float *X, int rows, int cols, float *out);
int qgtl_hip_gather_rows(const float *src, const int *idx, int rows, int cols, float *dst);
int qgtl_hip_adam(float *w, uint16_t *w_bf16, float *m, float *v, const float *g, size_t n, float lr, float b1,
                  float b2, float eps, int step);
int qgtl_hip_to_bf16(const float *w, uint16_t *out, size_t n);
/* per sample b: loss[b] (FP64) and dY = scale * dL/dy for As_all[idx[b]] (P x P fp32) */
int qgtl_hip_kaporin(int P, int r, int B, const float *Y, const float *As_all, const int *idx, double *loss,
                     float *dY, float scale);

/* ---- FP64 natural-gradient solves, one workgroup per system ---- */
int qgtl_hip_lu_solve_c128(void *A, const void *b, void *x, int *piv, int P, int batch, int *status);
int qgtl_hip_herm_eig_c128(void *A, void *V, double *ev, int P, int batch, int max_sweeps, int *sweeps);
int qgtl_hip_cholesky_solve_f64(double *A, const double *b, double *x, int P, int batch, double lambda, int *status);
int qgtl_hip_pcg_f64(const double *A, const double *b, double *x, int P, int batch, double lambda, const double *d,
                     const double *U, int r, double tol, int maxit, int *iters, double *relres);

/* ---- fixed working memory of the no-allocation solve ----
 * One file-scope device array of QGTL_NOALLOC_SCRATCH_BYTES per GPU (held
 * until the process ends).  _init looks its device address up on the current
 * device (called by qgtl_hip_init); _scratch returns that address and size;
 * _jacobi_diag_f64 writes d = 1 / (diag(A_s) + lambda) for every system. */
#define QGTL_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int qgtl_hip_noalloc_init(void);
int qgtl_hip_noalloc_scratch(void **ptr, size_t *bytes);
int qgtl_hip_jacobi_diag_f64(const double *A, int P, int batch, double lambda, double *d);
int qgtl_hip_gram_small_c128(const void *M, void *B, int P);
/* device-resident natural-gradient step: A = Re Q (slot -> parameter order) + lambda I, D = diag A,
 * g = gradient (eg = [E, grad in slot order]); learned-preconditioner features and output transform */
int qgtl_hip_ng_assemble(const double *Q, const int *slot, int P, double lambda, const double *eg, double *A,
                         double *D, double *g);
int qgtl_hip_pc_features(const double *theta, const double *D, double field, int P, float *x);
int qgtl_hip_pc_transform(const float *y, const double *D, int P, int r, double *d, double *U);
int qgtl_hip_reciprocal(const double *D, int P, double *d);
int qgtl_hip_pinv_apply_c128(const void *V, const double *ev, const void *g, void *x, int P, double cutoff);

/* ---- whole-vector operations on complex64 vectors of any length (state container, compute backend) ----
 * partials: device scratch of qgtl_hip_vec_partials_len() doubles; out4: device, 4 doubles */
int qgtl_hip_vec_partials_len(void);
/* out4 = Re <a|b>, Im <a|b>, |a|^2, |b|^2 (FP64, fixed summation order) */
int qgtl_hip_vec_dotc_c64(const void *a, const void *b, size_t n, double *partials, double *out4);
/* out4[0] = sum |psi_i|^2 obs_i, out4[1] = |psi|^2 (obs: n real values) */
int qgtl_hip_vec_expect_diag_c64(const void *psi, const float *obs, size_t n, double *partials, double *out4);
int qgtl_hip_vec_scale_c64(void *x, size_t n, float s);
/* y = U x, U row-major m x ncols complex64, FP64 accumulation (y must not alias x) */
int qgtl_hip_vec_cgemv_c64(const void *U, const void *x, void *y, size_t m, size_t ncols);
/* C = A B, row-major complex64 (A m x n, B n x k), FP32 accumulation */
int qgtl_hip_vec_cgemm_c64(const void *A, const void *B, void *C, size_t m, size_t n, size_t k);
/* the kind of memory behind a pointer: 1 device memory of this process, 0 otherwise */
int qgtl_hip_is_device_ptr(const void *p);
int qgtl_hip_memset(void *dst, int value, size_t bytes);
/* a copy between any two of host, pinned and device memory (direction inferred), on the library stream */
int qgtl_hip_copy(void *dst, const void *src, size_t bytes);
int qgtl_hip_device_total_mem(size_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
