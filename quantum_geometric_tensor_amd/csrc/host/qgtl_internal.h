/*
 * qgtl_internal.h -- host-side structures of libqgtl_mi355x (C11).
 *
 * A circuit is compiled into blocks.  A block is a U-layer (one 2x2 matrix
 * per qubit, the product of that qubit's single-qubit gates, FP64) followed
 * by an M-stage (a list of CNOT/SWAP gates, a GF(2)-linear permutation of
 * basis indices).  Derivative states of the block's parameters are created
 * after the block from the block's output state, and carried through the
 * later blocks by LDS-tiled sweep passes.
 */
#ifndef QGTL_INTERNAL_H
#define QGTL_INTERNAL_H

#include <stdint.h>
#include <stdio.h>
#include <stddef.h>

#include "qgtl_hip_abi.h"
#include "qgtl_mi355x.h"
#include "qgtl_types.h"

typedef struct { double re, im; } qgtl_cd;
typedef struct { qgtl_cd m[4]; } qgtl_mat2;       /* row-major 2x2 */

#define QGTL_MAX_PASSES_PER_BLOCK 6
/* Default Gram split: two fp16 pieces of 2^15 x, three products.  Its
 * representation error is <= 2^-22 per value whatever the state, whereas
 * two bf16 pieces leave 2^-17, which only averages out for unstructured
 * states (shallow circuits at n >= 14 measured 1.8e-6 normwise). */
static inline int qgtl_default_gram_terms(int n) { (void)n; return QGTL_GRAM_F16X3; }
static inline int qgtl_valid_gram_terms(int t)
{
    return t == 3 || t == 4 || t == 6 || t == QGTL_GRAM_F16X3 || t == QGTL_GRAM_F16X4;
}

typedef struct {
    int has_u[QGTL_MAX_QUBITS];
    qgtl_mat2 u[QGTL_MAX_QUBITS];
    int nperm;
    int8_t perm_type[QGTL_PERM_MAX];     /* 0 CNOT(a -> b), 1 SWAP(a, b) */
    int8_t perm_a[QGTL_PERM_MAX];
    int8_t perm_b[QGTL_PERM_MAX];
    int is_ring;                         /* M-stage is CNOT(i -> i+1 mod n), i = 0..n-1 */
    int param_begin, param_end;          /* parameters created in this block */
    /* device passes that apply this block */
    int npasses;
    qgtl_pass_desc_t pass[QGTL_MAX_PASSES_PER_BLOCK];
    int pass_flips[QGTL_MAX_PASSES_PER_BLOCK];  /* 1: output goes to the other buffer */
    int mstage_kernels;                  /* 1: M-stage applied by per-gate kernels after the passes */
    /* the inverse block (the inverse M-stage, then U^dagger), for derivative
     * states carried backwards (n > 13, M-stage a CNOT ring or empty) */
    int inv_ok;
    int ninv;
    qgtl_pass_desc_t ipass[QGTL_MAX_PASSES_PER_BLOCK];
    int ipass_flips[QGTL_MAX_PASSES_PER_BLOCK];
    int tab0;                            /* phase tables of pass i at tab0 + i, of ipass i at tab0 + npasses + i */
} qgtl_block_t;

typedef struct {
    int block;
    int qubit;
    int op_index;
    qgtl_mat2 A;                         /* d_k (pre-M) = A psi_block (pre-M) */
    uint32_t v, w;                       /* injection masks after the M-stage */
    qgtl_mat2 B;                         /* u_q^dagger A u_q: the block's inverse applied to d_k, a 2 x 2 on
                                            the parameter's qubit acting on the block's input state */
} qgtl_param_t;

struct qgtl_plan {
    int n;
    int mode;
    int nops;
    int P;
    qgtl_op_t *ops;
    int nblocks;
    qgtl_block_t *blocks;
    int blocks_cap;
    qgtl_param_t *params;                /* indexed by parameter number */
    int *slot_param;                     /* D slot -> parameter number */
    qgtl_tune_t tune;
    /* device buffers */
    size_t N;
    void *psi[2];                        /* primal (c64 or c128) */
    void *D;                             /* P states */
    void *tmp;                           /* chunk ping-pong (c64) */
    size_t tmp_states;
    double *inj_partials;
    size_t inj_partials_len;
    double *inj_sums;                    /* P x 4 */
    float *gram_partials;
    int gram_chunks;           /* split-K chunks of the real-form Gram (complex64 states) */
    int gram_chunks_cplx;      /* ... of the complex-form Gram (split-plane states) */
    int gram_cplx;             /* the last execute ran the complex-form Gram: C holds G (R x R complex128) */
    double *C;                           /* mixed: R x R;  plain: G (P x P complex) */
    double *Qdev;                        /* P x P complex */
    double *Qhost;                       /* pinned host copy of Qdev (download staging) */
    double *plain_partials;
    size_t plain_partials_len;
    uint64_t structure_hash;
    void *psi_final;                     /* primal after the last execute (mixed) */
    qgtl_hip_event_t ev_psi, ev_inj;     /* ordering of the primal chain on the second stream */
    /* depth-first plans whose block-boundary states fit a small budget: the
     * primal state after block k (k = 0..nblocks) in chain[k], the chain's
     * pass scratch, and one event per chain[k], so that the chain runs ahead
     * on the second stream and each chunk waits only for the state it reads */
    void **chain;
    int nchain;                          /* nblocks + 1, or 0 (two-buffer chain) */
    void *chain_tmp;
    void *tmp_side;                      /* chunk buffer of the chain's stream (backward chunks), or NULL */
    int capturing;                       /* the launches being issued are captured into the plan's graph */
    qgtl_hip_event_t *ev_chain;
    int nev_chain;
    int presplit;                        /* D holds split-plane states (the Gram's fp16 operand pieces) */
    int psi_in_gram;                     /* D slot P holds the primal state at the meeting point; the Gram of
                                            P + 1 states gives the overlaps and norms (depth-first plans) */
    int meet;                            /* D holds the derivative states after block meet (two-sided schedule:
                                            later blocks' states are carried backwards); nblocks - 1: the end */
    /* rotation-form phase tables of every pass (block order), built on the host */
    int tabs_dirty;                      /* angles changed since the last upload */
    int tabs_passes;
    float *tabs_host;                    /* tabs_passes x 2 x QGTL_PHASE_TAB complex64 */
    void *tabs_dev;
    /* per D slot: the injection as a load-time formula, forward (slots 0..P-1)
     * and backward (P..2P-1), rebuilt with the tables */
    qgtl_inject_slot_t *inj_slots_host;
    void *inj_slots_dev;
    /* Hamiltonian for energy / gradient */
    int nterms;
    void *terms_dev;
    void *hpsi;
    double *eg_partials;
    double *eg_sums;                     /* (P + 1) x 4 */
    double *eg_out;                      /* 1 + P */
    /* device-resident natural-gradient step (qgtl_ngvqe.c) */
    int *ng_slot;                        /* parameter -> D slot */
    double *ng_A, *ng_D, *ng_g, *ng_x, *ng_d, *ng_U, *ng_theta, *ng_res;
    int *ng_it;
    int ng_rank;
    /* the launch sequence as a HIP graph: build_gen counts plan builds (angles or
     * tune changed); the graph is captured on the second execute of one build and
     * replayed while the build is unchanged */
    uint64_t build_gen, exec_gen, graph_gen;
    qgtl_hip_graph_t graph;
    qgtl_stats_t graph_st;               /* counts of the captured sequence */
};

/* plan construction (qgtl_plan.c) */
int qgtl_plan_build(struct qgtl_plan *plan);
int qgtl_gate_matrix(const qgtl_op_t *op, qgtl_mat2 *m);
int qgtl_is_param_gate(int type);
uint64_t qgtl_structure_hash(int n, int nops, const qgtl_op_t *ops, int mode);

/* errors (qgtl_engine.c) */
int qgtl_set_error(int code, const char *fmt, ...);

/* One process-wide recursive lock (qgtl_engine.c) held by every exported
 * entry point that touches the device or state shared between calls: the
 * library's streams and lane selector, the plan cache behind qgtl_qgt and the
 * QGTL network calls, the solvers' scratch buffers and the profiling events.
 * Calls from several host threads therefore run one after another on the GPU,
 * each with its own inputs and outputs.  QGTL_LOCKED() at the top of a
 * function holds the lock until the function returns. */
void qgtl_lock(void);
void qgtl_unlock(void);
static inline void qgtl_unlock_scope(int *unused) { (void)unused; qgtl_unlock(); }
#define QGTL_LOCKED() \
    int qgtl_lock_guard_ __attribute__((cleanup(qgtl_unlock_scope), unused)) = (qgtl_lock(), 0)

/* qgtl_plan_set_angles without taking the lock (host-only plan rebuild): for a
 * caller that holds the lock and rebuilds several plans on worker threads, one
 * thread per plan -- a worker taking the lock the caller holds would wait for
 * the caller, which waits for the workers. */
int qgtl_plan_set_angles_nolock(qgtl_plan_t *p, int nops, const double *angles);

/* The block after which the derivative states of the plan meet on the two-sided
 * schedule (the last block when every state is carried forward). */
int qgtl_plan_meet_block(const struct qgtl_plan *p);



/* engine (qgtl_engine.c) */
int qgtl_init_if_needed(void);

/* learned preconditioner (qgtl_mlp.c) */
void qgtl_mlp_shape(const qgtl_mlp_t *m, int *in_dim, int *P, int *rank);
void qgtl_pass_phase_tables(const qgtl_pass_desc_t *p, float *out);
int qgtl_carry_to_meet(struct qgtl_plan *p, void **state, void **scratch);

/* file container of qgtl_io.c (prefix, kind header, payload, FNV-1a 64 checksum) */
uint64_t qgtl_fnv1a(uint64_t h, const void *data, size_t n);
int qgtl_io_write(const char *path, const char magic[8], const void *hdr, uint32_t hdr_bytes, int count,
                  const void *const *pieces, const size_t *sizes);
int qgtl_io_open(const char *path, const char magic[8], void *hdr, uint32_t hdr_bytes, uint64_t *payload_bytes,
                 uint64_t *checksum, FILE **out);
int qgtl_io_read_payload(FILE *f, const char *path, const void *hdr, uint32_t hdr_bytes, uint64_t payload_bytes,
                         uint64_t checksum, int count, void *const *pieces, const size_t *sizes);
/* device input row of a batch of one, and the forward pass leaving y on the device */
int qgtl_mlp_device_input(qgtl_mlp_t *m, float **x_dev);
int qgtl_mlp_forward_device(qgtl_mlp_t *m, const float **y_dev);

/* Hamiltonian (qgtl_hamiltonian.c): E and grad into p->eg_out on the device, no sync */
int qgtl_energy_gradient_device(qgtl_plan_t *p);
#endif
