/*
 * qgtl_mi355x.h -- additive entry points of libqgtl_mi355x.so.
 *
 * QGTL has no counterpart for these symbols: they expose FP64 output, plan
 * reuse, batching, timing and the natural-gradient solvers that QGTL's
 * ComplexFloat API cannot carry.  The drop-in API in qgtl_api.h is built on
 * top of them.  All functions return 0 on success and a negative qgt_error_t
 * code on failure; qgtl_last_error() then holds a message for the calling
 * thread.
 *
 * Conventions: complex arrays are interleaved (re, im) doubles; qubit q is
 * bit (1 << q) of a basis index; parameters are numbered in gate order
 * (src/quantum_geometric/core/quantum_parameter_shift.c:278-337).
 */
#ifndef QGTL_MI355X_H
#define QGTL_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One gate of a flat circuit.  type is a gate_type_t value.  Single-qubit
 * gates: q0 = target, q1 = -1.  CNOT/CY/CZ: q0 = control, q1 = target
 * (create_cnot_gate's {control, target}, quantum_gate_operations.c:420-423).
 * SWAP: q0, q1.  param: variational parameter index carried by an RX/RY/RZ
 * gate, or -1. */
typedef struct {
    int32_t type;
    int32_t q0;
    int32_t q1;
    int32_t param;
    double angle;
} qgtl_op_t;

enum {
    QGTL_MODE_MIXED = 0,      /* complex64 sweeps, split-BF16x3 MFMA Gram, FP64 reductions/correction */
    QGTL_MODE_FP64_PLAIN = 1  /* complex128, one kernel per gate, one derivative state at a time, FP64 Gram */
};

/* Per-call timing and traffic (device times come from HIP events recorded on
 * the library's stream around each launch when profiling is enabled). */
typedef struct {
    double wall_ms;                 /* host wall time of the call */
    double sweep_ms;                /* sum of sweep-pass kernel durations */
    double inject_ms;               /* derivative-state injection kernels */
    double gram_ms;                 /* Gram kernels (incl. split-K reduction) */
    double finalize_ms;             /* FP64 correction / normalisation */
    double other_ms;                /* everything else that was timed */
    int64_t sweep_launches;
    int64_t sweep_state_passes;     /* states x passes over the full state */
    double sweep_bytes;             /* algorithmic bytes moved by sweep passes (16 B per amp per state-pass) */
    int64_t gram_launches;
    double gram_flops;              /* 4 N P (P+1) */
    double gram_kernel_ms;          /* duration of the main Gram kernel */
    int64_t kernel_launches;        /* all launches of the call */
    int32_t P;
    int32_t n;
    int32_t presplit;               /* 1: final derivative states were written as the Gram's fp16 pieces */
    int32_t meet_block;             /* derivative states met after this block (the last one: all carried forward) */
    int32_t graph;                  /* 1: the launches ran as a replayed HIP graph (counts are the captured ones) */
    int32_t reserved;
} qgtl_stats_t;

/* Tunables chosen by the autotuner (or defaults when NULL / zero fields). */
typedef struct {
    int32_t chunk_states;     /* derivative states propagated together (0 = auto) */
    int32_t gram_chunks;      /* split-K factor of the Gram (0 = auto) */
    int32_t gram_variant;     /* Gram operand split: 3 / 4 = two bf16 pieces with 3 / 4 products, 6 = exact
                                 three-piece bf16 split, 33 / 34 = two fp16 pieces of 2^15 x with 3 / 4
                                 products; 0 = default (33) */
    int32_t sweep_form;       /* 0 = rotation form (real rotations + phase tables), 1 = general complex matrices */
    int32_t no_presplit;      /* 0: where the plan allows it, the final pass of every derivative state (and the
                                 last block's injection) writes the Gram's fp16 operand pieces directly, so the
                                 Gram does no conversion; 1: states stay complex64 and the Gram splits them */
    int32_t sweep_schedule;   /* 0: two-sided where the plan allows it (n > 13, CNOT-ring or empty M-stages):
                                 the derivative states of late blocks are carried backwards through the
                                 inverse blocks and all states meet after the block that minimises the
                                 sweep work; 1: every derivative state is carried forward to the end */
    int32_t no_fused_inject;  /* 0: a derivative-state chunk that is carried through passes is formed by its
                                 first pass while loading (from the primal state, read through the caches),
                                 never written or read as injected; 1: a separate injection kernel writes it */
    int32_t no_graph;         /* 0: from the second execute of an unchanged plan (same angles and tune) on,
                                 the launch sequence runs as one captured HIP graph (per-launch profiling
                                 keeps it eager); 1: every execute launches its kernels one by one */
} qgtl_tune_t;

int qgtl_init(int device);
const char *qgtl_last_error(void);
int qgtl_device_count(void);
/* Enable per-launch HIP event timing (fills the *_ms fields of qgtl_stats_t). */
void qgtl_set_profiling(int enable);

/* Hardware-efficient ansatz builder: per layer RY(q), RZ(q) on every qubit,
 * then CNOT(i -> i+1 mod n) for i = 0..n-1.  Returns the op count written
 * (ops needs qgtl_hea_num_ops(n, layers) entries). */
int qgtl_hea_num_ops(int n, int layers);
int qgtl_hea_ops(int n, int layers, const double *theta, qgtl_op_t *ops);
int qgtl_num_params(int nops, const qgtl_op_t *ops);

/* Full P x P QGT, Q[mu][nu] = <d_mu psi|d_nu psi> - <d_mu psi|psi><psi|d_nu psi>,
 * written as complex128 row-major.  Q may be NULL (timing runs). */
int qgtl_qgt(int n, int nops, const qgtl_op_t *ops, int mode, double *Q, qgtl_stats_t *stats);

/* Final state |psi> as complex128 (mixed mode computes in complex64). */
int qgtl_statevector(int n, int nops, const qgtl_op_t *ops, int mode, double *psi);

/* Reusable plan: structure fixed at creation, angles replaceable. */
typedef struct qgtl_plan qgtl_plan_t;
qgtl_plan_t *qgtl_plan_create(int n, int nops, const qgtl_op_t *ops, int mode, const qgtl_tune_t *tune);
int qgtl_plan_set_angles(qgtl_plan_t *plan, int nops, const double *angles);
int qgtl_plan_execute(qgtl_plan_t *plan, double *Q, qgtl_stats_t *stats);
int qgtl_plan_num_params(const qgtl_plan_t *plan);
void qgtl_plan_destroy(qgtl_plan_t *plan);
/* The overlaps a_k = <d_k psi|psi> (2P doubles: Re, Im; parameter order) that the
 * last mixed-mode execute subtracted in Q = G - a a^H (the quantity QGTL forms in
 * quantum_geometric_tensor_network.c:1139-1166), normalised by the exact
 * ||d_k psi|| = 1/2 and ||psi|| = 1 as the QGT is. */
int qgtl_plan_overlaps(qgtl_plan_t *plan, double *a);

/* Release every cached device buffer of the calling process. */
void qgtl_release_cache(void);

/* ---- batched QGTs (mixed mode) ----
 * Many independent circuits (uneven qubit counts and depths) run together:
 * each step of the sweep (block b, pass j), the injections, the Gram, its
 * split-K reduction and the FP64 correction is ONE grouped launch over every
 * instance, with a persistent grid that loops over the work units; the whole
 * sequence is captured once in a HIP graph and replayed per execution.
 * QGTL evaluates one circuit and one matrix element at a time
 * (src/quantum_geometric/core/quantum_geometric_curvature.c:289-320); the
 * batch has no counterpart there.
 *
 * ops: the instances' gate lists one after another (nops[i] each); every
 * instance needs n in 1..30, at least one parameter, and M-stages that are
 * CNOT rings when n > 13 (any CNOT/SWAP list when n <= 13).  Q output: the
 * instances' P_i x P_i complex128 matrices one after another (parameter
 * order), at complex offset qgtl_batch_q_offset(b, i). */
typedef struct qgtl_batch qgtl_batch_t;

typedef struct {
    int32_t pass_grid;        /* workgroups of the sweep launches (0: 2 per CU; -1: one per unit) */
    int32_t inject_grid;      /* injection launches (0: 8 per CU; -1: one per unit) */
    int32_t gram_grid;        /* Gram launches (0: 2 per CU; -1: one per unit); rounded to a multiple of 8 */
    int32_t gram_chunk_len;   /* amplitudes per split-K chunk of a Gram (0: 16384) */
    int32_t pass_variant;     /* 0: register phases where they apply; 1: LDS gate groups only */
    int32_t no_graph;         /* 1: issue the launches directly instead of replaying the graph */
    int32_t sweep_form;       /* 0: rotation form of the pass matrices; 1: general complex matrices */
    int32_t sweep_schedule;   /* 0: two-sided per instance where its blocks allow it (derivative states of
                                 late blocks carried backwards through inverse blocks to a meeting block);
                                 1: every state carried forward to the end (fixed at creation) */
} qgtl_batch_tune_t;

typedef struct {
    double wall_ms;           /* host wall time of qgtl_batch_execute */
    double device_ms;         /* HIP-event time of the launch sequence (graph replay) */
    double upload_ms;         /* angle-dependent descriptors to the GPU (host wall) */
    double download_ms;       /* Q to the host and reordering (host wall) */
    int64_t launches;         /* kernel launches in the sequence */
    int64_t units;            /* work units over all launches */
    double sweep_bytes;       /* algorithmic sweep bytes (16 B per amplitude per state-pass) */
    double gram_flops;        /* sum of 4 N P (P + 1) */
    int32_t instances;
    int32_t graph;            /* 1 when the graph was replayed */
} qgtl_batch_stats_t;

qgtl_batch_t *qgtl_batch_create(int count, const int *n, const int *nops, const qgtl_op_t *ops,
                                const qgtl_batch_tune_t *tune);
int qgtl_batch_count(const qgtl_batch_t *b);
int qgtl_batch_num_params(const qgtl_batch_t *b, int i);
size_t qgtl_batch_q_offset(const qgtl_batch_t *b, int i);   /* complex elements before instance i */
size_t qgtl_batch_q_size(const qgtl_batch_t *b);            /* sum of P_i^2 */
size_t qgtl_batch_device_bytes(const qgtl_batch_t *b);
/* angles: one per op, same layout as the creation ops (non-rotation entries ignored) */
int qgtl_batch_set_angles(qgtl_batch_t *b, const double *angles);
int qgtl_batch_execute(qgtl_batch_t *b, double *Q, qgtl_batch_stats_t *stats);
int qgtl_batch_set_tune(qgtl_batch_t *b, const qgtl_batch_tune_t *tune);
int qgtl_batch_get_tune(const qgtl_batch_t *b, qgtl_batch_tune_t *tune);
/* Autotuner: times candidate launch shapes on this batch (coordinate
 * descent, median of `reps` replays each), keeps the fastest in the batch
 * and, when json_path is given, writes it there as a JSON object. */
int qgtl_batch_autotune(qgtl_batch_t *b, int reps, const char *json_path, qgtl_batch_tune_t *best);
int qgtl_batch_tune_load(const char *json_path, qgtl_batch_tune_t *tune);
void qgtl_batch_destroy(qgtl_batch_t *b);

/* ---- Hamiltonians: sums of Pauli strings ----
 * term = coeff * i^{ny} X^xmask Z^zmask with ny = popcount(xmask & zmask)
 * (a Y on qubit q sets bit q in both masks). */
typedef struct {
    uint32_t xmask;
    uint32_t zmask;
    double coeff_re;
    double coeff_im;
} qgtl_pauli_term_t;
/* QGTL's Pauli-string convention: character 0 acts on the highest qubit
 * (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:339). */
int qgtl_pauli_term_from_string(const char *pauli, double coeff, qgtl_pauli_term_t *out);
/* Transverse-field Ising chain H = -J sum_{i<n-1} Z_i Z_{i+1} - h sum_i X_i (2n-1 terms). */
int qgtl_tfim_terms(int n, double J, double h, qgtl_pauli_term_t *terms);
int qgtl_plan_set_hamiltonian(qgtl_plan_t *plan, int nterms, const qgtl_pauli_term_t *terms);
/* After qgtl_plan_execute (mixed mode): E = <psi|H|psi> and
 * grad_k = 2 Re <d_k psi|H|psi> from the derivative states left on the GPU. */
int qgtl_plan_energy_gradient(qgtl_plan_t *plan, double *energy, double *grad);

/* FP64 copy (P x P complex128) of the QGT that the drop-in API caches on a
 * network created by create_quantum_geometric_tensor_network. */
struct quantum_geometric_tensor_network;
int qgtl_network_qgt_f64(const struct quantum_geometric_tensor_network *qgtn, double *Q, int *P);

/* The device copy (complex64, `dimension` values) of a quantum_state_t moved to
 * the GPU with quantum_state_to_device, or NULL: for callers that write or read
 * it with their own kernels or the compute backend's memcpy. */
struct quantum_state_t;
void *qgtl_state_device_ptr(const struct quantum_state_t *state);

/* ---- natural-gradient solves (FP64 on the GPU) ----
 * (A + lambda I) x = b for `batch` real symmetric P x P systems stored one
 * after another (row-major).  CHOLESKY is exact (the yardstick of the iterative
 * methods); CG has no preconditioner; JACOBI_PCG uses M^-1 = diag(1/(A_ii + lambda));
 * PCG uses M^-1 = diag(d) + U U^T (d: P per system, U: P x rank per system,
 * rank <= 64).  Iterative methods start at x = 0 and stop at
 * ||r|| / ||b|| <= tol or maxit.  P <= 1024. */
enum { QGTL_SOLVE_CHOLESKY = 0, QGTL_SOLVE_CG = 1, QGTL_SOLVE_JACOBI_PCG = 2, QGTL_SOLVE_PCG = 3 };
typedef struct {
    int32_t status;          /* 0 ok, -1 not PD / not converged */
    int32_t iterations;
    double rel_residual;
    double device_ms;        /* kernel time of the whole batch (HIP events) */
} qgtl_solve_info_t;
int qgtl_solve_f64(int P, int batch, const double *A, const double *b, double lambda, int method, double tol,
                   int maxit, const double *d, const double *U, int rank, double *x, qgtl_solve_info_t *info);
/* No counterpart in QGTL: qgtl_solve_f64 for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results; every device
 * buffer is carved from one fixed per-GPU array of QGTL_NOALLOC_SCRATCH_BYTES
 * (32 MiB, held from library init until the process ends), whose address is
 * looked up at init.  Needs 8 (P^2 + 3 P + rank P) batch bytes plus 16 batch
 * (each rounded up to 256); larger calls, and batch > QGTL_NOALLOC_MAX_BATCH,
 * fail with QGT_ERROR_INVALID_ARGUMENT.  info[].device_ms is host wall time.
 * Calls share the array and keep the last call's contents; the library lock
 * orders calls from one process, but work captured into graphs and replayed
 * concurrently on one GPU would race on it. */
#define QGTL_NOALLOC_MAX_BATCH 4096
int qgtl_solve_f64_noalloc(int P, int batch, const double *A, const double *b, double lambda, int method,
                           double tol, int maxit, const double *d, const double *U, int rank, double *x,
                           qgtl_solve_info_t *info);
/* Eigenvalues (ascending) and eigenvectors (columns of V, row-major) of a
 * Hermitian complex128 matrix, parallel Jacobi on the GPU.  V may be NULL. */
int qgtl_hermitian_eig(int P, const double *A, double *evals, double *V);

/* ---- learned preconditioner for the metric solve ----
 * A small MLP maps features x of one parameter point to the diagonal d and
 * low-rank factor U of M^-1 = diag(d) + U U^T for the PCG solve of
 * (Re Q + lambda I) x = grad E (QGTL_SOLVE_PCG).  hidden_layers dense layers
 * of width `hidden` (tanh-form GELU), BF16 weights and MFMA operands with
 * FP32 accumulation, FP32 Adam on Kaporin's condition number of the
 * preconditioned, Jacobi-scaled matrix As = D^-1/2 A D^-1/2 (D = diag A):
 * with y the network output, s_k = y[k(1+r)], u_kj = y[k(1+r)+1+j] / sqrt(P),
 * d = e^s / D, U = D^-1/2 u (s = 0, u = 0 is Jacobi).  No QGTL counterpart;
 * it preconditions compute_natural_gradient_cpu's CG
 * (src/quantum_geometric/hybrid/classical_optimization_engine.c:543-610). */
typedef struct qgtl_mlp qgtl_mlp_t;
qgtl_mlp_t *qgtl_mlp_create(int in_dim, int hidden, int hidden_layers, int P, int rank, uint64_t seed);
size_t qgtl_mlp_num_params(const qgtl_mlp_t *m);
/* FP32 parameter vector: W_0 (in x out, row-major), b_0, W_1, b_1, ...; set resets Adam */
int qgtl_mlp_get_params(const qgtl_mlp_t *m, float *params);
int qgtl_mlp_set_params(qgtl_mlp_t *m, const float *params);
/* training set on the GPU: S feature rows (S x in_dim) and S Jacobi-scaled matrices As (S x P x P) */
int qgtl_mlp_set_data(qgtl_mlp_t *m, int S, const float *X, const float *As);
/* mean loss over the samples idx[0..B) and its gradient (grad may be NULL) */
int qgtl_mlp_loss_grad(qgtl_mlp_t *m, int B, const int *idx, double *loss, float *grad);
/* one Adam step (beta 0.9 / 0.999) on the samples idx; *loss: batch mean before the step */
int qgtl_mlp_train_step(qgtl_mlp_t *m, int B, const int *idx, float lr, double *loss);
/* network output Y (B x P(1+r)) for B feature rows X */
int qgtl_mlp_forward(qgtl_mlp_t *m, int B, const float *X, float *Y);
/* features of one point: sin theta (P), cos theta (P), log D (P), field (in_dim = 3P + 1) */
int qgtl_pc_features(int P, const double *theta, const double *D, double field, float *x);
/* d (P) and U (P x rank) of M^-1 for one system from its features and the diagonal D of A */
int qgtl_mlp_preconditioner(qgtl_mlp_t *m, const float *x, const double *D, double *d, double *U);
void qgtl_mlp_destroy(qgtl_mlp_t *m);

/* ---- files: computed QGTs, training sets, network checkpoints ----
 * QGTL declares geometric_tensor_save / geometric_tensor_load
 * (include/quantum_geometric/core/quantum_geometric_tensor.h:144-147) without
 * an implementation or a format; these are this library's formats for what the
 * QGT path produces.  One container for all three: magic, version, sizes and an
 * FNV-1a 64 checksum over header and payload, all checked on load (a mismatch
 * is QGT_ERROR_VALIDATION_FAILED, a different P/S/F QGT_ERROR_DIMENSION_MISMATCH,
 * an unreadable file QGT_ERROR_IO_ERROR). */
int qgtl_qgt_save(const char *path, int n, int P, const double *theta /* P or NULL */,
                  const double *Q /* P x P complex128 */);
int qgtl_qgt_file_info(const char *path, int *n, int *P, int *has_theta);
int qgtl_qgt_load(const char *path, int P, double *theta /* P or NULL */, double *Q);
int qgtl_dataset_save(const char *path, int S, int F, int P, const float *X /* S x F */,
                      const float *As /* S x P x P */);
int qgtl_dataset_file_info(const char *path, int *S, int *F, int *P);
int qgtl_dataset_load(const char *path, int S, int F, int P, float *X, float *As);
/* fp32 master weights [+ Adam moments and step]; load builds a network of the saved shape */
int qgtl_mlp_save(const qgtl_mlp_t *m, const char *path, int with_optimizer);
qgtl_mlp_t *qgtl_mlp_load(const char *path);
int qgtl_mlp_get_shape(const qgtl_mlp_t *m, int *in_dim, int *hidden, int *hidden_layers, int *P, int *rank);

/* ---- natural-gradient VQE step (QGTL: run_vqe / update_natural_gradient,
 * src/quantum_geometric/hybrid/quantum_classical_algorithms.c:854,
 * src/quantum_geometric/hybrid/classical_optimization_engine.c:390-416) ----
 * theta (P, parameter order) -> E, grad E and Q at theta on the GPU; solve
 * (Re Q + lambda I) x = grad E with `method`; theta -= eta x.  The plan must be
 * mixed-mode with a Hamiltonian set; QGTL_SOLVE_PCG uses the learned
 * preconditioner `mlp` (features: theta, diag(Re Q) + lambda, `field`) with
 * Jacobi-PCG as the fallback when it misses the tolerance. */
typedef struct {
    double lambda;
    double eta;
    int32_t method;          /* QGTL_SOLVE_* */
    int32_t maxit;
    double tol;              /* relative residual of the iterative methods */
} qgtl_ng_config_t;
typedef struct {
    double energy;           /* at the incoming theta */
    double grad_norm;
    double step_norm;
    int32_t solve_iterations;  /* including a fallback solve */
    int32_t solve_status;
    int32_t fallback;          /* 1: the learned PCG missed the tolerance and Jacobi-PCG solved */
    int32_t pad;
    double solve_rel_residual;
    double solve_device_ms;
    double qgt_ms;           /* host wall: angles, QGT, energy and gradient */
    double solve_ms;         /* host wall: preconditioner and solve */
    double step_ms;          /* host wall of the whole step */
} qgtl_ng_info_t;
int qgtl_ng_step(qgtl_plan_t *plan, double *theta, const qgtl_ng_config_t *cfg, qgtl_mlp_t *mlp, double field,
                 qgtl_ng_info_t *info);

/* Diagnostics: v_mfma_f32_32x32x16_bf16 lane-layout probe; writes the 32 x 32
 * product D[i][j] = (1 + i/16)(i%16 + 16 (j%8)) as the Gram kernel reads it. */
int qgtl_debug_mfma_layout(float *host_out);
/* Diagnostics: v_mfma_f32_16x16x32_f16 lane-layout probe (the 16 x 16 form of the two-tile
 * Gram); writes the 16 x 16 product D[i][j] = i + 16 (i & 1) + 32 j as that kernel reads it. */
int qgtl_debug_mfma_layout16(float *host_out);

/* Diagnostics: one plain copy kernel of `bytes` (multiple of 16) between two
 * fresh device buffers; *ms = its HIP-event time.  Reference of known size
 * for the HBM traffic counters (FETCH_SIZE / WRITE_SIZE). */
int qgtl_debug_copy_kernel(size_t bytes, double *ms);
/* Device and pinned-host allocations and frees the library has made so far
 * (every one goes through one runtime function): lets a test show that a call
 * allocated nothing. */
unsigned long long qgtl_debug_alloc_calls(void);

/* Diagnostics: largest |diag(post) R diag(pre) - U| over the sweep passes'
 * gates (the rotation form of the pass descriptors) and how many passes
 * carry a pre-phase layer.  Host only. */
int qgtl_plan_debug_form(const qgtl_plan_t *plan, double *max_err, int *passes_with_pre);

/* Diagnostics: raw FP64 Gram C = Z Z^T (R x R) and injection sums (P x 4)
 * of the last mixed-mode execution of plan, in D-slot order. */
int qgtl_plan_debug_buffers(qgtl_plan_t *plan, double *C_out, double *inj_out, int *R_out);

#ifdef __cplusplus
}
#endif
#endif
