/*
 * qgt_oracle.h -- FP64 CPU restatement of QGTL's quantum-geometric-tensor path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP library
 * (quantum_geometric_tensor_amd/csrc).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links or
 * calls it.  It is compiled with -O3 and WITHOUT -ffast-math so that every
 * reduction is plain IEEE double arithmetic in a fixed order.
 *
 * What it restates (reference paths relative to the QGTL source root):
 *   gates       src/quantum_geometric/core/quantum_gate_operations.c:11-47
 *               src/quantum_geometric/hybrid/quantum_classical_algorithms.c:174-265
 *   parameters  src/quantum_geometric/core/quantum_parameter_shift.c:278-337
 *               (parameterised gates are numbered in gate order)
 *   QGT         src/quantum_geometric/core/quantum_geometric_tensor_network.c:1120-1175
 *               Q_mn = <d_m psi|d_n psi> - <d_m psi|psi><psi|d_n psi>
 *               src/quantum_geometric/distributed/differential_geometry.c:2819-2906
 *               (P x P FP64 loop over explicit derivative states)
 *   full QGT    src/quantum_geometric/core/quantum_geometric_curvature.c:289-320
 *   metric      src/quantum_geometric/core/quantum_geometric_metric.c:353-396 (Re Q)
 *   Berry       src/quantum_geometric/core/quantum_geometric_curvature.c:201-247 (Im Q)
 *   Pauli <H>   src/quantum_geometric/hybrid/quantum_classical_algorithms.c:320-380
 *   NG solves   src/quantum_geometric/core/quantum_geometric_gradient.c:2721-2964
 *               src/quantum_geometric/hybrid/classical_optimization_engine.c:543-610
 *               src/quantum_geometric/learning/quantum_pipeline_impl.c:1601-1682
 *
 * Derivative states are the exact ones, |d_k psi> = U_{>k} (-i/2 G_k) U_{<=k}|0>,
 * not the reference's shifted-difference approximation
 * (quantum_geometric_gradient.c:2356-2366), which is not a derivative.
 *
 * Pins (tests/test_oracle.py): the pi-shift identity d_k psi = psi(theta+pi e_k)/2,
 * central finite differences, dense Kronecker brute force for n <= 4, the
 * single-qubit closed form, the layer-1 product-state block, Hermiticity,
 * PSD, ||d_k psi||^2 = 1/4, Re<d_k psi|psi> = 0, and the reference's own
 * state-magnitude tests (Bell, GHZ).
 */
#ifndef QGT_ORACLE_H
#define QGT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate codes are QGTL's gate_type_t values
 * (include/quantum_geometric/core/quantum_base_types.h:32-72). */
enum {
    ORC_I = 0, ORC_X = 1, ORC_Y = 2, ORC_Z = 3, ORC_H = 4, ORC_S = 5, ORC_T = 6,
    ORC_RX = 7, ORC_RY = 8, ORC_RZ = 9, ORC_CNOT = 10, ORC_CY = 11, ORC_CZ = 12,
    ORC_SWAP = 13, ORC_U1 = 15, ORC_PHASE = 19, ORC_SDG = 26, ORC_TDG = 27, ORC_SX = 29
};

/* One gate.  Single-qubit: q0 = target, q1 = -1.  CNOT/CY/CZ: q0 = control,
 * q1 = target (the {control, target} order of create_cnot_gate,
 * quantum_gate_operations.c:420-423).  SWAP: q0, q1.  param = index of the
 * variational parameter carried by the gate, or -1. */
typedef struct {
    int32_t type;
    int32_t q0;
    int32_t q1;
    int32_t param;
    double angle;
} oracle_op_t;

/* --- synthetic inputs -------------------------------------------------- */
uint64_t oracle_splitmix64_next(uint64_t *state);
double   oracle_uniform01(uint64_t *state);               /* (x>>11)*2^-53 */
void     oracle_uniform_angles(uint64_t seed, size_t count, double *theta); /* [0, 2pi) */

/* Hardware-efficient ansatz: per layer RY(q) for q=0..n-1, RZ(q) for
 * q=0..n-1, then CNOT(i -> i+1 mod n) for i = 0..n-1 (no ring for n = 1).
 * Parameter k = 2nl+q (RY), 2nl+n+q (RZ).  Returns the op count. */
int oracle_hea_ops(int n, int layers, const double *theta, oracle_op_t *ops);
int oracle_hea_num_ops(int n, int layers);

/* --- state vectors (complex128, interleaved re,im; qubit q <-> bit 1<<q) - */
int  oracle_num_params(int nops, const oracle_op_t *ops);
int  oracle_gate_matrix(const oracle_op_t *op, double *m /* 8 doubles */);
void oracle_apply_1q(int n, double *psi, int q, const double *m);
int  oracle_apply_op(int n, double *psi, const oracle_op_t *op);
int  oracle_run(int n, int nops, const oracle_op_t *ops, double *psi);

/* D is [P][2^n] complex.  Generator insertion (exact). */
int oracle_derivative_states(int n, int nops, const oracle_op_t *ops, double *psi, double *D);
/* Same states from the identity R_G(a+pi) = -i G R_G(a): d_k psi = psi(theta+pi e_k)/2. */
int oracle_derivative_states_shift(int n, int nops, const oracle_op_t *ops, double *D);
/* Only the listed parameters (for sub-block checks at large n). */
int oracle_derivative_subset(int n, int nops, const oracle_op_t *ops, int count,
                             const int *params, double *psi, double *D);

/* QGT from explicit states.  Q is [P][P] complex row-major. */
void oracle_qgt_pairwise(size_t dim, int P, const double *psi, const double *D, double *Q);
void oracle_qgt_elementwise(size_t dim, int P, const double *psi, const double *D, double *Q);
int  oracle_qgt(int n, int nops, const oracle_op_t *ops, double *Q);
/* Same values, overlaps hoisted and upper triangle only (timed CPU baseline). */
int  oracle_qgt_hoisted(int n, int nops, const oracle_op_t *ops, double *Q);
/* CPU baseline timing: the derivative states on `threads` OpenMP threads
 * (<= 0: the default) with their wall time in *seconds; and rows [0, rows) of
 * the per-element loop (variant 0: mu <= nu with three dot products per
 * element, the structure of quantum_geometric_curvature.c:300-317 over
 * quantum_geometric_tensor_network.c:1127-1166) or of the P x P loop (variant
 * 1: differential_geometry.c:2832-2858) over given states, `reps` times,
 * seconds in sample_s[]; *elements = Q elements one repetition computes. */
int  oracle_derivative_states_timed(int n, int nops, const oracle_op_t *ops, int threads, double *psi, double *D,
                                    double *seconds);
int  oracle_qgt_loop_time(size_t dim, int P, const double *psi, const double *D, int variant, int threads, int rows,
                          int reps, double *sample_s, long *elements);
/* Q[r][c] for the listed rows/cols. */
int  oracle_qgt_block(int n, int nops, const oracle_op_t *ops, int nr, const int *rows,
                      int nc, const int *cols, double *Q);

/* Dense 2^n x 2^n unitary by Kronecker products (n <= 6), row-major complex. */
int oracle_dense_unitary(int n, int nops, const oracle_op_t *ops, double *U);

/* --- Hamiltonians: sum_t coeff[t] * pauli[t], pauli strings with char 0 on
 *     the highest qubit (quantum_classical_algorithms.c:339). ---------- */
double oracle_pauli_expectation(int n, const double *psi, const char *pauli);
/* TFIM open chain H = -J sum_{i<n-1} Z_i Z_{i+1} - h sum_i X_i. */
void   oracle_tfim_apply(int n, double J, double h, const double *psi, double *hpsi);
double oracle_tfim_energy(int n, double J, double h, const double *psi);
/* grad_k = 2 Re <d_k psi| H |psi>. */
int    oracle_tfim_gradient(int n, int nops, const oracle_op_t *ops, double J, double h,
                            double *energy, double *grad);

/* --- solves (FP64) ------------------------------------------------------ */
int oracle_cholesky_solve(int P, const double *A, const double *b, double *x);
/* Preconditioned CG on SPD A, x0 = 0, stop at ||r||/||b|| <= tol.
 * M^-1 = diag(d) + U U^T (U is P x r row-major); d = NULL means identity;
 * r = 0 means no low-rank part.  Returns iterations (or -1). */
int oracle_pcg(int P, const double *A, const double *b, double *x, double tol, int maxit,
               const double *d, const double *U, int r, double *final_rel_res);
/* Eigenvalues (ascending) and eigenvectors (columns of V, row-major) of a
 * complex Hermitian matrix by cyclic Jacobi. */
int oracle_hermitian_eig(int P, const double *A, double *evals, double *V);
/* compute_regularized_natural_gradient semantics
 * (quantum_geometric_gradient.c:2887-2964), complex, FP64:
 * adaptive lambda when kappa > threshold; (G + lambda I)^-1 g; pseudo-inverse
 * fallback with singular cutoff. Returns lambda used (negative on failure). */
double oracle_regularized_natural_gradient(int P, const double *G, const double *g, double *x,
                                           double lambda, double cond_threshold, int adaptive,
                                           int pinv_fallback, double cutoff);

#ifdef __cplusplus
}
#endif
#endif
