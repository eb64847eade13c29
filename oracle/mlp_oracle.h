/*
 * mlp_oracle.h -- FP64 CPU restatement of the learned preconditioner's
 * network and training loss.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP training and inference
 * kernels (quantum_geometric_tensor_amd/csrc/hip/qgtl_mlp.hip).  Only tests/
 * load it.  QGTL has no learned preconditioner; its natural-gradient solves
 * (src/quantum_geometric/core/quantum_geometric_gradient.c:2887-2964,
 * src/quantum_geometric/hybrid/classical_optimization_engine.c:543-610) are
 * the methods this network preconditions.
 *
 * Network (nh hidden layers of width H):
 *   h_1 = g(x W_0 + b_0), h_{i+1} = g(h_i W_i + b_i), y = h_nh W_nh + b_nh,
 *   W_i row-major (in x out), g(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
 *   Parameter vector: W_0, b_0, W_1, b_1, ..., W_nh, b_nh.
 * Output of one sample (O = P (1 + r)): s_k = y[k (1 + r)],
 *   u_kj = y[k (1 + r) + 1 + j] / sqrt(P).
 * Preconditioner for A = Re Q + lambda I with D = diag(A) and the unit-
 *   diagonal As = D^-1/2 A D^-1/2:  Mt = diag(e^s) + u u^T,
 *   M^-1 = D^-1/2 Mt D^-1/2 = diag(e^s / D) + (D^-1/2 u)(D^-1/2 u)^T.
 * Loss: Kaporin's condition number of Mt As, without its constant -log det As:
 *   L = P log(tr(Mt As) / P) - log det Mt,
 *   log det Mt = sum_k s_k + log det(I_r + u^T diag(e^-s) u).
 * It is invariant to the scale of Mt and bounds the conjugate-gradient
 * iteration count (I. Kaporin, Numer. Linear Algebra Appl. 1 (1994) 179-210).
 *
 * Pins (tests/test_mlp_oracle.py): central finite differences of the loss
 * with respect to every output and to sampled weights, Kaporin's value
 * against a dense evaluation (numpy slogdet / trace), invariance of L under
 * Mt -> c Mt.
 */
#ifndef MLP_ORACLE_H
#define MLP_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

size_t oracle_mlp_num_params(int F, int H, int nh, int O);
/* Y (B x O) = network(X (B x F)) */
void oracle_mlp_forward(int F, int H, int nh, int O, int B, const double *params, const double *X, double *Y);
/* Loss of one sample and dL/dy (dy may be NULL). */
double oracle_kaporin(int P, int r, const double *y, const double *As, double *dy);
/* Mean Kaporin loss over B samples (As: B x P x P) and its gradient with
 * respect to the parameters (grad may be NULL). */
double oracle_mlp_loss_grad(int F, int H, int nh, int P, int r, int B, const double *params, const double *X,
                            const double *As, double *grad);

#ifdef __cplusplus
}
#endif
#endif
