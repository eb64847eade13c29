/*
 * qgtl_hamiltonian.c -- Pauli-sum Hamiltonians, energy and gradient on the GPU.
 *
 * QGTL evaluates E = sum_t c_t <psi|P_t|psi> with one state copy and one
 * pass per Pauli string (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:320-427).
 * Here H psi is one gather kernel over all terms, and the gradient
 * grad_k = 2 Re <d_k psi|H|psi> is an FP64-accumulated product of H psi with
 * the derivative states the QGT left on the GPU, so a natural-gradient step
 * gets Q, grad and E from one set of sweeps.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_internal.h"

#define TRYH(call) do { int _rc = (call); if (_rc) return qgtl_set_error(_rc, "%s: %s", #call, qgtl_hip_last_error()); } while (0)

int qgtl_pauli_term_from_string(const char *pauli, double coeff, qgtl_pauli_term_t *out)
{
    if (!pauli || !out) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null Pauli string");
    const int n = (int)strlen(pauli);
    if (n < 1 || n > QGTL_MAX_QUBITS) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "Pauli string length %d", n);
    qgtl_pauli_term_t t = {0u, 0u, coeff, 0.0};
    for (int q = 0; q < n; q++) {
        const char c = pauli[n - 1 - q];
        switch (c) {
        case 'I': case 'i': break;
        case 'X': case 'x': t.xmask |= 1u << q; break;
        case 'Z': case 'z': t.zmask |= 1u << q; break;
        case 'Y': case 'y': t.xmask |= 1u << q; t.zmask |= 1u << q; break;
        default: return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "bad Pauli character '%c'", c);
        }
    }
    *out = t;
    return 0;
}

int qgtl_tfim_terms(int n, double J, double h, qgtl_pauli_term_t *terms)
{
    if (n < 1 || n > QGTL_MAX_QUBITS || !terms) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "bad TFIM size");
    int k = 0;
    for (int i = 0; i + 1 < n; i++) {
        qgtl_pauli_term_t t = {0u, (1u << i) | (1u << (i + 1)), -J, 0.0};
        terms[k++] = t;
    }
    for (int i = 0; i < n; i++) {
        qgtl_pauli_term_t t = {1u << i, 0u, -h, 0.0};
        terms[k++] = t;
    }
    return k;
}

int qgtl_plan_set_hamiltonian(qgtl_plan_t *p, int nterms, const qgtl_pauli_term_t *terms)
{
    QGTL_LOCKED();
    if (!p || nterms < 1 || !terms) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "empty Hamiltonian");
    const uint32_t full = (p->n >= 32) ? 0xffffffffu : ((1u << p->n) - 1u);
    for (int t = 0; t < nterms; t++)
        if ((terms[t].xmask | terms[t].zmask) & ~full)
            return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "term %d acts on a qubit >= %d", t, p->n);
    int rc = qgtl_init_if_needed();
    if (rc) return rc;
    if (nterms > p->nterms) {
        qgtl_hip_free(p->terms_dev);
        p->terms_dev = NULL;
        TRYH(qgtl_hip_malloc(&p->terms_dev, sizeof(qgtl_pauli_term_t) * (size_t)nterms));
    }
    TRYH(qgtl_hip_h2d(p->terms_dev, terms, sizeof(qgtl_pauli_term_t) * (size_t)nterms));
    TRYH(qgtl_hip_sync());
    p->nterms = nterms;
    return 0;
}

int qgtl_energy_gradient_device(qgtl_plan_t *p)
{
    if (!p || p->mode != QGTL_MODE_MIXED) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "needs a mixed-mode plan");
    if (!p->psi_final) return qgtl_set_error(QGT_ERROR_INVALID_STATE, "execute the plan first");
    if (p->nterms < 1) return qgtl_set_error(QGT_ERROR_INVALID_STATE, "no Hamiltonian set");
    const size_t N = p->N, P = (size_t)p->P;
    const size_t tiles = (N + 2047) / 2048;
    if (!p->hpsi) {
        TRYH(qgtl_hip_malloc(&p->hpsi, N * 8));
        TRYH(qgtl_hip_malloc((void **)&p->eg_partials, (P + 1) * tiles * 4 * sizeof(double)));
        TRYH(qgtl_hip_malloc((void **)&p->eg_sums, (P + 1) * 4 * sizeof(double)));
        TRYH(qgtl_hip_malloc((void **)&p->eg_out, (P + 1) * sizeof(double)));
    }
    int t1 = 0, t2 = 0;
    TRYH(qgtl_hip_pauli_apply_c64(p->psi_final, p->hpsi, p->n, p->nterms, p->terms_dev));
    TRYH(qgtl_hip_dot_states_c64(p->psi_final, N, 1, p->hpsi, p->n, p->eg_partials + P * tiles * 4, &t2, 0));
    TRYH(qgtl_hip_reduce_partials_f64(p->eg_partials + P * tiles * 4, 1, t2, p->eg_sums + 4 * P));
    if (P > 0) {
        /* the derivative states sit after block meet: <d_k|H psi> there needs
         * H psi carried back by the same inverse blocks (the other primal
         * buffer is free scratch after the execution) */
        void *h = p->hpsi, *scratch = (p->psi_final == p->psi[0]) ? p->psi[1] : p->psi[0];
        int rc = qgtl_carry_to_meet(p, &h, &scratch);
        if (rc) return rc;
        TRYH(qgtl_hip_dot_states_c64(p->D, N, (int)P, h, p->n, p->eg_partials, &t1, p->presplit));
        TRYH(qgtl_hip_reduce_partials_f64(p->eg_partials, (int)P, t1, p->eg_sums));
    }
    TRYH(qgtl_hip_energy_gradient_finalize(p->eg_sums, p->eg_sums + 4 * P, (int)P, p->eg_out));
    return 0;
}

int qgtl_plan_energy_gradient(qgtl_plan_t *p, double *energy, double *grad)
{
    QGTL_LOCKED();
    int rc0 = qgtl_energy_gradient_device(p);
    if (rc0) return rc0;
    const size_t P = (size_t)p->P;
    double *host = (double *)malloc((P + 1) * sizeof(double));
    if (!host) return QGT_ERROR_NO_MEMORY;
    int rc = qgtl_hip_d2h(host, p->eg_out, (P + 1) * sizeof(double));
    if (!rc) rc = qgtl_hip_sync();
    if (rc) { free(host); return qgtl_set_error(rc, "energy/gradient: %s", qgtl_hip_last_error()); }
    if (energy) *energy = host[0];
    if (grad)
        for (size_t s = 0; s < P; s++) grad[p->slot_param[s]] = host[1 + s];
    free(host);
    return 0;
}
