/*
 * qgtl_ngvqe.c -- one natural-gradient VQE step on the GPU.
 *
 * The caller of the QGT path in QGTL is the VQE loop: run_vqe
 * (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:854) with
 * update_natural_gradient
 * (src/quantum_geometric/hybrid/classical_optimization_engine.c:390-416),
 * which solves (G + lambda I) x = grad E by conjugate gradient
 * (compute_natural_gradient_cpu, :543-610) and steps theta -= eta x.  Here one
 * step is: new angles into the plan, the mixed-precision QGT and, from the
 * same derivative states, E and grad E on the GPU; A = Re Q + lambda I; the
 * solve on the GPU (Cholesky, CG, Jacobi-PCG, or PCG with the learned
 * diagonal + low-rank preconditioner, falling back to Jacobi-PCG when it
 * misses the tolerance); the step.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qgtl_internal.h"

static double wall_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

int qgtl_ng_step(qgtl_plan_t *p, double *theta, const qgtl_ng_config_t *cfg, qgtl_mlp_t *mlp, double field,
                 qgtl_ng_info_t *info)
{
    if (!p || !theta || !cfg) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: null argument");
    if (p->mode != QGTL_MODE_MIXED) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: mixed-mode plan");
    if (p->nterms < 1) return qgtl_set_error(QGT_ERROR_INVALID_STATE, "qgtl_ng_step: no Hamiltonian set");
    if (cfg->method == QGTL_SOLVE_PCG && !mlp)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: PCG needs the learned preconditioner");
    const int P = p->P;
    const double t0 = wall_ms();
    double *angles = (double *)malloc(sizeof(double) * (size_t)(p->nops > 0 ? p->nops : 1));
    double *Q = (double *)malloc(sizeof(double) * 2 * (size_t)P * P);
    double *A = (double *)malloc(sizeof(double) * (size_t)P * P);
    double *g = (double *)malloc(sizeof(double) * (size_t)P);
    double *x = (double *)calloc((size_t)P, sizeof(double));
    double *D = (double *)malloc(sizeof(double) * (size_t)P);
    double *d = NULL, *U = NULL;
    float *feat = NULL;
    int rc = 0;
    if (!angles || !Q || !A || !g || !x || !D) { rc = qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); goto out; }
    for (int i = 0; i < p->nops; i++)
        angles[i] = p->ops[i].param >= 0 ? theta[p->ops[i].param] : p->ops[i].angle;
    rc = qgtl_plan_set_angles(p, p->nops, angles);
    if (rc) goto out;
    qgtl_stats_t st;
    rc = qgtl_plan_execute(p, Q, &st);
    if (rc) goto out;
    double E = 0.0;
    rc = qgtl_plan_energy_gradient(p, &E, g);
    if (rc) goto out;
    const double t1 = wall_ms();
    for (int k = 0; k < P; k++)
        for (int l = 0; l < P; l++) A[(size_t)k * P + l] = Q[2 * ((size_t)k * P + l)];
    for (int k = 0; k < P; k++) D[k] = A[(size_t)k * P + k] + cfg->lambda;
    qgtl_solve_info_t si;
    memset(&si, 0, sizeof si);
    int rank = 0;
    if (cfg->method == QGTL_SOLVE_PCG) {
        int mF = 0, mP = 0;
        qgtl_mlp_shape(mlp, &mF, &mP, &rank);
        d = (double *)malloc(sizeof(double) * (size_t)P);
        U = (double *)malloc(sizeof(double) * (size_t)P * (rank > 0 ? rank : 1));
        feat = (float *)malloc(sizeof(float) * (size_t)(3 * P + 1));
        if (!d || !U || !feat) { rc = qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); goto out; }
        if (mP != P || mF != 3 * P + 1) {
            rc = qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "learned preconditioner built for another size");
            goto out;
        }
        rc = qgtl_pc_features(P, theta, D, field, feat);
        if (!rc) rc = qgtl_mlp_preconditioner(mlp, feat, D, d, U);
        if (rc) goto out;
    }
    rc = qgtl_solve_f64(P, 1, A, g, cfg->lambda, cfg->method, cfg->tol, cfg->maxit, d, U, rank, x, &si);
    int fallback = 0;
    if (!rc && cfg->method == QGTL_SOLVE_PCG && si.status != 0) {
        /* the learned preconditioner did not reach the tolerance: Jacobi-PCG from scratch */
        const int it = si.iterations;
        memset(x, 0, sizeof(double) * (size_t)P);
        rc = qgtl_solve_f64(P, 1, A, g, cfg->lambda, QGTL_SOLVE_JACOBI_PCG, cfg->tol, cfg->maxit, NULL, NULL, 0, x, &si);
        si.iterations += it;
        fallback = 1;
    }
    if (rc) goto out;
    const double t2 = wall_ms();
    double gn = 0.0, xn = 0.0;
    for (int k = 0; k < P; k++) {
        gn += g[k] * g[k];
        xn += x[k] * x[k];
        theta[k] -= cfg->eta * x[k];
    }
    if (info) {
        memset(info, 0, sizeof *info);
        info->energy = E;
        info->grad_norm = sqrt(gn);
        info->step_norm = cfg->eta * sqrt(xn);
        info->solve_iterations = si.iterations;
        info->solve_status = si.status;
        info->fallback = fallback;
        info->solve_rel_residual = si.rel_residual;
        info->solve_device_ms = si.device_ms;
        info->qgt_ms = t1 - t0;
        info->solve_ms = t2 - t1;
        info->step_ms = wall_ms() - t0;
    }
out:
    free(angles); free(Q); free(A); free(g); free(x); free(D); free(d); free(U); free(feat);
    return rc;
}
