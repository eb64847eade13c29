/*
 * qgtl_ngvqe.c -- one natural-gradient VQE step on the GPU.
 *
 * The caller of the QGT path in QGTL is the VQE loop: run_vqe
 * (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:854) with
 * update_natural_gradient
 * (src/quantum_geometric/hybrid/classical_optimization_engine.c:390-416),
 * which solves (G + lambda I) x = grad E by conjugate gradient
 * (compute_natural_gradient_cpu, :543-610) and steps theta -= eta x.  Here one
 * step stays on the device from the QGT to the solution: new angles into the
 * plan (kernel descriptors), the mixed-precision QGT and, from the same
 * derivative states, E and grad E; A = Re Q + lambda I in parameter order;
 * the solve (Cholesky, CG, Jacobi-PCG, or PCG with the learned diagonal +
 * low-rank preconditioner, whose network runs on the device too, falling
 * back to Jacobi-PCG when it misses the tolerance); then x and E come back
 * (P + 1 doubles) and theta -= eta x on the host.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qgtl_internal.h"

#define NTRY(call)                                                                              \
    do {                                                                                        \
        int _rc = (call);                                                                       \
        if (_rc) { rc = qgtl_set_error(_rc, "%s: %s", #call, qgtl_hip_last_error()); goto out; } \
    } while (0)

static double wall_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static int ensure_ng_buffers(qgtl_plan_t *p, int rank)
{
    const size_t P = (size_t)p->P;
    int rc = 0;
    if (!p->ng_A) {
        int *slot = (int *)malloc(sizeof(int) * P);
        if (!slot) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        for (size_t s = 0; s < P; s++) slot[p->slot_param[s]] = (int)s;
        NTRY(qgtl_hip_malloc((void **)&p->ng_slot, sizeof(int) * P));
        NTRY(qgtl_hip_h2d(p->ng_slot, slot, sizeof(int) * P));
        NTRY(qgtl_hip_malloc((void **)&p->ng_A, sizeof(double) * P * P));
        NTRY(qgtl_hip_malloc((void **)&p->ng_D, sizeof(double) * P));
        NTRY(qgtl_hip_malloc((void **)&p->ng_g, sizeof(double) * P));
        NTRY(qgtl_hip_malloc((void **)&p->ng_x, sizeof(double) * P));
        NTRY(qgtl_hip_malloc((void **)&p->ng_d, sizeof(double) * P));
        NTRY(qgtl_hip_malloc((void **)&p->ng_theta, sizeof(double) * P));
        NTRY(qgtl_hip_malloc((void **)&p->ng_res, sizeof(double)));
        NTRY(qgtl_hip_malloc((void **)&p->ng_it, sizeof(int)));
        NTRY(qgtl_hip_sync());
    out:
        free(slot);
        if (rc) return rc;
    }
    if (rank > p->ng_rank) {
        qgtl_hip_free(p->ng_U);
        p->ng_U = NULL;
        p->ng_rank = 0;
        int r2 = qgtl_hip_malloc((void **)&p->ng_U, sizeof(double) * P * (size_t)rank);
        if (r2) return qgtl_set_error(r2, "ng buffers: %s", qgtl_hip_last_error());
        p->ng_rank = rank;
    }
    return 0;
}

/* one solve on the device; *iters / *relres / *status from the device */
static int solve_device(qgtl_plan_t *p, const qgtl_ng_config_t *cfg, int method, int rank, int *iters, double *relres,
                        int *status)
{
    const int P = p->P;
    int rc = 0;
    NTRY(qgtl_hip_fill_f32(p->ng_x, (size_t)P * 2, 0.f));       /* x0 = 0 (P doubles = 2P floats) */
    if (method == QGTL_SOLVE_CHOLESKY) {
        NTRY(qgtl_hip_cholesky_solve_f64(p->ng_A, p->ng_g, p->ng_x, P, 1, 0.0, p->ng_it));
        int st = 0;
        NTRY(qgtl_hip_d2h(&st, p->ng_it, sizeof(int)));
        NTRY(qgtl_hip_sync());
        *iters = 1;
        *relres = 0.0;
        *status = st;
        return 0;
    }
    const double *d = NULL, *U = NULL;
    if (method == QGTL_SOLVE_JACOBI_PCG) {
        NTRY(qgtl_hip_reciprocal(p->ng_D, P, p->ng_d));
        d = p->ng_d;
    } else if (method == QGTL_SOLVE_PCG) {
        d = p->ng_d;
        U = rank > 0 ? p->ng_U : NULL;
    }
    NTRY(qgtl_hip_pcg_f64(p->ng_A, p->ng_g, p->ng_x, P, 1, 0.0, d, U, U ? rank : 0, cfg->tol, cfg->maxit, p->ng_it,
                          p->ng_res));
    NTRY(qgtl_hip_d2h(iters, p->ng_it, sizeof(int)));
    NTRY(qgtl_hip_d2h(relres, p->ng_res, sizeof(double)));
    NTRY(qgtl_hip_sync());
    *status = (*relres <= cfg->tol) ? 0 : -1;
out:
    return rc;
}

int qgtl_ng_step(qgtl_plan_t *p, double *theta, const qgtl_ng_config_t *cfg, qgtl_mlp_t *mlp, double field,
                 qgtl_ng_info_t *info)
{
    QGTL_LOCKED();
    if (!p || !theta || !cfg) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: null argument");
    if (p->mode != QGTL_MODE_MIXED) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: mixed-mode plan");
    if (p->nterms < 1) return qgtl_set_error(QGT_ERROR_INVALID_STATE, "qgtl_ng_step: no Hamiltonian set");
    if (cfg->method == QGTL_SOLVE_PCG && !mlp)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: PCG needs the learned preconditioner");
    if (cfg->method < QGTL_SOLVE_CHOLESKY || cfg->method > QGTL_SOLVE_PCG)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: unknown method %d", cfg->method);
    const int P = p->P;
    if (P < 1 || P > 1024) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_ng_step: needs 1 <= P <= 1024");
    int rank = 0, mF = 0, mP = 0;
    if (cfg->method == QGTL_SOLVE_PCG) {
        qgtl_mlp_shape(mlp, &mF, &mP, &rank);
        if (mP != P || mF != 3 * P + 1)
            return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "learned preconditioner built for another size");
    }
    const double t0 = wall_ms();
    int rc = ensure_ng_buffers(p, rank);
    if (rc) return rc;
    double *angles = (double *)malloc(sizeof(double) * (size_t)(p->nops > 0 ? p->nops : 1));
    double *x = (double *)malloc(sizeof(double) * (size_t)P);
    if (!angles || !x) { free(angles); free(x); return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); }
    for (int i = 0; i < p->nops; i++) angles[i] = p->ops[i].param >= 0 ? theta[p->ops[i].param] : p->ops[i].angle;
    rc = qgtl_plan_set_angles(p, p->nops, angles);
    if (rc) goto out;
    qgtl_stats_t st;
    rc = qgtl_plan_execute(p, NULL, &st);             /* Q stays on the device (slot order) */
    if (rc) goto out;
    rc = qgtl_energy_gradient_device(p);
    if (rc) goto out;
    NTRY(qgtl_hip_ng_assemble(p->Qdev, p->ng_slot, P, cfg->lambda, p->eg_out, p->ng_A, p->ng_D, p->ng_g));
    double E = 0.0;
    NTRY(qgtl_hip_d2h(&E, p->eg_out, sizeof(double)));
    NTRY(qgtl_hip_sync());
    const double t1 = wall_ms();
    if (cfg->method == QGTL_SOLVE_PCG) {
        float *xin = NULL;
        const float *y = NULL;
        NTRY(qgtl_hip_h2d(p->ng_theta, theta, sizeof(double) * (size_t)P));
        rc = qgtl_mlp_device_input(mlp, &xin);
        if (rc) goto out;
        NTRY(qgtl_hip_pc_features(p->ng_theta, p->ng_D, field, P, xin));
        rc = qgtl_mlp_forward_device(mlp, &y);
        if (rc) goto out;
        NTRY(qgtl_hip_pc_transform(y, p->ng_D, P, rank, p->ng_d, p->ng_U));
    }
    int iters = 0, status = 0, fallback = 0;
    double relres = 0.0;
    rc = solve_device(p, cfg, cfg->method, rank, &iters, &relres, &status);
    if (rc) goto out;
    if (cfg->method == QGTL_SOLVE_PCG && status != 0) {
        /* the learned preconditioner missed the tolerance: Jacobi-PCG from scratch */
        const int it = iters;
        rc = solve_device(p, cfg, QGTL_SOLVE_JACOBI_PCG, 0, &iters, &relres, &status);
        if (rc) goto out;
        iters += it;
        fallback = 1;
    }
    NTRY(qgtl_hip_d2h(x, p->ng_x, sizeof(double) * (size_t)P));
    NTRY(qgtl_hip_sync());
    const double t2 = wall_ms();
    double gn = 0.0, xn = 0.0;
    {
        double *g = (double *)malloc(sizeof(double) * (size_t)P);
        if (g) {
            if (!qgtl_hip_d2h(g, p->ng_g, sizeof(double) * (size_t)P) && !qgtl_hip_sync())
                for (int k = 0; k < P; k++) gn += g[k] * g[k];
            free(g);
        }
    }
    for (int k = 0; k < P; k++) {
        xn += x[k] * x[k];
        theta[k] -= cfg->eta * x[k];
    }
    if (info) {
        memset(info, 0, sizeof *info);
        info->energy = E;
        info->grad_norm = sqrt(gn);
        info->step_norm = cfg->eta * sqrt(xn);
        info->solve_iterations = iters;
        info->solve_status = status;
        info->fallback = fallback;
        info->solve_rel_residual = relres;
        info->solve_device_ms = t2 - t1;
        info->qgt_ms = t1 - t0;
        info->solve_ms = t2 - t1;
        info->step_ms = wall_ms() - t0;
    }
out:
    free(angles);
    free(x);
    return rc;
}
