/*
 * qgtl_solve.c -- natural-gradient solves on the GPU (FP64).
 *
 * Additive: qgtl_solve_f64 (Cholesky / CG / Jacobi-PCG / diag+low-rank PCG on
 * (A + lambda I) x = b, batched), qgtl_hermitian_eig.
 * QGTL API: get_default_natural_gradient_config,
 * compute_complex_matrix_condition_number, compute_pseudoinverse_natural_gradient,
 * compute_regularized_natural_gradient
 * (src/quantum_geometric/core/quantum_geometric_gradient.c:2721-2964), with the
 * LU, SVD and matrix products moved to FP64 GPU kernels (qgtl_solve.hip).
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qgtl_api.h"
#include "qgtl_internal.h"

/* grow-only device scratch for the solves */
enum { S_A, S_B, S_X, S_V, S_EV, S_INT, S_D, S_U, S_IT, S_RES, S_COUNT };
static void *g_scr[S_COUNT];
static size_t g_scr_len[S_COUNT];

static int scratch(int slot, size_t bytes, void **out)
{
    if (g_scr_len[slot] < bytes) {
        qgtl_hip_free(g_scr[slot]);
        g_scr[slot] = NULL;
        g_scr_len[slot] = 0;
        int rc = qgtl_hip_malloc(&g_scr[slot], bytes);
        if (rc) return qgtl_set_error(rc, "solve scratch: %s", qgtl_hip_last_error());
        g_scr_len[slot] = bytes;
    }
    *out = g_scr[slot];
    return 0;
}

#define TRY(call) do { int _rc = (call); if (_rc) return _rc < -100 ? qgtl_set_error(_rc, "%s: %s", #call, qgtl_hip_last_error()) : _rc; } while (0)

int qgtl_solve_f64(int P, int batch, const double *A, const double *b, double lambda, int method, double tol,
                   int maxit, const double *d, const double *U, int rank, double *x, qgtl_solve_info_t *info)
{
    QGTL_LOCKED();
    if (P < 1 || P > 1024 || batch < 1 || !A || !b || !x) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_solve_f64: bad arguments");
    if (method == QGTL_SOLVE_PCG && (rank < 0 || rank > 64 || (rank > 0 && !U)))
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_solve_f64: low-rank factor needs 0 <= rank <= 64");
    TRY(qgtl_init_if_needed());
    const size_t PP = (size_t)P * P * batch, Pb = (size_t)P * batch;
    void *dA, *db, *dx, *dst, *dd = NULL, *dU = NULL, *dit, *dres;
    TRY(scratch(S_A, PP * sizeof(double), &dA));
    TRY(scratch(S_B, Pb * sizeof(double), &db));
    TRY(scratch(S_X, Pb * sizeof(double), &dx));
    TRY(scratch(S_INT, (size_t)batch * sizeof(int), &dst));
    TRY(scratch(S_IT, (size_t)batch * sizeof(int), &dit));
    TRY(scratch(S_RES, (size_t)batch * sizeof(double), &dres));
    TRY(qgtl_hip_h2d(dA, A, PP * sizeof(double)));
    TRY(qgtl_hip_h2d(db, b, Pb * sizeof(double)));
    qgtl_hip_event_t e0 = NULL, e1 = NULL;
    TRY(qgtl_hip_event_create(&e0));
    TRY(qgtl_hip_event_create(&e1));
    TRY(qgtl_hip_event_record(e0));
    if (method == QGTL_SOLVE_CHOLESKY) {
        TRY(qgtl_hip_cholesky_solve_f64((double *)dA, (const double *)db, (double *)dx, P, batch, lambda, (int *)dst));
    } else {
        double *dh = NULL;
        if (method == QGTL_SOLVE_JACOBI_PCG) {
            /* d = 1 / (diag(A) + lambda) */
            dh = (double *)malloc(Pb * sizeof(double));
            if (!dh) return QGT_ERROR_NO_MEMORY;
            for (int s = 0; s < batch; s++)
                for (int i = 0; i < P; i++) dh[(size_t)s * P + i] = 1.0 / (A[(size_t)s * P * P + (size_t)i * P + i] + lambda);
            d = dh;
        }
        if (d && method != QGTL_SOLVE_CG) {
            TRY(scratch(S_D, Pb * sizeof(double), &dd));
            TRY(qgtl_hip_h2d(dd, d, Pb * sizeof(double)));
        }
        if (method == QGTL_SOLVE_PCG && rank > 0) {
            TRY(scratch(S_U, Pb * (size_t)rank * sizeof(double), &dU));
            TRY(qgtl_hip_h2d(dU, U, Pb * (size_t)rank * sizeof(double)));
        }
        int rc = qgtl_hip_pcg_f64((const double *)dA, (const double *)db, (double *)dx, P, batch, lambda,
                                  method == QGTL_SOLVE_CG ? NULL : (const double *)dd,
                                  method == QGTL_SOLVE_PCG ? (const double *)dU : NULL,
                                  method == QGTL_SOLVE_PCG ? rank : 0, tol, maxit, (int *)dit, (double *)dres);
        if (dh) { free(dh); d = NULL; }
        TRY(rc);
    }
    TRY(qgtl_hip_event_record(e1));
    TRY(qgtl_hip_d2h(x, dx, Pb * sizeof(double)));
    int *hst = (int *)malloc((size_t)batch * sizeof(int));
    int *hit = (int *)malloc((size_t)batch * sizeof(int));
    double *hres = (double *)malloc((size_t)batch * sizeof(double));
    if (!hst || !hit || !hres) { free(hst); free(hit); free(hres); return QGT_ERROR_NO_MEMORY; }
    if (method == QGTL_SOLVE_CHOLESKY) TRY(qgtl_hip_d2h(hst, dst, (size_t)batch * sizeof(int)));
    else {
        TRY(qgtl_hip_d2h(hit, dit, (size_t)batch * sizeof(int)));
        TRY(qgtl_hip_d2h(hres, dres, (size_t)batch * sizeof(double)));
    }
    TRY(qgtl_hip_sync());
    float ms = 0.f;
    qgtl_hip_event_elapsed_ms(e0, e1, &ms);
    qgtl_hip_event_destroy(e0);
    qgtl_hip_event_destroy(e1);
    int fail = 0;
    for (int s = 0; s < batch; s++) {
        qgtl_solve_info_t inf = {0, 0, 0.0, ms};
        if (method == QGTL_SOLVE_CHOLESKY) {
            inf.status = hst[s];
            inf.iterations = 1;
        } else {
            inf.iterations = hit[s];
            inf.rel_residual = hres[s];
            inf.status = (hres[s] <= tol) ? 0 : -1;
        }
        if (inf.status) fail = 1;
        if (info) info[s] = inf;
    }
    free(hst); free(hit); free(hres);
    if (fail) return qgtl_set_error(method == QGTL_SOLVE_CHOLESKY ? QGT_ERROR_MATRIX_SINGULAR : QGT_ERROR_CONVERGENCE_FAILED,
                                    method == QGTL_SOLVE_CHOLESKY ? "matrix not positive definite" : "CG did not reach the tolerance");
    return 0;
}

/* ---------------------------------------------- no-allocation variant */
/* Per-system results come back through these (the call holds the library
 * lock, so one call uses them at a time). */
static int g_na_status[QGTL_NOALLOC_MAX_BATCH];
static int g_na_iters[QGTL_NOALLOC_MAX_BATCH];
static double g_na_res[QGTL_NOALLOC_MAX_BATCH];

static size_t carve(size_t *off, size_t bytes)
{
    const size_t at = *off;
    *off += (bytes + 255) / 256 * 256;
    return at;
}

static double wall_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* qgtl_solve_f64 with every device buffer carved from the library's fixed
 * per-GPU scratch array (QGTL_NOALLOC_SCRATCH_BYTES) and no allocation of any
 * kind: the Jacobi preconditioner is formed on the device, per-system results
 * return through static host arrays, timing is host wall time. */
int qgtl_solve_f64_noalloc(int P, int batch, const double *A, const double *b, double lambda, int method, double tol,
                           int maxit, const double *d, const double *U, int rank, double *x, qgtl_solve_info_t *info)
{
    QGTL_LOCKED();
    if (P < 1 || P > 1024 || batch < 1 || !A || !b || !x)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_solve_f64_noalloc: bad arguments");
    if (batch > QGTL_NOALLOC_MAX_BATCH)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_solve_f64_noalloc: batch %d exceeds the maximum %d",
                              batch, QGTL_NOALLOC_MAX_BATCH);
    if (method != QGTL_SOLVE_CHOLESKY && method != QGTL_SOLVE_CG && method != QGTL_SOLVE_JACOBI_PCG &&
        method != QGTL_SOLVE_PCG)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_solve_f64_noalloc: unknown method %d", method);
    if (method == QGTL_SOLVE_PCG && (rank < 0 || rank > 64 || (rank > 0 && !U)))
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT,
                              "qgtl_solve_f64_noalloc: low-rank factor needs 0 <= rank <= 64");
    const size_t PP = (size_t)P * P * batch, Pb = (size_t)P * batch;
    const int pre = method == QGTL_SOLVE_JACOBI_PCG || (method == QGTL_SOLVE_PCG && d);
    const int lr = method == QGTL_SOLVE_PCG && rank > 0;
    size_t off = 0;
    const size_t oA = carve(&off, PP * sizeof(double)), ob = carve(&off, Pb * sizeof(double));
    const size_t ox = carve(&off, Pb * sizeof(double));
    const size_t od = pre ? carve(&off, Pb * sizeof(double)) : 0;
    const size_t oU = lr ? carve(&off, Pb * (size_t)rank * sizeof(double)) : 0;
    const size_t ost = carve(&off, (size_t)batch * sizeof(int)), oit = carve(&off, (size_t)batch * sizeof(int));
    const size_t ores = carve(&off, (size_t)batch * sizeof(double));
    if (off > QGTL_NOALLOC_SCRATCH_BYTES)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT,
                              "qgtl_solve_f64_noalloc: P = %d, batch = %d needs %zu bytes of device scratch, the "
                              "maximum is %zu", P, batch, off, (size_t)QGTL_NOALLOC_SCRATCH_BYTES);
    TRY(qgtl_init_if_needed());
    void *base = NULL;
    size_t cap = 0;
    TRY(qgtl_hip_noalloc_scratch(&base, &cap));
    if (off > cap) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_solve_f64_noalloc: scratch too small");
    char *s = (char *)base;
    double *dA = (double *)(s + oA), *db = (double *)(s + ob), *dx = (double *)(s + ox);
    double *dd = pre ? (double *)(s + od) : NULL, *dU = lr ? (double *)(s + oU) : NULL;
    int *dst = (int *)(s + ost), *dit = (int *)(s + oit);
    double *dres = (double *)(s + ores);
    const double t0 = wall_ms();
    TRY(qgtl_hip_h2d(dA, A, PP * sizeof(double)));
    TRY(qgtl_hip_h2d(db, b, Pb * sizeof(double)));
    if (method == QGTL_SOLVE_CHOLESKY) {
        TRY(qgtl_hip_cholesky_solve_f64(dA, db, dx, P, batch, lambda, dst));
    } else {
        if (method == QGTL_SOLVE_JACOBI_PCG) TRY(qgtl_hip_jacobi_diag_f64(dA, P, batch, lambda, dd));
        if (method == QGTL_SOLVE_PCG && d) TRY(qgtl_hip_h2d(dd, d, Pb * sizeof(double)));
        if (lr) TRY(qgtl_hip_h2d(dU, U, Pb * (size_t)rank * sizeof(double)));
        TRY(qgtl_hip_pcg_f64(dA, db, dx, P, batch, lambda, dd, dU, lr ? rank : 0, tol, maxit, dit, dres));
    }
    TRY(qgtl_hip_d2h(x, dx, Pb * sizeof(double)));
    if (method == QGTL_SOLVE_CHOLESKY) TRY(qgtl_hip_d2h(g_na_status, dst, (size_t)batch * sizeof(int)));
    else {
        TRY(qgtl_hip_d2h(g_na_iters, dit, (size_t)batch * sizeof(int)));
        TRY(qgtl_hip_d2h(g_na_res, dres, (size_t)batch * sizeof(double)));
    }
    TRY(qgtl_hip_sync());
    const float ms = (float)(wall_ms() - t0);
    int fail = 0;
    for (int k = 0; k < batch; k++) {
        qgtl_solve_info_t inf = {0, 0, 0.0, ms};
        if (method == QGTL_SOLVE_CHOLESKY) {
            inf.status = g_na_status[k];
            inf.iterations = 1;
        } else {
            inf.iterations = g_na_iters[k];
            inf.rel_residual = g_na_res[k];
            inf.status = (g_na_res[k] <= tol) ? 0 : -1;
        }
        if (inf.status) fail = 1;
        if (info) info[k] = inf;
    }
    if (fail) return qgtl_set_error(method == QGTL_SOLVE_CHOLESKY ? QGT_ERROR_MATRIX_SINGULAR : QGT_ERROR_CONVERGENCE_FAILED,
                                    method == QGTL_SOLVE_CHOLESKY ? "matrix not positive definite" : "CG did not reach the tolerance");
    return 0;
}

/* Hermitian eigen-decomposition on the GPU (complex128 in, ascending). */
int qgtl_hermitian_eig(int P, const double *A, double *evals, double *V)
{
    QGTL_LOCKED();
    if (P < 1 || P > 1024 || !A || !evals) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_hermitian_eig: bad arguments");
    TRY(qgtl_init_if_needed());
    const size_t PP = (size_t)P * P;
    void *dA, *dV, *dev, *dsw;
    TRY(scratch(S_A, PP * 16, &dA));
    TRY(scratch(S_V, PP * 16, &dV));
    TRY(scratch(S_EV, (size_t)P * 8, &dev));
    TRY(scratch(S_INT, sizeof(int), &dsw));
    TRY(qgtl_hip_h2d(dA, A, PP * 16));
    TRY(qgtl_hip_herm_eig_c128(dA, dV, (double *)dev, P, 1, 60, (int *)dsw));
    TRY(qgtl_hip_d2h(evals, dev, (size_t)P * 8));
    if (V) TRY(qgtl_hip_d2h(V, dV, PP * 16));
    TRY(qgtl_hip_sync());
    return 0;
}

/* ------------------------------------------------------------ QGTL API */
natural_gradient_config_t get_default_natural_gradient_config(void)
{
    natural_gradient_config_t c = {1e-4f, 1e8f, true, true, 1e-10f};
    return c;
}

static double *to_c128(const ComplexFloat *m, size_t count)
{
    double *o = (double *)malloc(count * 2 * sizeof(double));
    if (!o) return NULL;
    for (size_t i = 0; i < count; i++) { o[2 * i] = m[i].real; o[2 * i + 1] = m[i].imag; }
    return o;
}

static int is_hermitian(const double *M, size_t P)
{
    double mx = 0.0, dev = 0.0;
    for (size_t i = 0; i < P; i++)
        for (size_t j = 0; j < P; j++) {
            const double ar = M[2 * (i * P + j)], ai = M[2 * (i * P + j) + 1];
            const double br = M[2 * (j * P + i)], bi = M[2 * (j * P + i) + 1];
            mx = fmax(mx, hypot(ar, ai));
            dev = fmax(dev, hypot(ar - br, ai + bi));
        }
    return dev <= 1e-6 * (mx > 0 ? mx : 1.0);
}

/* Singular values of M (FP64 on the GPU): |eigenvalues| when Hermitian,
 * sqrt(eigenvalues of M^H M) otherwise.  Leaves the eigenvectors in S_V. */
static int singular_values(const double *M, size_t P, double *sv, int *hermitian)
{
    const size_t PP = P * P;
    void *dA, *dV, *dev, *dsw, *dB;
    TRY(scratch(S_A, PP * 16, &dA));
    TRY(scratch(S_V, PP * 16, &dV));
    TRY(scratch(S_EV, P * 8, &dev));
    TRY(scratch(S_INT, sizeof(int), &dsw));
    *hermitian = is_hermitian(M, P);
    if (*hermitian) {
        TRY(qgtl_hip_h2d(dA, M, PP * 16));
    } else {
        TRY(scratch(S_B, PP * 16, &dB));
        TRY(qgtl_hip_h2d(dB, M, PP * 16));
        TRY(qgtl_hip_gram_small_c128(dB, dA, (int)P));
    }
    TRY(qgtl_hip_herm_eig_c128(dA, dV, (double *)dev, (int)P, 1, 60, (int *)dsw));
    TRY(qgtl_hip_d2h(sv, dev, P * 8));
    TRY(qgtl_hip_sync());
    for (size_t i = 0; i < P; i++) sv[i] = *hermitian ? fabs(sv[i]) : sqrt(fmax(sv[i], 0.0));
    return 0;
}

bool compute_complex_matrix_condition_number(const ComplexFloat *matrix, size_t dimension, float *condition_number)
{
    QGTL_LOCKED();
    if (!matrix || !condition_number || dimension == 0 || dimension > 1024) return false;
    if (qgtl_init_if_needed()) return false;
    double *M = to_c128(matrix, dimension * dimension);
    double *sv = (double *)malloc(dimension * sizeof(double));
    int herm = 0;
    if (!M || !sv || singular_values(M, dimension, sv, &herm)) { free(M); free(sv); return false; }
    double smax = 0.0, smin = INFINITY;
    for (size_t i = 0; i < dimension; i++) {
        if (sv[i] > smax) smax = sv[i];
        if (sv[i] > 0.0 && sv[i] < smin) smin = sv[i];
    }
    *condition_number = (smin > 1e-15) ? (float)(smax / smin) : INFINITY;
    free(M);
    free(sv);
    return true;
}

/* x = M^+ g with singular values <= cutoff dropped. */
static int pinv_solve(const double *M, const double *g, double *x, size_t P, double cutoff)
{
    double *sv = (double *)malloc(P * sizeof(double));
    int herm = 0;
    if (!sv) return QGT_ERROR_NO_MEMORY;
    int rc = singular_values(M, P, sv, &herm);
    free(sv);
    if (rc) return rc;
    void *dV = g_scr[S_V], *dev = g_scr[S_EV], *dg, *dx, *dB;
    TRY(scratch(S_D, P * 16, &dg));
    TRY(scratch(S_X, P * 16, &dx));
    if (herm) {
        TRY(qgtl_hip_h2d(dg, g, P * 16));
        TRY(qgtl_hip_pinv_apply_c128(dV, (const double *)dev, dg, dx, (int)P, cutoff));
    } else {
        /* M^+ g = (M^H M)^+ M^H g, cutoff on sigma^2 */
        double *mhg = (double *)calloc(2 * P, sizeof(double));
        if (!mhg) return QGT_ERROR_NO_MEMORY;
        for (size_t i = 0; i < P; i++)
            for (size_t k = 0; k < P; k++) {
                const double mr = M[2 * (k * P + i)], mi = -M[2 * (k * P + i) + 1];
                mhg[2 * i] += mr * g[2 * k] - mi * g[2 * k + 1];
                mhg[2 * i + 1] += mr * g[2 * k + 1] + mi * g[2 * k];
            }
        (void)dB;
        rc = qgtl_hip_h2d(dg, mhg, P * 16);
        if (!rc) rc = qgtl_hip_pinv_apply_c128(dV, (const double *)dev, dg, dx, (int)P, cutoff * cutoff);
        if (!rc) rc = qgtl_hip_sync();
        free(mhg);
        TRY(rc);
    }
    TRY(qgtl_hip_d2h(x, dx, P * 16));
    TRY(qgtl_hip_sync());
    return 0;
}

bool compute_pseudoinverse_natural_gradient(const ComplexFloat *gradient, const ComplexFloat *metric,
                                            ComplexFloat *natural_gradient, size_t dimension, float singular_cutoff)
{
    QGTL_LOCKED();
    if (!gradient || !metric || !natural_gradient || dimension == 0 || dimension > 1024) return false;
    if (qgtl_init_if_needed()) return false;
    double *M = to_c128(metric, dimension * dimension), *g = to_c128(gradient, dimension);
    double *x = (double *)malloc(dimension * 16);
    const int rc = (M && g && x) ? pinv_solve(M, g, x, dimension, singular_cutoff) : QGT_ERROR_NO_MEMORY;
    if (rc == 0)
        for (size_t i = 0; i < dimension; i++) natural_gradient[i] = (ComplexFloat){(float)x[2 * i], (float)x[2 * i + 1]};
    free(M); free(g); free(x);
    return rc == 0;
}

bool compute_regularized_natural_gradient(const ComplexFloat *gradient, const ComplexFloat *metric,
                                          ComplexFloat *natural_gradient, size_t dimension,
                                          const natural_gradient_config_t *config)
{
    QGTL_LOCKED();
    if (!gradient || !metric || !natural_gradient || !config || dimension == 0 || dimension > 1024) return false;
    if (qgtl_init_if_needed()) return false;
    double lambda = config->regularization_param;
    if (config->use_adaptive_regularization) {
        float kappa;
        if (compute_complex_matrix_condition_number(metric, dimension, &kappa) && isfinite(kappa) &&
            kappa > config->condition_threshold) {
            /* quantum_geometric_gradient.c:2901-2914; an exactly singular metric
             * (kappa = inf) keeps lambda instead of QGTL's lambda = inf */
            const float al = 1e-6f * sqrtf(kappa);
            if (al > lambda) lambda = al;
        }
    }
    const size_t P = dimension, PP = P * P;
    double *A = to_c128(metric, PP), *g = to_c128(gradient, P), *x = (double *)malloc(P * 16);
    int ok = 0;
    if (A && g && x) {
        for (size_t i = 0; i < P; i++) A[2 * (i * P + i)] += lambda;
        void *dA, *db, *dx, *dpiv, *dst;
        int st = -1;
        if (!scratch(S_A, PP * 16, &dA) && !scratch(S_B, P * 16, &db) && !scratch(S_X, P * 16, &dx) &&
            !scratch(S_U, P * sizeof(int), &dpiv) && !scratch(S_INT, sizeof(int), &dst) &&
            !qgtl_hip_h2d(dA, A, PP * 16) && !qgtl_hip_h2d(db, g, P * 16) &&
            !qgtl_hip_lu_solve_c128(dA, db, dx, (int *)dpiv, (int)P, 1, (int *)dst) &&
            !qgtl_hip_d2h(&st, dst, sizeof(int)) && !qgtl_hip_d2h(x, dx, P * 16) && !qgtl_hip_sync()) {
            if (st == 0) ok = 1;
        }
        if (!ok && config->use_pseudoinverse_fallback) {
            double *M = to_c128(metric, PP);
            ok = M && pinv_solve(M, g, x, P, config->singular_value_cutoff) == 0;
            free(M);
        }
        if (ok)
            for (size_t i = 0; i < P; i++) natural_gradient[i] = (ComplexFloat){(float)x[2 * i], (float)x[2 * i + 1]};
    }
    free(A); free(g); free(x);
    return ok;
}
