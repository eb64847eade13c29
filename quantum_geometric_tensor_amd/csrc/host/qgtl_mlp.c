/*
 * qgtl_mlp.c -- the learned preconditioner: a small MLP that maps the
 * features of one parameter point (sin/cos of the angles, log of the
 * metric diagonal, the Hamiltonian field) to the diagonal and low-rank
 * factor of a preconditioner for (Re Q + lambda I) x = grad E, trained on
 * the GPU by the kernels of qgtl_mlp.hip with Adam on Kaporin's condition
 * number of the preconditioned matrix.
 *
 * Network: nh hidden layers of width H with the tanh-form GELU, linear
 * output y of size P (1 + r).  Weights W_i are (in x out) row-major, each
 * followed by its bias, in one FP32 master vector (Adam) with a BF16
 * mirror that the GEMMs read.  With D = diag(A) and As = D^-1/2 A D^-1/2:
 *   s_k = y[k (1 + r)], u_kj = y[k (1 + r) + 1 + j] / sqrt(P),
 *   M^-1 = diag(e^s / D) + (D^-1/2 u)(D^-1/2 u)^T.
 * The Jacobi preconditioner is the special case s = 0, u = 0, which is
 * where a small output layer starts training.
 *
 * QGTL has no learned preconditioner; the solve it accelerates is
 * compute_natural_gradient_cpu's conjugate gradient
 * (src/quantum_geometric/hybrid/classical_optimization_engine.c:543-610).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_internal.h"

#define MTRY(call)                                                                              \
    do {                                                                                        \
        int _rc = (call);                                                                       \
        if (_rc) return qgtl_set_error(_rc, "%s: %s", #call, qgtl_hip_last_error());            \
    } while (0)

#define MLP_MAX_LAYERS 8

struct qgtl_mlp {
    int F, H, nh, P, r, O;
    size_t nparams;
    size_t off[MLP_MAX_LAYERS + 1];      /* parameter offset of layer i */
    int in[MLP_MAX_LAYERS + 1], out[MLP_MAX_LAYERS + 1];
    float *w, *m, *v, *g;                /* device: master, Adam moments, gradient */
    uint16_t *wb;                        /* device: bf16 mirror */
    int step;
    /* activations for a batch of up to cap rows */
    int cap;
    float *x;                            /* cap x F */
    float *pre[MLP_MAX_LAYERS + 1];      /* cap x out_i */
    float *act[MLP_MAX_LAYERS + 1];
    float *dact[MLP_MAX_LAYERS + 1];     /* gradient w.r.t. act (reused as dz) */
    double *loss;                        /* cap */
    int *idx;                            /* cap */
    /* training data on the device */
    int S;
    float *X_all, *As_all;
};

static uint64_t splitmix(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void free_batch(struct qgtl_mlp *m)
{
    qgtl_hip_free(m->x);
    m->x = NULL;
    for (int i = 0; i <= m->nh; i++) {
        qgtl_hip_free(m->pre[i]);
        qgtl_hip_free(m->act[i]);
        qgtl_hip_free(m->dact[i]);
        m->pre[i] = m->act[i] = m->dact[i] = NULL;
    }
    qgtl_hip_free(m->loss);
    qgtl_hip_free(m->idx);
    m->loss = NULL;
    m->idx = NULL;
    m->cap = 0;
}

static int ensure_batch(struct qgtl_mlp *m, int B)
{
    if (B <= m->cap) return 0;
    free_batch(m);
    MTRY(qgtl_hip_malloc((void **)&m->x, sizeof(float) * (size_t)B * m->F));
    for (int i = 0; i <= m->nh; i++) {
        const size_t n = sizeof(float) * (size_t)B * m->out[i];
        MTRY(qgtl_hip_malloc((void **)&m->pre[i], n));
        MTRY(qgtl_hip_malloc((void **)&m->act[i], n));
        MTRY(qgtl_hip_malloc((void **)&m->dact[i], n));
    }
    MTRY(qgtl_hip_malloc((void **)&m->loss, sizeof(double) * (size_t)B));
    MTRY(qgtl_hip_malloc((void **)&m->idx, sizeof(int) * (size_t)B));
    m->cap = B;
    return 0;
}

void qgtl_mlp_destroy(qgtl_mlp_t *m)
{
    QGTL_LOCKED();
    if (!m) return;
    free_batch(m);
    qgtl_hip_free(m->w);
    qgtl_hip_free(m->m);
    qgtl_hip_free(m->v);
    qgtl_hip_free(m->g);
    qgtl_hip_free(m->wb);
    qgtl_hip_free(m->X_all);
    qgtl_hip_free(m->As_all);
    free(m);
}

qgtl_mlp_t *qgtl_mlp_create(int in_dim, int hidden, int hidden_layers, int P, int rank, uint64_t seed)
{
    QGTL_LOCKED();
    if (in_dim < 1 || hidden < 1 || hidden_layers < 1 || hidden_layers >= MLP_MAX_LAYERS || P < 1 || P > 1024 ||
        rank < 0 || rank > 32 || P * rank > 8192) {
        qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT,
                       "mlp: bad shape (hidden layers 1..7, P <= 1024, rank <= 32, P x rank <= 8192)");
        return NULL;
    }
    if (qgtl_init_if_needed()) return NULL;
    struct qgtl_mlp *m = (struct qgtl_mlp *)calloc(1, sizeof *m);
    if (!m) { qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); return NULL; }
    m->F = in_dim; m->H = hidden; m->nh = hidden_layers; m->P = P; m->r = rank; m->O = P * (1 + rank);
    size_t o = 0;
    for (int i = 0; i <= m->nh; i++) {
        m->in[i] = (i == 0) ? m->F : m->H;
        m->out[i] = (i == m->nh) ? m->O : m->H;
        m->off[i] = o;
        o += (size_t)m->in[i] * m->out[i] + m->out[i];
    }
    m->nparams = o;
    /* uniform(+-sqrt(3/in)) weights (variance 1/in), output layer x0.05, zero biases */
    float *h = (float *)calloc(o, sizeof(float));
    if (!h) { free(m); qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); return NULL; }
    uint64_t st = seed;
    for (int i = 0; i <= m->nh; i++) {
        const double a = sqrt(3.0 / m->in[i]) * (i == m->nh ? 0.05 : 1.0);
        const size_t nw = (size_t)m->in[i] * m->out[i];
        for (size_t e = 0; e < nw; e++) {
            const double u = (splitmix(&st) >> 11) * (1.0 / 9007199254740992.0);
            h[m->off[i] + e] = (float)((2.0 * u - 1.0) * a);
        }
    }
    int rc = qgtl_hip_malloc((void **)&m->w, o * sizeof(float));
    if (!rc) rc = qgtl_hip_malloc((void **)&m->m, o * sizeof(float));
    if (!rc) rc = qgtl_hip_malloc((void **)&m->v, o * sizeof(float));
    if (!rc) rc = qgtl_hip_malloc((void **)&m->g, o * sizeof(float));
    if (!rc) rc = qgtl_hip_malloc((void **)&m->wb, o * sizeof(uint16_t));
    if (!rc) rc = qgtl_mlp_set_params(m, h);
    free(h);
    if (rc) {
        qgtl_set_error(rc, "mlp create: %s", qgtl_hip_last_error());
        qgtl_mlp_destroy(m);
        return NULL;
    }
    return m;
}

size_t qgtl_mlp_num_params(const qgtl_mlp_t *m) { return m ? m->nparams : 0; }

void qgtl_mlp_shape(const qgtl_mlp_t *m, int *in_dim, int *P, int *rank)
{
    *in_dim = m->F;
    *P = m->P;
    *rank = m->r;
}

int qgtl_mlp_set_params(qgtl_mlp_t *m, const float *params)
{
    QGTL_LOCKED();
    if (!m || !params) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null mlp or params");
    MTRY(qgtl_hip_h2d(m->w, params, m->nparams * sizeof(float)));
    MTRY(qgtl_hip_fill_f32(m->m, m->nparams, 0.f));
    MTRY(qgtl_hip_fill_f32(m->v, m->nparams, 0.f));
    MTRY(qgtl_hip_to_bf16(m->w, m->wb, m->nparams));
    MTRY(qgtl_hip_sync());
    m->step = 0;
    return 0;
}

int qgtl_mlp_get_params(const qgtl_mlp_t *m, float *params)
{
    QGTL_LOCKED();
    if (!m || !params) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null mlp or params");
    MTRY(qgtl_hip_d2h(params, m->w, m->nparams * sizeof(float)));
    MTRY(qgtl_hip_sync());
    return 0;
}

int qgtl_mlp_set_data(qgtl_mlp_t *m, int S, const float *X, const float *As)
{
    QGTL_LOCKED();
    if (!m || S < 1 || !X || !As) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "empty training set");
    qgtl_hip_free(m->X_all);
    qgtl_hip_free(m->As_all);
    m->X_all = m->As_all = NULL;
    m->S = 0;
    MTRY(qgtl_hip_malloc((void **)&m->X_all, sizeof(float) * (size_t)S * m->F));
    MTRY(qgtl_hip_malloc((void **)&m->As_all, sizeof(float) * (size_t)S * m->P * m->P));
    MTRY(qgtl_hip_h2d(m->X_all, X, sizeof(float) * (size_t)S * m->F));
    MTRY(qgtl_hip_h2d(m->As_all, As, sizeof(float) * (size_t)S * m->P * m->P));
    MTRY(qgtl_hip_sync());
    m->S = S;
    return 0;
}

/* forward on the B rows already in m->x */
static int forward_dev(struct qgtl_mlp *m, int B)
{
    const float *in = m->x;
    for (int i = 0; i <= m->nh; i++) {
        const uint16_t *W = m->wb + m->off[i];
        const float *b = m->w + m->off[i] + (size_t)m->in[i] * m->out[i];
        MTRY(qgtl_hip_mlp_gemm(B, m->out[i], m->in[i], in, 0, 0, m->in[i], W, 1, 0, m->out[i], m->pre[i],
                               m->out[i], b, 0));
        if (i < m->nh) MTRY(qgtl_hip_gelu_fwd(m->pre[i], m->act[i], (size_t)B * m->out[i]));
        in = (i < m->nh) ? m->act[i] : m->pre[i];
    }
    return 0;
}

/* loss (mean over B) and gradient into m->g for the dataset rows idx */
static int loss_grad_dev(struct qgtl_mlp *m, int B, const int *idx, double *loss_mean)
{
    if (B < 1 || !idx) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "empty batch");
    if (m->S < 1) return qgtl_set_error(QGT_ERROR_INVALID_STATE, "no training data (qgtl_mlp_set_data)");
    for (int b = 0; b < B; b++)
        if (idx[b] < 0 || idx[b] >= m->S) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "sample %d out of range", idx[b]);
    int rc = ensure_batch(m, B);
    if (rc) return rc;
    MTRY(qgtl_hip_h2d(m->idx, idx, sizeof(int) * (size_t)B));
    MTRY(qgtl_hip_gather_rows(m->X_all, m->idx, B, m->F, m->x));
    rc = forward_dev(m, B);
    if (rc) return rc;
    const int L = m->nh;
    /* output layer is linear: y = pre[L]; dact[L] = dL/dy / B */
    MTRY(qgtl_hip_kaporin(m->P, m->r, B, m->pre[L], m->As_all, m->idx, m->loss, m->dact[L], 1.0f / (float)B));
    for (int i = L; i >= 0; i--) {
        float *dz = m->dact[i];
        if (i < L) MTRY(qgtl_hip_gelu_bwd(m->pre[i], dz, (size_t)B * m->out[i]));
        const float *in = (i == 0) ? m->x : m->act[i - 1];
        float *gW = m->g + m->off[i], *gb = gW + (size_t)m->in[i] * m->out[i];
        /* gW (in x out) = in^T (in x B) dz (B x out) */
        MTRY(qgtl_hip_mlp_gemm(m->in[i], m->out[i], B, in, 0, 1, m->in[i], dz, 0, 0, m->out[i], gW, m->out[i], NULL, 0));
        MTRY(qgtl_hip_colsum(dz, B, m->out[i], gb));
        if (i > 0)   /* dact[i-1] (B x in) = dz (B x out) W^T */
            MTRY(qgtl_hip_mlp_gemm(B, m->in[i], m->out[i], dz, 0, 0, m->out[i], m->wb + m->off[i], 1, 1, m->out[i],
                                   m->dact[i - 1], m->in[i], NULL, 0));
    }
    if (loss_mean) {
        double *h = (double *)malloc(sizeof(double) * (size_t)B);
        if (!h) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        rc = qgtl_hip_d2h(h, m->loss, sizeof(double) * (size_t)B);
        if (!rc) rc = qgtl_hip_sync();
        double s = 0.0;
        for (int b = 0; b < B; b++) s += h[b];
        free(h);
        if (rc) return qgtl_set_error(rc, "mlp loss: %s", qgtl_hip_last_error());
        *loss_mean = s / B;
    }
    return 0;
}

int qgtl_mlp_loss_grad(qgtl_mlp_t *m, int B, const int *idx, double *loss, float *grad)
{
    QGTL_LOCKED();
    if (!m) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null mlp");
    int rc = loss_grad_dev(m, B, idx, loss);
    if (rc) return rc;
    if (grad) {
        MTRY(qgtl_hip_d2h(grad, m->g, m->nparams * sizeof(float)));
        MTRY(qgtl_hip_sync());
    }
    return 0;
}

int qgtl_mlp_train_step(qgtl_mlp_t *m, int B, const int *idx, float lr, double *loss)
{
    QGTL_LOCKED();
    if (!m) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null mlp");
    int rc = loss_grad_dev(m, B, idx, loss);
    if (rc) return rc;
    m->step++;
    MTRY(qgtl_hip_adam(m->w, m->wb, m->m, m->v, m->g, m->nparams, lr, 0.9f, 0.999f, 1e-8f, m->step));
    if (!loss) MTRY(qgtl_hip_sync());
    return 0;
}

int qgtl_mlp_device_input(qgtl_mlp_t *m, float **x_dev)
{
    int rc = ensure_batch(m, 1);
    if (rc) return rc;
    *x_dev = m->x;
    return 0;
}

int qgtl_mlp_forward_device(qgtl_mlp_t *m, const float **y_dev)
{
    int rc = forward_dev(m, 1);
    if (rc) return rc;
    *y_dev = m->pre[m->nh];
    return 0;
}

int qgtl_mlp_forward(qgtl_mlp_t *m, int B, const float *X, float *Y)
{
    QGTL_LOCKED();
    if (!m || B < 1 || !X || !Y) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "bad forward arguments");
    int rc = ensure_batch(m, B);
    if (rc) return rc;
    MTRY(qgtl_hip_h2d(m->x, X, sizeof(float) * (size_t)B * m->F));
    rc = forward_dev(m, B);
    if (rc) return rc;
    MTRY(qgtl_hip_d2h(Y, m->pre[m->nh], sizeof(float) * (size_t)B * m->O));
    MTRY(qgtl_hip_sync());
    return 0;
}

/* Features of one parameter point: sin theta (P), cos theta (P), log D (P), field. */
int qgtl_pc_features(int P, const double *theta, const double *D, double field, float *x)
{
    if (P < 1 || !theta || !D || !x) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "bad feature arguments");
    for (int k = 0; k < P; k++) {
        if (!(D[k] > 0.0)) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "metric diagonal must be positive");
        x[k] = (float)sin(theta[k]);
        x[P + k] = (float)cos(theta[k]);
        x[2 * P + k] = (float)log(D[k]);
    }
    x[3 * P] = (float)field;
    return 0;
}

int qgtl_mlp_preconditioner(qgtl_mlp_t *m, const float *x, const double *D, double *d, double *U)
{
    QGTL_LOCKED();
    if (!m || !x || !D || !d) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "bad preconditioner arguments");
    float *y = (float *)malloc(sizeof(float) * (size_t)m->O);
    if (!y) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
    int rc = qgtl_mlp_forward(m, 1, x, y);
    if (rc) { free(y); return rc; }
    const int st = 1 + m->r;
    const double us = 1.0 / sqrt((double)m->P);
    for (int k = 0; k < m->P; k++) {
        d[k] = exp((double)y[(size_t)k * st]) / D[k];
        if (U)
            for (int j = 0; j < m->r; j++) U[(size_t)k * m->r + j] = (double)y[(size_t)k * st + 1 + j] * us / sqrt(D[k]);
    }
    free(y);
    return 0;
}

/* ---- checkpoints (container of qgtl_io.c, magic "QGTLMLP1") ----
 * kind header: shape, Adam step, whether the Adam moments follow;
 * payload: fp32 master weights [+ first and second moments]. */
typedef struct {
    int32_t F, H, nh, P, r, step, has_opt, reserved;
} mlp_hdr_t;

int qgtl_mlp_save(const qgtl_mlp_t *m, const char *path, int with_optimizer)
{
    QGTL_LOCKED();
    if (!m || !path) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_mlp_save: null argument");
    const size_t nb = m->nparams * sizeof(float);
    const int k = with_optimizer ? 3 : 1;
    float *h = (float *)malloc(nb * (size_t)k);
    if (!h) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
    int rc = 0;
    if (qgtl_hip_d2h(h, m->w, nb) || (k == 3 && (qgtl_hip_d2h(h + m->nparams, m->m, nb) ||
                                                  qgtl_hip_d2h(h + 2 * m->nparams, m->v, nb))) ||
        qgtl_hip_sync()) {
        rc = qgtl_set_error(QGT_ERROR_GPU_TRANSFER_FAILED, "mlp checkpoint: %s", qgtl_hip_last_error());
    } else {
        const mlp_hdr_t hd = {m->F, m->H, m->nh, m->P, m->r, m->step, k == 3, 0};
        const void *pieces[1] = {h};
        const size_t sizes[1] = {nb * (size_t)k};
        rc = qgtl_io_write(path, "QGTLMLP1", &hd, sizeof hd, 1, pieces, sizes);
    }
    free(h);
    return rc;
}

qgtl_mlp_t *qgtl_mlp_load(const char *path)
{
    QGTL_LOCKED();
    mlp_hdr_t hd;
    uint64_t pb = 0, ck = 0;
    FILE *f = NULL;
    if (qgtl_io_open(path, "QGTLMLP1", &hd, sizeof hd, &pb, &ck, &f)) return NULL;
    /* the payload size the header implies, before anything is allocated for it */
    uint64_t np = 0;
    const int ok_shape = hd.F >= 1 && hd.F <= (1 << 20) && hd.H >= 1 && hd.H <= 65536 && hd.nh >= 1 &&
                         hd.nh < MLP_MAX_LAYERS && hd.P >= 1 && hd.P <= 1024 && hd.r >= 0 && hd.r <= 16;
    if (ok_shape)
        for (int i = 0; i <= hd.nh; i++) {
            const uint64_t in = (i == 0) ? (uint64_t)hd.F : (uint64_t)hd.H;
            const uint64_t out = (i == hd.nh) ? (uint64_t)hd.P * (1u + (uint64_t)hd.r) : (uint64_t)hd.H;
            np += in * out + out;
        }
    if (!ok_shape || pb != np * sizeof(float) * (hd.has_opt ? 3u : 1u)) {
        fclose(f);
        qgtl_set_error(QGT_ERROR_VALIDATION_FAILED, "%s: inconsistent checkpoint header", path);
        return NULL;
    }
    qgtl_mlp_t *m = qgtl_mlp_create(hd.F, hd.H, hd.nh, hd.P, hd.r, 0);
    if (!m) { fclose(f); return NULL; }
    const size_t nb = m->nparams * sizeof(float);
    const int k = hd.has_opt ? 3 : 1;
    float *h = (float *)malloc(nb * (size_t)k);
    if (!h) {
        fclose(f);
        qgtl_mlp_destroy(m);
        qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        return NULL;
    }
    void *pieces[1] = {h};
    const size_t sizes[1] = {nb * (size_t)k};
    int rc = qgtl_io_read_payload(f, path, &hd, sizeof hd, pb, ck, 1, pieces, sizes);
    if (!rc) rc = qgtl_mlp_set_params(m, h);       /* also refreshes the bf16 mirror, zeroes Adam */
    if (!rc && k == 3) {
        if (qgtl_hip_h2d(m->m, h + m->nparams, nb) || qgtl_hip_h2d(m->v, h + 2 * m->nparams, nb) || qgtl_hip_sync())
            rc = qgtl_set_error(QGT_ERROR_GPU_TRANSFER_FAILED, "mlp checkpoint: %s", qgtl_hip_last_error());
        else
            m->step = hd.step;
    }
    free(h);
    if (rc) { qgtl_mlp_destroy(m); return NULL; }
    return m;
}

int qgtl_mlp_get_shape(const qgtl_mlp_t *m, int *in_dim, int *hidden, int *hidden_layers, int *P, int *rank)
{
    if (!m) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null mlp");
    if (in_dim) *in_dim = m->F;
    if (hidden) *hidden = m->H;
    if (hidden_layers) *hidden_layers = m->nh;
    if (P) *P = m->P;
    if (rank) *rank = m->r;
    return 0;
}
