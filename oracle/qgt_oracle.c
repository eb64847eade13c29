/*
 * qgt_oracle.c -- FP64 CPU restatement of QGTL's QGT path.
 * TEST INFRASTRUCTURE ONLY (see qgt_oracle.h for scope and citations).
 * Built with -O3, no -ffast-math.  OpenMP only splits independent outputs
 * (amplitude pairs, matrix rows); every individual sum runs sequentially,
 * so results do not depend on the thread count.
 */
#include "qgt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OMP_MIN_DIM (1u << 14)

/* ------------------------------------------------------------------ RNG */
uint64_t oracle_splitmix64_next(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double oracle_uniform01(uint64_t *state)
{
    return (double)(oracle_splitmix64_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

void oracle_uniform_angles(uint64_t seed, size_t count, double *theta)
{
    uint64_t s = seed;
    for (size_t i = 0; i < count; i++)
        theta[i] = 2.0 * M_PI * oracle_uniform01(&s);
}

/* ------------------------------------------------------------------ HEA */
int oracle_hea_num_ops(int n, int layers)
{
    int ring = (n >= 2) ? n : 0;
    return layers * (2 * n + ring);
}

int oracle_hea_ops(int n, int layers, const double *theta, oracle_op_t *ops)
{
    int k = 0;
    for (int l = 0; l < layers; l++) {
        for (int q = 0; q < n; q++) {
            int p = 2 * n * l + q;
            ops[k++] = (oracle_op_t){ORC_RY, q, -1, p, theta[p]};
        }
        for (int q = 0; q < n; q++) {
            int p = 2 * n * l + n + q;
            ops[k++] = (oracle_op_t){ORC_RZ, q, -1, p, theta[p]};
        }
        if (n >= 2)
            for (int q = 0; q < n; q++)
                ops[k++] = (oracle_op_t){ORC_CNOT, q, (q + 1) % n, -1, 0.0};
    }
    return k;
}

/* ---------------------------------------------------------------- gates */
int oracle_num_params(int nops, const oracle_op_t *ops)
{
    int P = 0;
    for (int i = 0; i < nops; i++)
        if (ops[i].param >= 0) P++;
    return P;
}

static void set_m(double *m, double a, double b, double c, double d,
                  double e, double f, double g, double h)
{
    m[0] = a; m[1] = b; m[2] = c; m[3] = d; m[4] = e; m[5] = f; m[6] = g; m[7] = h;
}

/* Row-major 2x2 (m00, m01, m10, m11), each complex as (re, im).
 * Conventions: quantum_gate_operations.c:11-47, quantum_classical_algorithms.c:228-265. */
int oracle_gate_matrix(const oracle_op_t *op, double *m)
{
    const double r = M_SQRT1_2;
    double c = cos(op->angle / 2.0), s = sin(op->angle / 2.0);
    switch (op->type) {
    case ORC_I:    set_m(m, 1, 0, 0, 0, 0, 0, 1, 0); return 0;
    case ORC_X:    set_m(m, 0, 0, 1, 0, 1, 0, 0, 0); return 0;
    case ORC_Y:    set_m(m, 0, 0, 0, -1, 0, 1, 0, 0); return 0;
    case ORC_Z:    set_m(m, 1, 0, 0, 0, 0, 0, -1, 0); return 0;
    case ORC_H:    set_m(m, r, 0, r, 0, r, 0, -r, 0); return 0;
    case ORC_S:    set_m(m, 1, 0, 0, 0, 0, 0, 0, 1); return 0;
    case ORC_SDG:  set_m(m, 1, 0, 0, 0, 0, 0, 0, -1); return 0;
    case ORC_T:    set_m(m, 1, 0, 0, 0, 0, 0, r, r); return 0;
    case ORC_TDG:  set_m(m, 1, 0, 0, 0, 0, 0, r, -r); return 0;
    case ORC_SX:   set_m(m, 0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5); return 0;
    case ORC_RX:   set_m(m, c, 0, 0, -s, 0, -s, c, 0); return 0;
    case ORC_RY:   set_m(m, c, 0, -s, 0, s, 0, c, 0); return 0;
    case ORC_RZ:   set_m(m, c, -s, 0, 0, 0, 0, c, s); return 0;
    case ORC_U1:
    case ORC_PHASE: set_m(m, 1, 0, 0, 0, 0, 0, cos(op->angle), sin(op->angle)); return 0;
    default: return -1;
    }
}

void oracle_apply_1q(int n, double *psi, int q, const double *m)
{
    const size_t half = (size_t)1 << (n - 1);
    const size_t lo_mask = ((size_t)1 << q) - 1;
    const size_t bit = (size_t)1 << q;
#pragma omp parallel for if (half >= OMP_MIN_DIM) schedule(static)
    for (size_t t = 0; t < half; t++) {
        size_t i = ((t & ~lo_mask) << 1) | (t & lo_mask);
        size_t j = i | bit;
        double ar = psi[2 * i], ai = psi[2 * i + 1];
        double br = psi[2 * j], bi = psi[2 * j + 1];
        psi[2 * i]     = m[0] * ar - m[1] * ai + m[2] * br - m[3] * bi;
        psi[2 * i + 1] = m[0] * ai + m[1] * ar + m[2] * bi + m[3] * br;
        psi[2 * j]     = m[4] * ar - m[5] * ai + m[6] * br - m[7] * bi;
        psi[2 * j + 1] = m[4] * ai + m[5] * ar + m[6] * bi + m[7] * br;
    }
}

/* Controlled single-qubit matrix: applies m to target where control bit is 1. */
static void apply_controlled(int n, double *psi, int c, int t, const double *m)
{
    const size_t dim = (size_t)1 << n;
    const size_t cb = (size_t)1 << c, tb = (size_t)1 << t;
#pragma omp parallel for if (dim >= 2 * OMP_MIN_DIM) schedule(static)
    for (size_t i = 0; i < dim; i++) {
        if (!(i & cb) || (i & tb)) continue;
        size_t j = i | tb;
        double ar = psi[2 * i], ai = psi[2 * i + 1];
        double br = psi[2 * j], bi = psi[2 * j + 1];
        psi[2 * i]     = m[0] * ar - m[1] * ai + m[2] * br - m[3] * bi;
        psi[2 * i + 1] = m[0] * ai + m[1] * ar + m[2] * bi + m[3] * br;
        psi[2 * j]     = m[4] * ar - m[5] * ai + m[6] * br - m[7] * bi;
        psi[2 * j + 1] = m[4] * ai + m[5] * ar + m[6] * bi + m[7] * br;
    }
}

int oracle_apply_op(int n, double *psi, const oracle_op_t *op)
{
    double m[8];
    const size_t dim = (size_t)1 << n;
    if (op->q0 < 0 || op->q0 >= n) return -1;
    switch (op->type) {
    case ORC_CNOT:
    case ORC_CY:
    case ORC_CZ: {
        if (op->q1 < 0 || op->q1 >= n || op->q1 == op->q0) return -1;
        oracle_op_t base = {op->type == ORC_CNOT ? ORC_X : op->type == ORC_CY ? ORC_Y : ORC_Z,
                            op->q1, -1, -1, 0.0};
        oracle_gate_matrix(&base, m);
        apply_controlled(n, psi, op->q0, op->q1, m);
        return 0;
    }
    case ORC_SWAP: {
        if (op->q1 < 0 || op->q1 >= n || op->q1 == op->q0) return -1;
        const size_t ab = (size_t)1 << op->q0, bb = (size_t)1 << op->q1;
        for (size_t i = 0; i < dim; i++) {
            if ((i & ab) && !(i & bb)) {
                size_t j = (i & ~ab) | bb;
                double tr = psi[2 * i], ti = psi[2 * i + 1];
                psi[2 * i] = psi[2 * j]; psi[2 * i + 1] = psi[2 * j + 1];
                psi[2 * j] = tr; psi[2 * j + 1] = ti;
            }
        }
        return 0;
    }
    default:
        if (oracle_gate_matrix(op, m) != 0) return -1;
        oracle_apply_1q(n, psi, op->q0, m);
        return 0;
    }
}

int oracle_run(int n, int nops, const oracle_op_t *ops, double *psi)
{
    const size_t dim = (size_t)1 << n;
    memset(psi, 0, 2 * dim * sizeof(double));
    psi[0] = 1.0;
    for (int g = 0; g < nops; g++)
        if (oracle_apply_op(n, psi, &ops[g]) != 0) return -1;
    return 0;
}

/* (-i/2) G for the rotation generator G in {X, Y, Z}. */
static int generator_matrix(int type, double *m)
{
    switch (type) {
    case ORC_RX: set_m(m, 0, 0, 0, -0.5, 0, -0.5, 0, 0); return 0;   /* -i/2 X */
    case ORC_RY: set_m(m, 0, 0, -0.5, 0, 0.5, 0, 0, 0); return 0;    /* -i/2 Y */
    case ORC_RZ: set_m(m, 0, -0.5, 0, 0, 0, 0, 0, 0.5); return 0;    /* -i/2 Z */
    default: return -1;
    }
}

/* Forward sweep: at each selected parameterised gate the current state is
 * copied, (-i/2)G is applied, and the copy then rides along with every later
 * gate.  want[k] selects parameter k; slot[k] is its output row. */
static int derivative_sweep(int n, int nops, const oracle_op_t *ops, const int *slot,
                            double *psi, double *D)
{
    const size_t dim = (size_t)1 << n;
    int P = oracle_num_params(nops, ops);
    int *active = (int *)malloc(sizeof(int) * (P > 0 ? P : 1));
    int nact = 0;
    if (!active) return -1;
    memset(psi, 0, 2 * dim * sizeof(double));
    psi[0] = 1.0;
    /* progress lines on stderr for long sweeps (2^24 amplitudes and up) */
    const int progress = n >= 24 && getenv("QGT_ORACLE_PROGRESS") != NULL;
    for (int g = 0; g < nops; g++) {
        if (progress && g % 32 == 0) {
            fprintf(stderr, "oracle derivative sweep: gate %d of %d, %d states\n", g, nops, nact + 1);
            fflush(stderr);
        }
        for (int a = 0; a < nact; a++)
            if (oracle_apply_op(n, D + (size_t)2 * dim * active[a], &ops[g]) != 0) { free(active); return -1; }
        if (oracle_apply_op(n, psi, &ops[g]) != 0) { free(active); return -1; }
        int k = ops[g].param;
        if (k >= 0 && k < P && slot[k] >= 0) {
            double m[8];
            if (generator_matrix(ops[g].type, m) != 0) { free(active); return -1; }
            double *d = D + (size_t)2 * dim * slot[k];
            memcpy(d, psi, 2 * dim * sizeof(double));
            oracle_apply_1q(n, d, ops[g].q0, m);
            active[nact++] = slot[k];
        }
    }
    free(active);
    return 0;
}

int oracle_derivative_states(int n, int nops, const oracle_op_t *ops, double *psi, double *D)
{
    int P = oracle_num_params(nops, ops);
    int *slot = (int *)malloc(sizeof(int) * (P > 0 ? P : 1));
    if (!slot) return -1;
    for (int k = 0; k < P; k++) slot[k] = k;
    int rc = derivative_sweep(n, nops, ops, slot, psi, D);
    free(slot);
    return rc;
}

int oracle_derivative_subset(int n, int nops, const oracle_op_t *ops, int count,
                             const int *params, double *psi, double *D)
{
    int P = oracle_num_params(nops, ops);
    int *slot = (int *)malloc(sizeof(int) * (P > 0 ? P : 1));
    if (!slot) return -1;
    for (int k = 0; k < P; k++) slot[k] = -1;
    for (int i = 0; i < count; i++) {
        if (params[i] < 0 || params[i] >= P) { free(slot); return -1; }
        slot[params[i]] = i;
    }
    int rc = derivative_sweep(n, nops, ops, slot, psi, D);
    free(slot);
    return rc;
}

int oracle_derivative_states_shift(int n, int nops, const oracle_op_t *ops, double *D)
{
    const size_t dim = (size_t)1 << n;
    int P = oracle_num_params(nops, ops);
    oracle_op_t *shifted = (oracle_op_t *)malloc(sizeof(oracle_op_t) * (nops > 0 ? nops : 1));
    if (!shifted) return -1;
    for (int k = 0; k < P; k++) {
        memcpy(shifted, ops, sizeof(oracle_op_t) * nops);
        for (int g = 0; g < nops; g++)
            if (shifted[g].param == k) shifted[g].angle += M_PI;
        double *d = D + (size_t)2 * dim * k;
        if (oracle_run(n, nops, shifted, d) != 0) { free(shifted); return -1; }
        for (size_t i = 0; i < 2 * dim; i++) d[i] *= 0.5;
    }
    free(shifted);
    return 0;
}

/* ------------------------------------------------------------------ QGT */
/* <a|b> = sum conj(a_i) b_i, sequential FP64 sum. */
static void dotc(size_t dim, const double *a, const double *b, double *re, double *im)
{
    double sr = 0.0, si = 0.0;
    for (size_t i = 0; i < dim; i++) {
        double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        sr += ar * br + ai * bi;
        si += ar * bi - ai * br;
    }
    *re = sr;
    *im = si;
}

/* differential_geometry.c:2832-2858: every (i, j), three sums per element;
 * rows [r0, r1) of the loop on `threads` OpenMP threads (0: the default). */
static void qgt_pairwise_rows(size_t dim, int P, const double *psi, const double *D, double *Q, int r0, int r1,
                              int threads)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int i = r0; i < r1; i++) {
        const double *di = D + 2 * dim * (size_t)i;
        for (int j = 0; j < P; j++) {
            const double *dj = D + 2 * dim * (size_t)j;
            double gr, gi, air, aii, bjr, bji;
            dotc(dim, di, dj, &gr, &gi);          /* <d_i|d_j> */
            dotc(dim, di, psi, &air, &aii);       /* <d_i|psi> */
            dotc(dim, psi, dj, &bjr, &bji);       /* <psi|d_j> */
            Q[2 * ((size_t)i * P + j)]     = gr - (air * bjr - aii * bji);
            Q[2 * ((size_t)i * P + j) + 1] = gi - (air * bji + aii * bjr);
        }
    }
}

void oracle_qgt_pairwise(size_t dim, int P, const double *psi, const double *D, double *Q)
{
    qgt_pairwise_rows(dim, P, psi, D, Q, 0, P, 0);
}

/* quantum_geometric_tensor_network.c:1127-1175 per element (mu <= nu), with the
 * Hermitian fill of quantum_geometric_curvature.c:300-317; rows [r0, r1). */
static void qgt_elementwise_rows(size_t dim, int P, const double *psi, const double *D, double *Q, int r0, int r1,
                                 int threads)
{
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int mu = r0; mu < r1; mu++) {
        const double *dm = D + 2 * dim * (size_t)mu;
        for (int nu = mu; nu < P; nu++) {
            const double *dn = D + 2 * dim * (size_t)nu;
            double ovr, ovi, pir, pii, pjr, pji;
            dotc(dim, dm, dn, &ovr, &ovi);        /* overlap <d_mu|d_nu> */
            dotc(dim, dm, psi, &pir, &pii);       /* proj_i = <d_mu|psi> */
            dotc(dim, dn, psi, &pjr, &pji);       /* <d_nu|psi>, conjugated below */
            pji = -pji;                           /* proj_j = <psi|d_nu> */
            double qr = ovr - (pir * pjr - pii * pji);
            double qi = ovi - (pir * pji + pii * pjr);
            Q[2 * ((size_t)mu * P + nu)] = qr;
            Q[2 * ((size_t)mu * P + nu) + 1] = qi;
            if (nu != mu) {
                Q[2 * ((size_t)nu * P + mu)] = qr;
                Q[2 * ((size_t)nu * P + mu) + 1] = -qi;
            }
        }
    }
}

void oracle_qgt_elementwise(size_t dim, int P, const double *psi, const double *D, double *Q)
{
    qgt_elementwise_rows(dim, P, psi, D, Q, 0, P, 0);
}

/* Same value as the elementwise form with the P projections hoisted. */
static void qgt_hoisted(size_t dim, int nr, const int *rows, int nc, const int *cols,
                        const double *psi, const double *Drow, const double *Dcol, double *Q)
{
    double *ar = (double *)malloc(sizeof(double) * 2 * (nr + nc));
    double *ac = ar + 2 * nr;
    (void)rows; (void)cols;
    for (int i = 0; i < nr; i++) dotc(dim, Drow + 2 * dim * (size_t)i, psi, &ar[2 * i], &ar[2 * i + 1]);
    for (int j = 0; j < nc; j++) dotc(dim, Dcol + 2 * dim * (size_t)j, psi, &ac[2 * j], &ac[2 * j + 1]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nr; i++) {
        for (int j = 0; j < nc; j++) {
            double gr, gi;
            dotc(dim, Drow + 2 * dim * (size_t)i, Dcol + 2 * dim * (size_t)j, &gr, &gi);
            /* a_i conj(a_j) with a = <d|psi> */
            double pr = ar[2 * i] * ac[2 * j] + ar[2 * i + 1] * ac[2 * j + 1];
            double pi = ar[2 * i + 1] * ac[2 * j] - ar[2 * i] * ac[2 * j + 1];
            Q[2 * ((size_t)i * nc + j)] = gr - pr;
            Q[2 * ((size_t)i * nc + j) + 1] = gi - pi;
        }
    }
    free(ar);
}

int oracle_qgt(int n, int nops, const oracle_op_t *ops, double *Q)
{
    const size_t dim = (size_t)1 << n;
    int P = oracle_num_params(nops, ops);
    double *psi = (double *)malloc(2 * dim * sizeof(double));
    double *D = (double *)malloc(2 * dim * sizeof(double) * (size_t)(P > 0 ? P : 1));
    if (!psi || !D) { free(psi); free(D); return -1; }
    int rc = oracle_derivative_states(n, nops, ops, psi, D);
    if (rc == 0) oracle_qgt_elementwise(dim, P, psi, D, Q);
    free(psi);
    free(D);
    return rc;
}

int oracle_qgt_block(int n, int nops, const oracle_op_t *ops, int nr, const int *rows,
                     int nc, const int *cols, double *Q)
{
    const size_t dim = (size_t)1 << n;
    double *psi = (double *)malloc(2 * dim * sizeof(double));
    double *Dr = (double *)malloc(2 * dim * sizeof(double) * (size_t)nr);
    double *Dc = (double *)malloc(2 * dim * sizeof(double) * (size_t)nc);
    if (!psi || !Dr || !Dc) { free(psi); free(Dr); free(Dc); return -1; }
    int rc = oracle_derivative_subset(n, nops, ops, nr, rows, psi, Dr);
    if (rc == 0) rc = oracle_derivative_subset(n, nops, ops, nc, cols, psi, Dc);
    if (rc == 0) qgt_hoisted(dim, nr, rows, nc, cols, psi, Dr, Dc, Q);
    free(psi);
    free(Dr);
    free(Dc);
    return rc;
}

/* ------------------------------------------------- dense brute force */
/* Embed a 2x2 matrix on qubit q as I_(2^(n-1-q)) (x) m (x) I_(2^q). */
static void embed(int n, int q, const double *m, double *E)
{
    const size_t dim = (size_t)1 << n;
    memset(E, 0, 2 * dim * dim * sizeof(double));
    for (size_t r = 0; r < dim; r++) {
        for (size_t c = 0; c < dim; c++) {
            /* Kronecker: identity factors require equal bits outside q */
            if ((r & ~((size_t)1 << q)) != (c & ~((size_t)1 << q))) continue;
            int br = (int)((r >> q) & 1), bc = (int)((c >> q) & 1);
            E[2 * (r * dim + c)] = m[2 * (2 * br + bc)];
            E[2 * (r * dim + c) + 1] = m[2 * (2 * br + bc) + 1];
        }
    }
}

static void matmul(size_t dim, const double *A, const double *B, double *C)
{
    for (size_t i = 0; i < dim; i++)
        for (size_t j = 0; j < dim; j++) {
            double sr = 0, si = 0;
            for (size_t k = 0; k < dim; k++) {
                double ar = A[2 * (i * dim + k)], ai = A[2 * (i * dim + k) + 1];
                double br = B[2 * (k * dim + j)], bi = B[2 * (k * dim + j) + 1];
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
            C[2 * (i * dim + j)] = sr;
            C[2 * (i * dim + j) + 1] = si;
        }
}

static void dense_gate(int n, const oracle_op_t *op, double *G, double *t1, double *t2)
{
    const size_t dim = (size_t)1 << n;
    double m[8], p0[8], p1[8], x[8], y[8], z[8];
    set_m(p0, 1, 0, 0, 0, 0, 0, 0, 0);
    set_m(p1, 0, 0, 0, 0, 0, 0, 1, 0);
    set_m(x, 0, 0, 1, 0, 1, 0, 0, 0);
    set_m(y, 0, 0, 0, -1, 0, 1, 0, 0);
    set_m(z, 1, 0, 0, 0, 0, 0, -1, 0);
    if (op->type == ORC_CNOT || op->type == ORC_CY || op->type == ORC_CZ) {
        const double *u = op->type == ORC_CNOT ? x : op->type == ORC_CY ? y : z;
        /* P0_c + P1_c U_t */
        embed(n, op->q0, p1, t1);
        embed(n, op->q1, u, t2);
        matmul(dim, t1, t2, G);
        embed(n, op->q0, p0, t1);
        for (size_t i = 0; i < 2 * dim * dim; i++) G[i] += t1[i];
    } else if (op->type == ORC_SWAP) {
        /* (I + XX + YY + ZZ) / 2 */
        const double *pa[3] = {x, y, z};
        memset(G, 0, 2 * dim * dim * sizeof(double));
        for (size_t i = 0; i < dim; i++) G[2 * (i * dim + i)] = 0.5;
        double *acc = (double *)malloc(2 * dim * dim * sizeof(double));
        for (int k = 0; k < 3; k++) {
            embed(n, op->q0, pa[k], t1);
            embed(n, op->q1, pa[k], t2);
            matmul(dim, t1, t2, acc);
            for (size_t i = 0; i < 2 * dim * dim; i++) G[i] += 0.5 * acc[i];
        }
        free(acc);
    } else {
        oracle_gate_matrix(op, m);
        embed(n, op->q0, m, G);
    }
}

int oracle_dense_unitary(int n, int nops, const oracle_op_t *ops, double *U)
{
    if (n < 1 || n > 6) return -1;
    const size_t dim = (size_t)1 << n, sz = 2 * dim * dim;
    double *G = (double *)malloc(sz * sizeof(double));
    double *T = (double *)malloc(sz * sizeof(double));
    double *t1 = (double *)malloc(sz * sizeof(double));
    double *t2 = (double *)malloc(sz * sizeof(double));
    memset(U, 0, sz * sizeof(double));
    for (size_t i = 0; i < dim; i++) U[2 * (i * dim + i)] = 1.0;
    for (int g = 0; g < nops; g++) {
        dense_gate(n, &ops[g], G, t1, t2);
        matmul(dim, G, U, T);           /* later gates act from the left */
        memcpy(U, T, sz * sizeof(double));
    }
    free(G); free(T); free(t1); free(t2);
    return 0;
}

/* ---------------------------------------------------------- Hamiltonians */
double oracle_pauli_expectation(int n, const double *psi, const char *pauli)
{
    const size_t dim = (size_t)1 << n;
    if ((int)strlen(pauli) != n) return 0.0;
    double *t = (double *)malloc(2 * dim * sizeof(double));
    memcpy(t, psi, 2 * dim * sizeof(double));
    for (int q = 0; q < n; q++) {
        char c = pauli[n - 1 - q];
        oracle_op_t op = {ORC_I, q, -1, -1, 0.0};
        if (c == 'X' || c == 'x') op.type = ORC_X;
        else if (c == 'Y' || c == 'y') op.type = ORC_Y;
        else if (c == 'Z' || c == 'z') op.type = ORC_Z;
        else continue;
        oracle_apply_op(n, t, &op);
    }
    double re, im;
    dotc(dim, psi, t, &re, &im);
    free(t);
    return re;
}

void oracle_tfim_apply(int n, double J, double h, const double *psi, double *hpsi)
{
    const size_t dim = (size_t)1 << n;
#pragma omp parallel for if (dim >= OMP_MIN_DIM) schedule(static)
    for (size_t x = 0; x < dim; x++) {
        double zz = 0.0;
        for (int i = 0; i + 1 < n; i++) {
            int zi = ((x >> i) & 1) ? -1 : 1, zj = ((x >> (i + 1)) & 1) ? -1 : 1;
            zz += (double)(zi * zj);
        }
        double re = -J * zz * psi[2 * x], im = -J * zz * psi[2 * x + 1];
        for (int i = 0; i < n; i++) {
            size_t y = x ^ ((size_t)1 << i);
            re -= h * psi[2 * y];
            im -= h * psi[2 * y + 1];
        }
        hpsi[2 * x] = re;
        hpsi[2 * x + 1] = im;
    }
}

double oracle_tfim_energy(int n, double J, double h, const double *psi)
{
    const size_t dim = (size_t)1 << n;
    double *hp = (double *)malloc(2 * dim * sizeof(double));
    oracle_tfim_apply(n, J, h, psi, hp);
    double re, im;
    dotc(dim, psi, hp, &re, &im);
    free(hp);
    return re;
}

int oracle_tfim_gradient(int n, int nops, const oracle_op_t *ops, double J, double h,
                         double *energy, double *grad)
{
    const size_t dim = (size_t)1 << n;
    int P = oracle_num_params(nops, ops);
    double *psi = (double *)malloc(2 * dim * sizeof(double));
    double *hp = (double *)malloc(2 * dim * sizeof(double));
    double *D = (double *)malloc(2 * dim * sizeof(double) * (size_t)(P > 0 ? P : 1));
    if (!psi || !hp || !D) { free(psi); free(hp); free(D); return -1; }
    int rc = oracle_derivative_states(n, nops, ops, psi, D);
    if (rc == 0) {
        oracle_tfim_apply(n, J, h, psi, hp);
        double re, im;
        dotc(dim, psi, hp, &re, &im);
        *energy = re;
        for (int k = 0; k < P; k++) {
            dotc(dim, D + 2 * dim * (size_t)k, hp, &re, &im);
            grad[k] = 2.0 * re;
        }
    }
    free(psi); free(hp); free(D);
    return rc;
}

/* --------------------------------------------------------------- solves */
int oracle_cholesky_solve(int P, const double *A, const double *b, double *x)
{
    double *L = (double *)calloc((size_t)P * P, sizeof(double));
    if (!L) return -1;
    for (int j = 0; j < P; j++) {
        double d = A[(size_t)j * P + j];
        for (int k = 0; k < j; k++) d -= L[(size_t)j * P + k] * L[(size_t)j * P + k];
        if (!(d > 0.0)) { free(L); return -1; }
        double ljj = sqrt(d);
        L[(size_t)j * P + j] = ljj;
        for (int i = j + 1; i < P; i++) {
            double s = A[(size_t)i * P + j];
            for (int k = 0; k < j; k++) s -= L[(size_t)i * P + k] * L[(size_t)j * P + k];
            L[(size_t)i * P + j] = s / ljj;
        }
    }
    for (int i = 0; i < P; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= L[(size_t)i * P + k] * x[k];
        x[i] = s / L[(size_t)i * P + i];
    }
    for (int i = P - 1; i >= 0; i--) {
        double s = x[i];
        for (int k = i + 1; k < P; k++) s -= L[(size_t)k * P + i] * x[k];
        x[i] = s / L[(size_t)i * P + i];
    }
    free(L);
    return 0;
}

static void apply_precond(int P, const double *d, const double *U, int r, const double *v, double *z,
                          double *tmp)
{
    for (int i = 0; i < P; i++) z[i] = d ? d[i] * v[i] : v[i];
    if (U && r > 0) {
        for (int c = 0; c < r; c++) {
            double s = 0.0;
            for (int i = 0; i < P; i++) s += U[(size_t)i * r + c] * v[i];
            tmp[c] = s;
        }
        for (int i = 0; i < P; i++) {
            double s = 0.0;
            for (int c = 0; c < r; c++) s += U[(size_t)i * r + c] * tmp[c];
            z[i] += s;
        }
    }
}

/* classical_optimization_engine.c:543-610 (CG, x0 = 0) with a preconditioner. */
int oracle_pcg(int P, const double *A, const double *b, double *x, double tol, int maxit,
               const double *d, const double *U, int r, double *final_rel_res)
{
    double *rv = (double *)malloc(sizeof(double) * (4 * P + (r > 0 ? r : 1)));
    double *z = rv + P, *p = z + P, *Ap = p + P, *tmp = Ap + P;
    double bn = 0.0;
    for (int i = 0; i < P; i++) { x[i] = 0.0; rv[i] = b[i]; bn += b[i] * b[i]; }
    bn = sqrt(bn);
    if (bn == 0.0) { free(rv); if (final_rel_res) *final_rel_res = 0.0; return 0; }
    apply_precond(P, d, U, r, rv, z, tmp);
    double rz = 0.0;
    for (int i = 0; i < P; i++) { p[i] = z[i]; rz += rv[i] * z[i]; }
    int it;
    double rel = 1.0;
    for (it = 0; it < maxit; it++) {
        double pAp = 0.0;
        for (int i = 0; i < P; i++) {
            double s = 0.0;
            for (int k = 0; k < P; k++) s += A[(size_t)i * P + k] * p[k];
            Ap[i] = s;
            pAp += p[i] * s;
        }
        double alpha = rz / pAp;
        double rn = 0.0;
        for (int i = 0; i < P; i++) { x[i] += alpha * p[i]; rv[i] -= alpha * Ap[i]; rn += rv[i] * rv[i]; }
        rel = sqrt(rn) / bn;
        if (rel <= tol) { it++; break; }
        apply_precond(P, d, U, r, rv, z, tmp);
        double rz_new = 0.0;
        for (int i = 0; i < P; i++) rz_new += rv[i] * z[i];
        double beta = rz_new / rz;
        rz = rz_new;
        for (int i = 0; i < P; i++) p[i] = z[i] + beta * p[i];
    }
    if (final_rel_res) *final_rel_res = rel;
    free(rv);
    return rel <= tol ? it : -1;
}

/* Cyclic complex Jacobi for Hermitian A (row-major complex). */
int oracle_hermitian_eig(int P, const double *A_in, double *evals, double *V)
{
    const size_t sz = 2 * (size_t)P * P;
    double *A = (double *)malloc(sz * sizeof(double));
    if (!A) return -1;
    memcpy(A, A_in, sz * sizeof(double));
    memset(V, 0, sz * sizeof(double));
    for (int i = 0; i < P; i++) V[2 * ((size_t)i * P + i)] = 1.0;
#define AR(i, j) A[2 * ((size_t)(i) * P + (j))]
#define AI(i, j) A[2 * ((size_t)(i) * P + (j)) + 1]
    double fro = 0.0;
    for (size_t i = 0; i < sz; i++) fro += A[i] * A[i];
    fro = sqrt(fro);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < P; p++)
            for (int q = p + 1; q < P; q++) off += AR(p, q) * AR(p, q) + AI(p, q) * AI(p, q);
        if (sqrt(off) <= 1e-15 * (fro > 0 ? fro : 1.0)) break;
        for (int p = 0; p < P; p++) {
            for (int q = p + 1; q < P; q++) {
                double apr = AR(p, q), api = AI(p, q);
                double mag = hypot(apr, api);
                if (mag == 0.0) continue;
                double er = apr / mag, ei = api / mag;          /* e = a_pq/|a_pq| */
                double app = AR(p, p), aqq = AR(q, q);
                double tau = (aqq - app) / (2.0 * mag);
                double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                /* J = [[c, s], [-s conj(e), c conj(e)]] on (p, q) */
                double jpp_r = c, jpp_i = 0, jpq_r = s, jpq_i = 0;
                double jqp_r = -s * er, jqp_i = s * ei, jqq_r = c * er, jqq_i = -c * ei;
                for (int k = 0; k < P; k++) {     /* columns: A <- A J */
                    double kpr = AR(k, p), kpi = AI(k, p), kqr = AR(k, q), kqi = AI(k, q);
                    AR(k, p) = kpr * jpp_r - kpi * jpp_i + kqr * jqp_r - kqi * jqp_i;
                    AI(k, p) = kpr * jpp_i + kpi * jpp_r + kqr * jqp_i + kqi * jqp_r;
                    AR(k, q) = kpr * jpq_r - kpi * jpq_i + kqr * jqq_r - kqi * jqq_i;
                    AI(k, q) = kpr * jpq_i + kpi * jpq_r + kqr * jqq_i + kqi * jqq_r;
                    double *vp = &V[2 * ((size_t)k * P + p)], *vq = &V[2 * ((size_t)k * P + q)];
                    double vpr = vp[0], vpi = vp[1], vqr = vq[0], vqi = vq[1];
                    vp[0] = vpr * jpp_r - vpi * jpp_i + vqr * jqp_r - vqi * jqp_i;
                    vp[1] = vpr * jpp_i + vpi * jpp_r + vqr * jqp_i + vqi * jqp_r;
                    vq[0] = vpr * jpq_r - vpi * jpq_i + vqr * jqq_r - vqi * jqq_i;
                    vq[1] = vpr * jpq_i + vpi * jpq_r + vqr * jqq_i + vqi * jqq_r;
                }
                for (int k = 0; k < P; k++) {     /* rows: A <- J^H A */
                    double pkr = AR(p, k), pki = AI(p, k), qkr = AR(q, k), qki = AI(q, k);
                    AR(p, k) = jpp_r * pkr + jpp_i * pki + jqp_r * qkr + jqp_i * qki;
                    AI(p, k) = jpp_r * pki - jpp_i * pkr + jqp_r * qki - jqp_i * qkr;
                    AR(q, k) = jpq_r * pkr + jpq_i * pki + jqq_r * qkr + jqq_i * qki;
                    AI(q, k) = jpq_r * pki - jpq_i * pkr + jqq_r * qki - jqq_i * qkr;
                }
                AI(p, p) = 0.0; AI(q, q) = 0.0; AR(p, q) = AI(p, q) = AR(q, p) = AI(q, p) = 0.0;
            }
        }
    }
    for (int i = 0; i < P; i++) evals[i] = AR(i, i);
    /* ascending, permuting columns of V alongside */
    for (int i = 0; i < P; i++) {
        int m = i;
        for (int j = i + 1; j < P; j++) if (evals[j] < evals[m]) m = j;
        if (m != i) {
            double t = evals[i]; evals[i] = evals[m]; evals[m] = t;
            for (int k = 0; k < P; k++) {
                double *a = &V[2 * ((size_t)k * P + i)], *b = &V[2 * ((size_t)k * P + m)];
                double tr = a[0], ti = a[1];
                a[0] = b[0]; a[1] = b[1]; b[0] = tr; b[1] = ti;
            }
        }
    }
#undef AR
#undef AI
    free(A);
    return 0;
}

/* Complex LU with partial pivoting; singular when |pivot|^2 < 1e-10 on
 * columns 0..P-2, as matrix_operations.c:60-110. */
static int lu_solve_complex(int P, const double *A_in, const double *b, double *x)
{
    const size_t sz = 2 * (size_t)P * P;
    double *a = (double *)malloc(sz * sizeof(double));
    int *piv = (int *)malloc(sizeof(int) * P);
    memcpy(a, A_in, sz * sizeof(double));
    for (int i = 0; i < P; i++) piv[i] = i;
#define LR(i, j) a[2 * ((size_t)(i) * P + (j))]
#define LI(i, j) a[2 * ((size_t)(i) * P + (j)) + 1]
    for (int i = 0; i < P - 1; i++) {
        int p = i;
        double mx = LR(i, i) * LR(i, i) + LI(i, i) * LI(i, i);
        for (int j = i + 1; j < P; j++) {
            double v = LR(j, i) * LR(j, i) + LI(j, i) * LI(j, i);
            if (v > mx) { mx = v; p = j; }
        }
        if (mx < 1e-10) { free(a); free(piv); return -1; }
        if (p != i) {
            int t = piv[i]; piv[i] = piv[p]; piv[p] = t;
            for (int j = 0; j < P; j++) {
                double tr = LR(i, j), ti = LI(i, j);
                LR(i, j) = LR(p, j); LI(i, j) = LI(p, j); LR(p, j) = tr; LI(p, j) = ti;
            }
        }
        double dr = LR(i, i), di = LI(i, i), dd = dr * dr + di * di;
        for (int j = i + 1; j < P; j++) {
            double nr = (LR(j, i) * dr + LI(j, i) * di) / dd;
            double ni = (LI(j, i) * dr - LR(j, i) * di) / dd;
            LR(j, i) = nr; LI(j, i) = ni;
            for (int k = i + 1; k < P; k++) {
                LR(j, k) -= nr * LR(i, k) - ni * LI(i, k);
                LI(j, k) -= nr * LI(i, k) + ni * LR(i, k);
            }
        }
    }
    for (int i = 0; i < P; i++) { x[2 * i] = b[2 * piv[i]]; x[2 * i + 1] = b[2 * piv[i] + 1]; }
    for (int i = 0; i < P; i++)
        for (int j = 0; j < i; j++) {
            double xr = x[2 * j], xi = x[2 * j + 1];
            x[2 * i] -= LR(i, j) * xr - LI(i, j) * xi;
            x[2 * i + 1] -= LR(i, j) * xi + LI(i, j) * xr;
        }
    for (int i = P - 1; i >= 0; i--) {
        for (int j = i + 1; j < P; j++) {
            double xr = x[2 * j], xi = x[2 * j + 1];
            x[2 * i] -= LR(i, j) * xr - LI(i, j) * xi;
            x[2 * i + 1] -= LR(i, j) * xi + LI(i, j) * xr;
        }
        double dr = LR(i, i), di = LI(i, i), dd = dr * dr + di * di;
        double xr = x[2 * i], xi = x[2 * i + 1];
        x[2 * i] = (xr * dr + xi * di) / dd;
        x[2 * i + 1] = (xi * dr - xr * di) / dd;
    }
#undef LR
#undef LI
    free(a);
    free(piv);
    return 0;
}

double oracle_regularized_natural_gradient(int P, const double *G, const double *g, double *x,
                                           double lambda, double cond_threshold, int adaptive,
                                           int pinv_fallback, double cutoff)
{
    const size_t sz = 2 * (size_t)P * P;
    double *ev = (double *)malloc(sizeof(double) * P);
    double *V = (double *)malloc(sz * sizeof(double));
    double *A = (double *)malloc(sz * sizeof(double));
    if (!ev || !V || !A) { free(ev); free(V); free(A); return -1.0; }
    oracle_hermitian_eig(P, G, ev, V);
    if (adaptive) {
        /* quantum_geometric_gradient.c:2733-2781, 2901-2914: kappa from singular
         * values (|eigenvalues| of a Hermitian metric); min over s > 0. */
        double smax = 0.0, smin = INFINITY;
        for (int i = 0; i < P; i++) {
            double s = fabs(ev[i]);
            if (s > smax) smax = s;
            if (s > 0.0 && s < smin) smin = s;
        }
        double kappa = (smin > 1e-15) ? smax / smin : INFINITY;
        /* An exactly singular metric gives kappa = inf; lambda is then kept
         * (the reference would set lambda = inf and return NaN). */
        if (isfinite(kappa) && kappa > cond_threshold) {
            double al = 1e-6 * sqrt(kappa);
            if (al > lambda) lambda = al;
        }
    }
    memcpy(A, G, sz * sizeof(double));
    for (int i = 0; i < P; i++) A[2 * ((size_t)i * P + i)] += lambda;
    if (lu_solve_complex(P, A, g, x) != 0) {
        if (!pinv_fallback) { free(ev); free(V); free(A); return -1.0; }
        /* quantum_geometric_gradient.c:2800-2884 on the unregularised metric:
         * x = V diag(1/s_i, s_i > cutoff) V^H g (singular values |lambda_i|,
         * sign carried by the eigenvalue). */
        double *t = (double *)calloc(2 * (size_t)P, sizeof(double));
        for (int i = 0; i < P; i++) {
            double tr = 0, ti = 0;
            for (int k = 0; k < P; k++) {
                double vr = V[2 * ((size_t)k * P + i)], vi = -V[2 * ((size_t)k * P + i) + 1];
                tr += vr * g[2 * k] - vi * g[2 * k + 1];
                ti += vr * g[2 * k + 1] + vi * g[2 * k];
            }
            if (fabs(ev[i]) > cutoff) { t[2 * i] = tr / ev[i]; t[2 * i + 1] = ti / ev[i]; }
        }
        for (int k = 0; k < P; k++) {
            double sr = 0, si = 0;
            for (int i = 0; i < P; i++) {
                double vr = V[2 * ((size_t)k * P + i)], vi = V[2 * ((size_t)k * P + i) + 1];
                sr += vr * t[2 * i] - vi * t[2 * i + 1];
                si += vr * t[2 * i + 1] + vi * t[2 * i];
            }
            x[2 * k] = sr; x[2 * k + 1] = si;
        }
        free(t);
        lambda = 0.0;
    }
    free(ev); free(V); free(A);
    return lambda;
}

/* Full QGT with the derivative states built once, the P overlaps
 * <d_k|psi> hoisted out of the element loop, and the Hermitian fill of
 * quantum_geometric_curvature.c:300-317: same values as oracle_qgt, the
 * loop order a fast CPU port would use (the timed CPU baseline). */
int oracle_qgt_hoisted(int n, int nops, const oracle_op_t *ops, double *Q)
{
    const size_t dim = (size_t)1 << n;
    int P = oracle_num_params(nops, ops);
    double *psi = (double *)malloc(2 * dim * sizeof(double));
    double *D = (double *)malloc(2 * dim * sizeof(double) * (size_t)(P > 0 ? P : 1));
    double *a = (double *)malloc(2 * sizeof(double) * (size_t)(P > 0 ? P : 1));
    if (!psi || !D || !a) { free(psi); free(D); free(a); return -1; }
    int rc = oracle_derivative_states(n, nops, ops, psi, D);
    if (rc == 0) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k < P; k++) dotc(dim, D + 2 * dim * (size_t)k, psi, &a[2 * k], &a[2 * k + 1]);
#pragma omp parallel for schedule(dynamic, 1)
        for (int mu = 0; mu < P; mu++) {
            for (int nu = mu; nu < P; nu++) {
                double gr, gi;
                dotc(dim, D + 2 * dim * (size_t)mu, D + 2 * dim * (size_t)nu, &gr, &gi);
                const double pr = a[2 * mu] * a[2 * nu] + a[2 * mu + 1] * a[2 * nu + 1];
                const double pi = a[2 * mu + 1] * a[2 * nu] - a[2 * mu] * a[2 * nu + 1];
                Q[2 * ((size_t)mu * P + nu)] = gr - pr;
                Q[2 * ((size_t)mu * P + nu) + 1] = gi - pi;
                Q[2 * ((size_t)nu * P + mu)] = gr - pr;
                Q[2 * ((size_t)nu * P + mu) + 1] = -(gi - pi);
            }
        }
    }
    free(psi); free(D); free(a);
    return rc;
}

/* ----------------------------------------------- CPU baseline timing */
static double wall_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int oracle_derivative_states_timed(int n, int nops, const oracle_op_t *ops, int threads, double *psi, double *D,
                                   double *seconds)
{
    const int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    const double t0 = wall_s();
    const int rc = oracle_derivative_states(n, nops, ops, psi, D);
    *seconds = wall_s() - t0;
    omp_set_num_threads(saved);
    return rc;
}

int oracle_qgt_loop_time(size_t dim, int P, const double *psi, const double *D, int variant, int threads, int rows,
                         int reps, double *sample_s, long *elements)
{
    if (P < 1 || rows < 1 || rows > P || reps < 1 || (variant != 0 && variant != 1)) return -1;
    double *Q = (double *)malloc(2 * sizeof(double) * (size_t)P * (size_t)P);
    if (!Q) return -1;
    long el = 0;
    for (int r = 0; r < rows; r++) el += variant == 0 ? (long)(P - r) : (long)P;
    *elements = el;
    for (int k = 0; k < reps; k++) {
        const double t0 = wall_s();
        if (variant == 0) qgt_elementwise_rows(dim, P, psi, D, Q, 0, rows, threads);
        else qgt_pairwise_rows(dim, P, psi, D, Q, 0, rows, threads);
        sample_s[k] = wall_s() - t0;
    }
    free(Q);
    return 0;
}
