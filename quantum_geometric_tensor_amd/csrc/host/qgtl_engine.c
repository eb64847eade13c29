/*
 * qgtl_engine.c -- runs a compiled plan on the GPU through the thin HIP ABI.
 *
 * Mixed mode (complex64 sweeps, MFMA split-BF16x3 Gram, FP64 reductions):
 *   for each block b: apply b to the primal; create the derivative states of
 *   b's parameters from the block output (inject); carry them through the
 *   later blocks; then G = D^H D on MFMA and Q = normalised G - a a^H in FP64.
 *   When every pass is in place (states of <= 2^13 amplitudes) the derivative
 *   states advance breadth-first, one launch per pass for all of them; wider
 *   states advance depth-first in balanced chunks of up to 1 GB (a chunk's states
 *   move together through the later blocks, so launches stay large while
 *   the ping-pong copy of a chunk is all the extra memory needed).
 * Plain FP64 mode: one complex128 kernel per gate, derivative states one at a
 *   time (gate insertion of -i/2 G), tiled FP64 Gram, FP64 correction.
 *
 * QGTL counterpart of the whole: compute_quantum_geometric_tensor looped by
 * geometric_compute_full_qgt (src/quantum_geometric/core/quantum_geometric_curvature.c:289-320).
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qgtl_internal.h"

/* ------------------------------------------------------------------ errors */
static _Thread_local char g_err[512] = "";

int qgtl_set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *qgtl_last_error(void) { return g_err; }

static int hip_fail(int rc, const char *where)
{
    return qgtl_set_error(rc, "%s: %s", where, qgtl_hip_last_error());
}

#define HIPCALL(call)                                          \
    do {                                                       \
        int _rc = (call);                                      \
        if (_rc) return hip_fail(_rc, #call);                  \
    } while (0)

/* ------------------------------------------------------------------ lock */
static pthread_mutex_t g_lock;
static pthread_once_t g_lock_once = PTHREAD_ONCE_INIT;

static void lock_init(void)
{
    pthread_mutexattr_t a;
    pthread_mutexattr_init(&a);
    pthread_mutexattr_settype(&a, PTHREAD_MUTEX_RECURSIVE);
    pthread_mutex_init(&g_lock, &a);
    pthread_mutexattr_destroy(&a);
}

void qgtl_lock(void)
{
    pthread_once(&g_lock_once, lock_init);
    pthread_mutex_lock(&g_lock);
}

void qgtl_unlock(void) { pthread_mutex_unlock(&g_lock); }

/* ------------------------------------------------------------------ init */
static int g_inited = 0;
static int g_profiling = 0;

int qgtl_init(int device)
{
    QGTL_LOCKED();
    int rc = qgtl_hip_init(device);
    if (rc) return hip_fail(rc, "qgtl_init");
    g_inited = 1;
    return 0;
}

static int ensure_init(void)
{
    if (g_inited) return 0;
    const char *e = getenv("QGTL_DEVICE");
    return qgtl_init(e ? atoi(e) : 0);
}

int qgtl_device_count(void) { return qgtl_hip_device_count(); }

void qgtl_set_profiling(int enable)
{
    QGTL_LOCKED();
    g_profiling = enable ? 1 : 0;
}

/* ------------------------------------------------------------- profiling */
enum { K_SWEEP = 0, K_INJECT, K_GRAM_MAIN, K_GRAM_AUX, K_FINAL, K_OTHER, K_KINDS };

typedef struct {
    qgtl_hip_event_t *ev;
    int nev, cap, used;
    int *kind;          /* per pair */
    int npairs, pcap;
} prof_t;

static prof_t g_prof;

static int prof_event(qgtl_hip_event_t *out)
{
    if (g_prof.used == g_prof.nev) {
        if (g_prof.nev == g_prof.cap) {
            int nc = g_prof.cap ? 2 * g_prof.cap : 1024;
            qgtl_hip_event_t *ne = (qgtl_hip_event_t *)realloc(g_prof.ev, sizeof(*ne) * (size_t)nc);
            if (!ne) return QGT_ERROR_NO_MEMORY;
            g_prof.ev = ne;
            g_prof.cap = nc;
        }
        HIPCALL(qgtl_hip_event_create(&g_prof.ev[g_prof.nev]));
        g_prof.nev++;
    }
    *out = g_prof.ev[g_prof.used++];
    return 0;
}

static int prof_start(void)
{
    if (!g_profiling) return 0;
    qgtl_hip_event_t e;
    int rc = prof_event(&e);
    if (rc) return rc;
    HIPCALL(qgtl_hip_event_record(e));
    return 0;
}

static int prof_stop(int kind)
{
    if (!g_profiling) return 0;
    qgtl_hip_event_t e;
    int rc = prof_event(&e);
    if (rc) return rc;
    HIPCALL(qgtl_hip_event_record(e));
    if (g_prof.npairs == g_prof.pcap) {
        int nc = g_prof.pcap ? 2 * g_prof.pcap : 512;
        int *nk = (int *)realloc(g_prof.kind, sizeof(int) * (size_t)nc);
        if (!nk) return QGT_ERROR_NO_MEMORY;
        g_prof.kind = nk;
        g_prof.pcap = nc;
    }
    g_prof.kind[g_prof.npairs++] = kind;
    return 0;
}

static void prof_reset(void)
{
    g_prof.used = 0;
    g_prof.npairs = 0;
}

/* After a stream sync: fold event pairs into stats. */
static int prof_collect(qgtl_stats_t *st)
{
    if (!g_profiling || !st) { prof_reset(); return 0; }
    for (int i = 0; i < g_prof.npairs; i++) {
        float ms = 0.f;
        HIPCALL(qgtl_hip_event_elapsed_ms(g_prof.ev[2 * i], g_prof.ev[2 * i + 1], &ms));
        switch (g_prof.kind[i]) {
        case K_SWEEP: st->sweep_ms += ms; break;
        case K_INJECT: st->inject_ms += ms; break;
        case K_GRAM_MAIN: st->gram_ms += ms; st->gram_kernel_ms += ms; break;
        case K_GRAM_AUX: st->gram_ms += ms; break;
        case K_FINAL: st->finalize_ms += ms; break;
        default: st->other_ms += ms; break;
        }
    }
    prof_reset();
    return 0;
}

#define LAUNCH(kind, call)                                     \
    do {                                                       \
        int _rc = prof_start();                                \
        if (_rc) return _rc;                                   \
        _rc = (call);                                          \
        if (_rc) return hip_fail(_rc, #call);                  \
        _rc = prof_stop(kind);                                 \
        if (_rc) return _rc;                                   \
        if (st) st->kernel_launches++;                         \
    } while (0)

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------ HEA */
int qgtl_hea_num_ops(int n, int layers) { return layers * (2 * n + (n >= 2 ? n : 0)); }

int qgtl_hea_ops(int n, int layers, const double *theta, qgtl_op_t *ops)
{
    int k = 0;
    for (int l = 0; l < layers; l++) {
        for (int q = 0; q < n; q++) {
            const int p = 2 * n * l + q;
            qgtl_op_t op = {GATE_TYPE_RY, q, -1, p, theta[p]};
            ops[k++] = op;
        }
        for (int q = 0; q < n; q++) {
            const int p = 2 * n * l + n + q;
            qgtl_op_t op = {GATE_TYPE_RZ, q, -1, p, theta[p]};
            ops[k++] = op;
        }
        if (n >= 2)
            for (int q = 0; q < n; q++) {
                qgtl_op_t op = {GATE_TYPE_CNOT, q, (q + 1) % n, -1, 0.0};
                ops[k++] = op;
            }
    }
    return k;
}

int qgtl_num_params(int nops, const qgtl_op_t *ops)
{
    int P = 0;
    for (int i = 0; i < nops; i++)
        if (ops[i].param >= 0) P++;
    return P;
}

/* ------------------------------------------------------------------ plans */
static void free_chain(struct qgtl_plan *p)
{
    for (int k = 0; k < p->nchain; k++) qgtl_hip_free(p->chain[k]);
    qgtl_hip_free(p->chain_tmp);
    qgtl_hip_free(p->tmp_side);
    free(p->chain);
    p->chain = NULL;
    p->chain_tmp = NULL;
    p->tmp_side = NULL;
    p->nchain = 0;
}

static void plan_free_device(struct qgtl_plan *p)
{
    free_chain(p);
    qgtl_hip_free(p->psi[0]);
    qgtl_hip_free(p->psi[1]);
    qgtl_hip_free(p->D);
    qgtl_hip_free(p->tmp);
    qgtl_hip_free(p->tabs_dev);
    p->tabs_dev = NULL;
    p->tabs_passes = 0;
    qgtl_hip_free(p->inj_slots_dev);
    p->inj_slots_dev = NULL;
    qgtl_hip_free(p->inj_partials);
    qgtl_hip_free(p->inj_sums);
    qgtl_hip_free(p->gram_partials);
    qgtl_hip_free(p->C);
    qgtl_hip_free(p->Qdev);
    qgtl_hip_host_free(p->Qhost);
    p->Qhost = NULL;
    qgtl_hip_free(p->plain_partials);
    qgtl_hip_free(p->terms_dev);
    qgtl_hip_free(p->hpsi);
    qgtl_hip_free(p->eg_partials);
    qgtl_hip_free(p->eg_sums);
    qgtl_hip_free(p->eg_out);
    qgtl_hip_free(p->ng_slot);
    qgtl_hip_free(p->ng_A);
    qgtl_hip_free(p->ng_D);
    qgtl_hip_free(p->ng_g);
    qgtl_hip_free(p->ng_x);
    qgtl_hip_free(p->ng_d);
    qgtl_hip_free(p->ng_U);
    qgtl_hip_free(p->ng_theta);
    qgtl_hip_free(p->ng_res);
    qgtl_hip_free(p->ng_it);
    p->ng_slot = p->ng_it = NULL;
    p->ng_A = p->ng_D = p->ng_g = p->ng_x = p->ng_d = p->ng_U = p->ng_theta = p->ng_res = NULL;
    p->ng_rank = 0;
    p->terms_dev = p->hpsi = NULL;
    p->eg_partials = p->eg_sums = p->eg_out = NULL;
    p->psi_final = NULL;
    p->psi[0] = p->psi[1] = p->D = p->tmp = NULL;
    p->inj_partials = p->inj_sums = p->C = p->Qdev = p->plain_partials = NULL;
    p->gram_partials = NULL;
}

qgtl_plan_t *qgtl_plan_create(int n, int nops, const qgtl_op_t *ops, int mode, const qgtl_tune_t *tune)
{
    QGTL_LOCKED();
    if (mode != QGTL_MODE_MIXED && mode != QGTL_MODE_FP64_PLAIN) {
        qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "unknown mode %d", mode);
        return NULL;
    }
    if (nops < 0 || (nops > 0 && !ops)) {
        qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null gate list");
        return NULL;
    }
    struct qgtl_plan *p = (struct qgtl_plan *)calloc(1, sizeof *p);
    if (!p) { qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); return NULL; }
    p->n = n;
    p->mode = mode;
    p->nops = nops;
    p->ops = (qgtl_op_t *)malloc(sizeof(qgtl_op_t) * (size_t)(nops > 0 ? nops : 1));
    if (!p->ops) { free(p); qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); return NULL; }
    if (nops) memcpy(p->ops, ops, sizeof(qgtl_op_t) * (size_t)nops);
    if (tune) p->tune = *tune;
    if (qgtl_plan_build(p) != 0) {
        free(p->ops); free(p->blocks); free(p->params); free(p->slot_param); free(p);
        return NULL;
    }
    p->N = (size_t)1 << n;
    p->structure_hash = qgtl_structure_hash(n, nops, ops, mode);
    return p;
}

int qgtl_plan_num_params(const qgtl_plan_t *p) { return p ? p->P : -1; }

int qgtl_plan_set_angles_nolock(qgtl_plan_t *p, int nops, const double *angles)
{
    if (!p || !angles || nops != p->nops) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "angle count mismatch");
    for (int i = 0; i < nops; i++) p->ops[i].angle = angles[i];
    return qgtl_plan_build(p);
}

int qgtl_plan_set_angles(qgtl_plan_t *p, int nops, const double *angles)
{
    QGTL_LOCKED();
    return qgtl_plan_set_angles_nolock(p, nops, angles);
}

void qgtl_plan_destroy(qgtl_plan_t *p)
{
    QGTL_LOCKED();
    if (!p) return;
    qgtl_hip_graph_destroy(p->graph);
    plan_free_device(p);
    free(p->inj_slots_host);
    qgtl_hip_event_destroy(p->ev_psi);
    qgtl_hip_event_destroy(p->ev_inj);
    for (int k = 0; k < p->nev_chain; k++) qgtl_hip_event_destroy(p->ev_chain[k]);
    free(p->ev_chain);
    free(p->tabs_host);
    free(p->ops);
    free(p->blocks);
    free(p->params);
    free(p->slot_param);
    free(p);
}

static int any_flip(const struct qgtl_plan *p)
{
    for (int b = 0; b < p->nblocks; b++)
        for (int i = 0; i < p->blocks[b].npasses; i++)
            if (p->blocks[b].pass_flips[i]) return 1;
    return 0;
}

static int chunk_states(const struct qgtl_plan *p)
{
    if (p->tune.chunk_states > 0) return p->tune.chunk_states < QGTL_INJECT_MAX ? p->tune.chunk_states : QGTL_INJECT_MAX;
    /* larger chunks mean fewer, larger launches: at n = 20 (8 MB states) one QGT
     * took 10.3 ms with 8 states per chunk, 9.4 ms with 16, 8.9 ms with 24 and
     * 8.5 ms with one chunk per block (40 states) */
    size_t c = ((size_t)1 << 30) / (p->N * 8);
    if (c < 1) c = 1;
    if (c > QGTL_INJECT_MAX) c = QGTL_INJECT_MAX;
    return (int)c;
}

/* Split-K chunks of ~2048 amplitudes, so that one chunk of up to 256 states
 * fits the 4 MB L2 of the XCD its tiles run on (measured at n = 20, P = 240:
 * 104 chunks 3.18 ms, 256 chunks 2.63 ms, 512 chunks 2.49 ms, 1024 chunks
 * 2.64 ms), at least ~1024 workgroups, and fp32 partials of at most 1 GB or,
 * for larger state sets, 1/16 of the states' bytes, within half of the free
 * device memory (free_bytes, read after the states were allocated): longer
 * chunks let the tile pairs that share rows drift apart, so their re-reads
 * miss L2 (n = 26, P = 416: 584 chunks 243 ms, 4096 chunks 227 ms, 8192 chunks
 * 215 ms). */
static int gram_chunks(const struct qgtl_plan *p, size_t free_bytes)
{
    if (p->tune.gram_chunks > 0) return (p->tune.gram_chunks + 7) / 8 * 8;
    const size_t T = (size_t)qgtl_hip_gram_tiles(p->P + 1);
    size_t want = p->N / 2048;
    const size_t minw = (1024 + T - 1) / T;
    if (want < minw) want = minw;
    size_t budget = ((size_t)p->P + 1) * p->N * 8 / 16;
    if (budget > free_bytes / 2) budget = free_bytes / 2;
    if (budget < ((size_t)1 << 30)) budget = (size_t)1 << 30;
    const size_t cap = budget / (T * 128 * 128 * sizeof(float));
    if (want > cap) want = cap;
    const size_t maxc = p->N / 256;
    if (want > maxc) want = maxc;
    want = (want + 7) / 8 * 8;
    if (want < 8) want = 8;
    return (int)want;
}

/* The same for the complex-form Gram of split-plane states (tile pairs of 128
 * states, both parts of G per pair): chunks of ~2048 amplitudes, at least ~512
 * workgroups (one per CU at a time), partials within the same budget. */
static int gram_chunks_cplx(const struct qgtl_plan *p, size_t free_bytes)
{
    if (p->tune.gram_chunks > 0) return (p->tune.gram_chunks + 7) / 8 * 8;
    const size_t per_chunk = qgtl_hip_gram_cplx_partial_floats(p->P + 1) * sizeof(float);
    const size_t T = per_chunk / (2 * 128 * 128 * sizeof(float));
    /* 129..256 states run one workgroup per chunk whose MFMA accumulators hold the whole
     * chunk (no outer fp32 sums): chunks of 1024 amplitudes bound the accumulation's rounding */
    size_t want = T == 3 ? p->N / 1024 : p->N / 2048;
    const size_t minw = (512 + T - 1) / T;
    if (want < minw) want = minw;
    size_t budget = ((size_t)p->P + 1) * p->N * 8 / 16;
    if (budget > free_bytes / 2) budget = free_bytes / 2;
    if (budget < ((size_t)1 << 30)) budget = (size_t)1 << 30;
    const size_t cap = budget / per_chunk;
    if (want > cap) want = cap;
    const size_t maxc = p->N / 256;
    if (want > maxc) want = maxc;
    want = (want + 7) / 8 * 8;
    if (want < 8) want = 8;
    return (int)want;
}

/* Split-BF16x3 leaves a ~2.7e-6 bias (the dropped lo*lo product) and x4
 * a 2^-17 representation error that averages out only over unstructured
 * states; the default is the two-piece fp16 split (see qgtl_internal.h). */
static int gram_terms(const struct qgtl_plan *p)
{
    if (qgtl_valid_gram_terms(p->tune.gram_variant)) return p->tune.gram_variant;
    return qgtl_default_gram_terms(p->n);
}

static int gram_products(const struct qgtl_plan *p) { return gram_terms(p) == QGTL_GRAM_F16X4 ? 4 : 3; }

/* Block-boundary states of the primal chain (depth-first plans), within 2 GiB;
 * otherwise (or if an allocation fails) the chain keeps two buffers.  With the
 * chain, a second chunk buffer (tmp_side, at most 1 GiB) lets the chain's
 * stream carry the backward chunks while the library stream carries the
 * forward ones; without it every chunk runs on the library stream. */
static void alloc_chain(struct qgtl_plan *p)
{
    const size_t bytes = p->N * 8, nchain = (size_t)p->nblocks + 1;
    if (p->chain || !any_flip(p) || (nchain + 1) * bytes > ((size_t)2 << 30)) return;
    p->chain = (void **)calloc(nchain, sizeof(void *));
    if (!p->chain) return;
    int ok = qgtl_hip_malloc(&p->chain_tmp, bytes) == 0;
    for (size_t k = 0; ok && k < nchain; k++) {
        ok = qgtl_hip_malloc(&p->chain[k], bytes) == 0;
        if (ok) p->nchain = (int)k + 1;
    }
    if (!ok) {
        free_chain(p);
        return;
    }
    if (p->tmp_states > 0 && p->tmp_states * bytes <= ((size_t)1 << 30) &&
        qgtl_hip_malloc(&p->tmp_side, p->tmp_states * bytes) != 0)
        p->tmp_side = NULL;
}

static int alloc_mixed(struct qgtl_plan *p)
{
    if (p->psi[0]) return 0;
    const size_t N = p->N, P = (size_t)p->P;
    HIPCALL(qgtl_hip_malloc(&p->psi[0], N * 8));
    HIPCALL(qgtl_hip_malloc(&p->psi[1], N * 8));
    if (P > 0) {
        /* one slot more than P: depth-first plans place the primal state at the
         * meeting point there, so the Gram also yields the overlaps and norms */
        HIPCALL(qgtl_hip_malloc(&p->D, (P + 1) * N * 8));
        if (any_flip(p)) {
            p->tmp_states = (size_t)chunk_states(p);
            HIPCALL(qgtl_hip_malloc(&p->tmp, p->tmp_states * N * 8));
        }
        const size_t tiles = (N + 2047) / 2048;
        p->inj_partials_len = (size_t)QGTL_INJECT_MAX * tiles * 4;
        HIPCALL(qgtl_hip_malloc((void **)&p->inj_partials, p->inj_partials_len * sizeof(double)));
        HIPCALL(qgtl_hip_malloc((void **)&p->inj_sums, P * 4 * sizeof(double)));
        size_t free_bytes = 0;
        HIPCALL(qgtl_hip_mem_free(&free_bytes));
        /* partials and C sized for both Gram forms (the plan picks one at execute) */
        p->gram_chunks = gram_chunks(p, free_bytes);
        p->gram_chunks_cplx = gram_chunks_cplx(p, free_bytes);
        const size_t T = (size_t)qgtl_hip_gram_tiles(p->P + 1);
        size_t pbytes = (size_t)p->gram_chunks * T * 128 * 128 * sizeof(float);
        const size_t pc = (size_t)p->gram_chunks_cplx * qgtl_hip_gram_cplx_partial_floats(p->P + 1) * sizeof(float);
        if (pc > pbytes) pbytes = pc;
        HIPCALL(qgtl_hip_malloc((void **)&p->gram_partials, pbytes));
        const size_t nt = (2 * (P + 1) + 127) / 128, R = nt * 128;
        const size_t ntc = (P + 1 + 127) / 128, Rc = ntc * 128;
        HIPCALL(qgtl_hip_malloc((void **)&p->C, (R * R > 2 * Rc * Rc ? R * R : 2 * Rc * Rc) * sizeof(double)));
        HIPCALL(qgtl_hip_malloc((void **)&p->Qdev, P * P * 2 * sizeof(double)));
        HIPCALL(qgtl_hip_host_alloc((void **)&p->Qhost, P * P * 2 * sizeof(double)));
    }
    alloc_chain(p);
    return 0;
}

static int alloc_plain(struct qgtl_plan *p)
{
    if (p->psi[0]) return 0;
    const size_t N = p->N, P = (size_t)p->P;
    HIPCALL(qgtl_hip_malloc(&p->psi[0], N * 16));
    if (P > 0) {
        HIPCALL(qgtl_hip_malloc(&p->D, P * N * 16));
        p->plain_partials_len = qgtl_hip_gram_f64_partials(p->P, N);
        const size_t dot = ((N + 4095) / 4096) * P * 4;
        if (dot > p->plain_partials_len) p->plain_partials_len = dot;
        HIPCALL(qgtl_hip_malloc((void **)&p->plain_partials, p->plain_partials_len * sizeof(double)));
        HIPCALL(qgtl_hip_malloc((void **)&p->inj_sums, P * 4 * sizeof(double)));
        HIPCALL(qgtl_hip_malloc((void **)&p->C, P * P * 2 * sizeof(double)));
        HIPCALL(qgtl_hip_malloc((void **)&p->Qdev, P * P * 2 * sizeof(double)));
        HIPCALL(qgtl_hip_host_alloc((void **)&p->Qhost, P * P * 2 * sizeof(double)));
    }
    alloc_chain(p);
    return 0;
}

/* device phase tables of pass i of block b (inv: of its inverse), NULL before they exist */
static const void *pass_tabs(const struct qgtl_plan *p, const qgtl_block_t *b, int i, int inv)
{
    if (!p->tabs_dev) return NULL;
    const int k = b->tab0 + (inv ? b->npasses : 0) + i;
    return (const char *)p->tabs_dev + (size_t)k * 2 * QGTL_PHASE_TAB * 8;
}

/* The float matrix of an injection slot.  A diagonal M needs no partner: v = 0
 * makes the partner the element itself (the same line, a zero coefficient).
 * An anti-diagonal M (no diagonal) lets the pass that forms the state load the
 * partner only.  Entries below 1e-15 of the matrix's largest are the rounding
 * residue of the products that formed it (B = u^H A u of a rotation whose
 * later gates turn within its plane) and are dropped, a change far below
 * float precision. */
static void slot_matrix(qgtl_inject_slot_t *q, const qgtl_mat2 *M)
{
    double big = 0.0;
    for (int e = 0; e < 4; e++) {
        const double a = fabs(M->m[e].re) + fabs(M->m[e].im);
        if (a > big) big = a;
    }
    for (int e = 0; e < 4; e++) {
        const int tiny = fabs(M->m[e].re) + fabs(M->m[e].im) <= 1e-15 * big;
        q->M[e].re = tiny ? 0.f : (float)M->m[e].re;
        q->M[e].im = tiny ? 0.f : (float)M->m[e].im;
    }
    if (q->M[1].re == 0.f && q->M[1].im == 0.f && q->M[2].re == 0.f && q->M[2].im == 0.f) q->v = 0;
}

/* Build and upload the phase tables of every pass after the angles changed. */
static int upload_tabs(struct qgtl_plan *p)
{
    if (!p->tabs_dirty && p->tabs_dev) return 0;
    int total = 0;
    for (int bi = 0; bi < p->nblocks; bi++) {
        p->blocks[bi].tab0 = total;
        total += p->blocks[bi].npasses + p->blocks[bi].ninv;
    }
    if (total == 0) { p->tabs_dirty = 0; return 0; }
    const size_t per = 2 * QGTL_PHASE_TAB * 2;                /* floats per pass */
    if (total != p->tabs_passes) {
        qgtl_hip_free(p->tabs_dev);
        p->tabs_dev = NULL;
        free(p->tabs_host);
        p->tabs_host = (float *)malloc(sizeof(float) * per * (size_t)total);
        if (!p->tabs_host) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        HIPCALL(qgtl_hip_malloc(&p->tabs_dev, sizeof(float) * per * (size_t)total));
        p->tabs_passes = total;
    }
    int k = 0;
    for (int bi = 0; bi < p->nblocks; bi++) {
        for (int i = 0; i < p->blocks[bi].npasses; i++, k++)
            qgtl_pass_phase_tables(&p->blocks[bi].pass[i], p->tabs_host + per * (size_t)k);
        for (int i = 0; i < p->blocks[bi].ninv; i++, k++)
            qgtl_pass_phase_tables(&p->blocks[bi].ipass[i], p->tabs_host + per * (size_t)k);
    }
    HIPCALL(qgtl_hip_h2d(p->tabs_dev, p->tabs_host, sizeof(float) * per * (size_t)total));
    if (p->P > 0) {
        const size_t P = (size_t)p->P;
        if (!p->inj_slots_dev) {
            free(p->inj_slots_host);
            p->inj_slots_host = (qgtl_inject_slot_t *)malloc(2 * P * sizeof(qgtl_inject_slot_t));
            if (!p->inj_slots_host) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
            HIPCALL(qgtl_hip_malloc(&p->inj_slots_dev, 2 * P * sizeof(qgtl_inject_slot_t)));
        }
        for (size_t s = 0; s < P; s++) {
            const qgtl_param_t *pp = &p->params[p->slot_param[s]];
            qgtl_inject_slot_t *f = &p->inj_slots_host[s], *bk = &p->inj_slots_host[P + s];
            f->v = pp->v;
            f->w = pp->w;
            bk->v = bk->w = 1u << pp->qubit;
            slot_matrix(f, &pp->A);
            slot_matrix(bk, &pp->B);
        }
        HIPCALL(qgtl_hip_h2d(p->inj_slots_dev, p->inj_slots_host, 2 * P * sizeof(qgtl_inject_slot_t)));
    }
    p->tabs_dirty = 0;
    return 0;
}

/* Run passes[0..np) (of block b; inv: its inverse) over nstates complex64
 * states at *cur (stride N); a pass that flips writes to *alt and swaps the
 * two.  split_final: the last pass writes the split-plane form, always to
 * *alt (it cannot run in place). */
static int run_passes(struct qgtl_plan *p, const qgtl_block_t *b, int inv, void **cur, void **alt, int nstates,
                      qgtl_stats_t *st, int split_final, const void *inj_psi, const void *inj_slots)
{
    const qgtl_pass_desc_t *passes = inv ? b->ipass : b->pass;
    const int *flips = inv ? b->ipass_flips : b->pass_flips;
    const int np = inv ? b->ninv : b->npasses;
    for (int i = 0; i < np; i++) {
        const int split = split_final && i == np - 1;
        const int flip = flips[i] || split;
        void *dst = flip ? *alt : *cur;
        if (i == 0 && inj_psi)   /* the chunk's states are formed from inj_psi while this pass loads */
            LAUNCH(K_SWEEP, qgtl_hip_pass_c64_inject(&passes[0], inj_psi, inj_slots, dst, p->N, nstates, split,
                                                     pass_tabs(p, b, 0, inv)));
        else
            LAUNCH(K_SWEEP, qgtl_hip_pass_c64(&passes[i], *cur, dst, p->N, nstates, 0, split, pass_tabs(p, b, i, inv)));
        if (st) {
            st->sweep_launches++;
            st->sweep_state_passes += nstates;
            /* a pass that forms its states while loading writes 8 B per amplitude
             * and reads the one primal state (twice: the element and its partner) */
            if (i == 0 && inj_psi) st->sweep_bytes += (double)p->N * (8.0 * nstates + 16.0);
            else st->sweep_bytes += (double)nstates * (double)p->N * 16.0;
        }
        if (flip) { void *t = *cur; *cur = *alt; *alt = t; }
    }
    return 0;
}

/* Apply block b (its passes, then any per-gate M-stage kernels). */
static int apply_block(struct qgtl_plan *p, const qgtl_block_t *b, void **cur, void **alt, int nstates,
                       qgtl_stats_t *st, int split_final, const void *inj_psi, const void *inj_slots)
{
    int rc = run_passes(p, b, 0, cur, alt, nstates, st, split_final, inj_psi, inj_slots);
    if (rc) return rc;
    if (b->mstage_kernels)
        for (int i = 0; i < b->nperm; i++) {
            if (b->perm_type[i] == 0)
                LAUNCH(K_SWEEP, qgtl_hip_cnot_c64(*cur, p->N, nstates, p->n, b->perm_a[i], b->perm_b[i]));
            else
                LAUNCH(K_SWEEP, qgtl_hip_swap_c64(*cur, p->N, nstates, p->n, b->perm_a[i], b->perm_b[i]));
            if (st) {
                st->sweep_launches++;
                st->sweep_state_passes += nstates;
                st->sweep_bytes += (double)nstates * (double)p->N * 16.0;
            }
        }
    return 0;
}

/* flips of block bi for a derivative-state chunk carried forward (inv = 0) or
 * backward (inv = 1), the forced one of a split final pass included: the last
 * forward block is meet, the last backward block meet + 1 */
static int block_flips(const struct qgtl_plan *p, int bi, int inv)
{
    const qgtl_block_t *b = &p->blocks[bi];
    const int np = inv ? b->ninv : b->npasses;
    const int *flips = inv ? b->ipass_flips : b->pass_flips;
    const int final_block = inv ? p->meet + 1 : p->meet;
    int f = 0;
    for (int i = 0; i < np; i++) f += flips[i] || (p->presplit && bi == final_block && i == np - 1);
    return f;
}

/* Sweep passes of block bi applied to one state forward (M-stage kernels included) */
static long fwd_cost(const qgtl_block_t *b) { return b->npasses + (b->mstage_kernels ? b->nperm : 0); }

/* State-passes of a two-sided schedule meeting after block c (-1: not possible):
 * parameters of blocks b <= c are injected after block b and carried forward
 * to c; those of blocks b > c are injected at block b's input (the block's
 * inverse applied to the derivative is a 2 x 2 on the parameter's qubit) and
 * carried backward through blocks b-1 .. c+1.  The inner products the Gram
 * forms are the same after any common unitary, so all states may meet at c. */
static long meet_cost(const struct qgtl_plan *p, int c)
{
    for (int b = c + 1; b < p->nblocks; b++)
        if (!p->blocks[b].inv_ok) return -1;       /* H psi is carried back through all of them */
    long cost = 0;
    for (int b = 0; b < p->nblocks; b++) {
        const long K = p->blocks[b].param_end - p->blocks[b].param_begin;
        if (b <= c)
            for (int b2 = b + 1; b2 <= c; b2++) cost += K * fwd_cost(&p->blocks[b2]);
        else
            for (int b2 = c + 1; b2 < b; b2++) cost += K * p->blocks[b2].ninv;
    }
    return cost;
}

static int choose_meet(const struct qgtl_plan *p, int depth_first)
{
    const int last = p->nblocks - 1;
    if (!depth_first || p->tune.sweep_schedule == 1 || last < 1) return last;
    int best = last;
    long best_cost = meet_cost(p, last);
    for (int c = last - 1; c >= 0; c--) {
        const long k = meet_cost(p, c);
        if (k >= 0 && k < best_cost) { best_cost = k; best = c; }
    }
    return best;
}

int qgtl_plan_meet_block(const struct qgtl_plan *p) { return choose_meet(p, any_flip(p)); }

/* Split-plane final states: depth-first plans in which every chunk that is
 * carried anywhere ends with a register-phase pass (the last pass of block
 * meet forward, the last inverse pass of block meet + 1 backward) and no
 * separate M-stage kernels, fp16 Gram pieces; chunks injected at the meeting
 * point are written split by the injection. */
static int presplit_ok(const struct qgtl_plan *p, int depth_first, int terms)
{
    if (p->tune.no_presplit || !depth_first || p->P == 0 || p->nblocks < 1 || p->N % 64 != 0) return 0;
    if (terms != QGTL_GRAM_F16X3 && terms != QGTL_GRAM_F16X4) return 0;
    int fwd = 0, bwd = 0;
    for (int b = 0; b < p->nblocks; b++) {
        if (p->blocks[b].param_end == p->blocks[b].param_begin) continue;
        if (b < p->meet) fwd = 1;
        if (b > p->meet + 1) bwd = 1;
    }
    if (fwd) {
        const qgtl_block_t *b = &p->blocks[p->meet];
        if (b->npasses < 1 || (b->mstage_kernels && b->nperm > 0)) return 0;
        if (!qgtl_hip_pass_split_ok(&b->pass[b->npasses - 1])) return 0;
    }
    if (bwd) {
        const qgtl_block_t *b = &p->blocks[p->meet + 1];
        if (b->ninv < 1 || !qgtl_hip_pass_split_ok(&b->ipass[b->ninv - 1])) return 0;
    }
    return 1;
}

/* Inject parameters slots [s0, s0+cnt) (all from block b) into dst (split-plane
 * form when split): after the block from its output, or (backward) at the
 * block's input from the state before it. */
static int inject(struct qgtl_plan *p, int s0, int cnt, const void *psi, void *dst, qgtl_stats_t *st, int split,
                  int backward, int store)
{
    qgtl_inject_desc_t d;
    memset(&d, 0, sizeof d);
    d.n = p->n;
    d.K = cnt;
    for (int j = 0; j < cnt; j++) {
        const qgtl_param_t *pp = &p->params[p->slot_param[s0 + j]];
        const qgtl_mat2 *M = backward ? &pp->B : &pp->A;
        d.v[j] = backward ? 1u << pp->qubit : pp->v;
        if (M->m[1].re == 0.0 && M->m[1].im == 0.0 && M->m[2].re == 0.0 && M->m[2].im == 0.0)
            d.v[j] = 0;   /* diagonal: no partner (the element's own line, zero coefficient) */
        d.w[j] = backward ? 1u << pp->qubit : pp->w;
        for (int e = 0; e < 4; e++) {
            d.A[j][e].re = (float)M->m[e].re;
            d.A[j][e].im = (float)M->m[e].im;
        }
    }
    int tiles = 0;
    if (p->psi_in_gram) {
        /* overlaps and norms come from the Gram: only the states are formed */
        if (store) LAUNCH(K_INJECT, qgtl_hip_inject_c64(&d, psi, dst, p->N, NULL, &tiles, split));
        return 0;
    }
    if (store)
        LAUNCH(K_INJECT, qgtl_hip_inject_c64(&d, psi, dst, p->N, p->inj_partials, &tiles, split));
    else   /* sums only: the chunk's first pass forms the states while loading */
        LAUNCH(K_INJECT, qgtl_hip_inject_sums_c64(&d, psi, p->inj_partials, &tiles));
    LAUNCH(K_INJECT, qgtl_hip_reduce_partials_f64(p->inj_partials, cnt, tiles, p->inj_sums + 4 * (size_t)s0));
    return 0;
}

/* Depth-first: inject the K parameters of block bi (D slots slot..slot+K-1)
 * from psi in balanced chunks of at most tmp_states states and carry each
 * chunk to the meeting point, forward through blocks bi+1 .. meet or
 * (backward) through the inverses of blocks bi-1 .. meet+1, so that it ends
 * in its D slot, in split-plane form when the plan says so. */
static int run_chunks(struct qgtl_plan *p, int bi, int slot, int K, const void *psi, qgtl_stats_t *st, int backward,
                      int lanes, void *tmp)
{
    int flips_after = 0;
    if (!backward)
        for (int b2 = bi + 1; b2 <= p->meet; b2++) flips_after += block_flips(p, b2, 0);
    else
        for (int b2 = bi - 1; b2 > p->meet; b2--) flips_after += block_flips(p, b2, 1);
    const int direct = backward ? bi == p->meet + 1 : bi == p->meet;   /* nothing to carry */
    /* a carried chunk is formed by its first pass while loading, when that is a register pass */
    int fuse = !direct && !p->tune.no_fused_inject && p->inj_slots_dev && p->n <= 28;
    if (fuse) {
        const qgtl_block_t *fb = &p->blocks[backward ? bi - 1 : bi + 1];
        const int np = backward ? fb->ninv : fb->npasses;
        fuse = np >= 1 && qgtl_hip_pass_split_ok(backward ? &fb->ipass[0] : &fb->pass[0]);
    }
    const int nch = (K + (int)p->tmp_states - 1) / (int)p->tmp_states;
    const int C = (K + nch - 1) / nch;
    for (int j0 = 0; j0 < K; j0 += C) {
        const int cnt = (K - j0 < C) ? K - j0 : C;
        void *dslot = (char *)p->D + (size_t)(slot + j0) * p->N * 8;
        void *c_cur = (flips_after % 2 == 0) ? dslot : tmp;
        void *c_alt = (c_cur == dslot) ? tmp : dslot;
        int rc = inject(p, slot + j0, cnt, psi, c_cur, st, p->presplit && direct, backward, !fuse);
        if (rc) return rc;
        const void *ipsi = fuse ? psi : NULL;
        const void *islots = fuse ? (const char *)p->inj_slots_dev +
                                        (size_t)((backward ? p->P : 0) + slot + j0) * sizeof(qgtl_inject_slot_t)
                                  : NULL;
        /* psi may be overwritten once its last reader of this chunk is issued */
        if (lanes && !fuse) HIPCALL(qgtl_hip_lane_signal(p->ev_inj));
        if (!backward) {
            for (int b2 = bi + 1; b2 <= p->meet; b2++) {
                rc = apply_block(p, &p->blocks[b2], &c_cur, &c_alt, cnt, st, p->presplit && b2 == p->meet,
                                 b2 == bi + 1 ? ipsi : NULL, islots);
                if (rc) return rc;
                if (lanes && fuse && b2 == bi + 1) HIPCALL(qgtl_hip_lane_signal(p->ev_inj));
            }
        } else {
            for (int b2 = bi - 1; b2 > p->meet; b2--) {
                rc = run_passes(p, &p->blocks[b2], 1, &c_cur, &c_alt, cnt, st, p->presplit && b2 == p->meet + 1,
                                b2 == bi - 1 ? ipsi : NULL, islots);
                if (rc) return rc;
                if (lanes && fuse && b2 == bi - 1) HIPCALL(qgtl_hip_lane_signal(p->ev_inj));
            }
        }
        if (c_cur != dslot) return qgtl_set_error(QGT_ERROR_COMPUTATION_FAILED, "internal: chunk ended off its slot");
    }
    return 0;
}

/* The mixed-precision QGT.  Depth-first plans run the primal chain (psi
 * through every block) on the second stream, so that its single-state passes
 * overlap the derivative chunks on the library stream: psi_b is published by
 * ev_psi, and the chain waits on ev_inj (the last injection issued) before a
 * pass overwrites a psi buffer an injection may still read.  With profiling on,
 * everything stays on one stream (per-launch timing). */
/* Block b applied to one state from src into dst, every pass out of place
 * (src is left as it is: the chunks read it), tmp the scratch between passes. */
static int apply_block_to(struct qgtl_plan *p, const qgtl_block_t *b, const void *src, void *dst, void *tmp,
                          qgtl_stats_t *st)
{
    const int np = b->npasses;
    if (np == 0) LAUNCH(K_OTHER, qgtl_hip_copy_f32x4(dst, src, p->N / 2));
    const void *in = src;
    for (int i = 0; i < np; i++) {
        void *out = ((np - 1 - i) % 2 == 0) ? dst : tmp;
        LAUNCH(K_SWEEP, qgtl_hip_pass_c64(&b->pass[i], in, out, p->N, 1, 0, 0, pass_tabs(p, b, i, 0)));
        if (st) {
            st->sweep_launches++;
            st->sweep_state_passes += 1;
            st->sweep_bytes += (double)p->N * 16.0;
        }
        in = out;
    }
    if (b->mstage_kernels)
        for (int i = 0; i < b->nperm; i++) {
            if (b->perm_type[i] == 0)
                LAUNCH(K_SWEEP, qgtl_hip_cnot_c64(dst, p->N, 1, p->n, b->perm_a[i], b->perm_b[i]));
            else
                LAUNCH(K_SWEEP, qgtl_hip_swap_c64(dst, p->N, 1, p->n, b->perm_a[i], b->perm_b[i]));
            if (st) {
                st->sweep_launches++;
                st->sweep_state_passes += 1;
                st->sweep_bytes += (double)p->N * 16.0;
            }
        }
    return 0;
}

/* Depth-first execution with the block-boundary states kept: the whole primal
 * chain is issued on the second stream (chain[k+1] = block k of chain[k], an
 * event after each), then the chunks on the library stream, each after the
 * event of the state it reads, so the chain never waits for the chunks. */
/* A chunk the library stream carries in run_mixed_chain: block bi's parameters
 * (slots from slot), forward from chain[bi + 1] or backward from chain[bi],
 * after the event that publishes that state. */
static int chain_lane0_chunk(struct qgtl_plan *p, int bi, int slot, qgtl_stats_t *st)
{
    const int K = p->blocks[bi].param_end - p->blocks[bi].param_begin;
    if (K <= 0) return 0;
    const int backward = bi > p->meet;
    const int src = backward ? bi : bi + 1;
    HIPCALL(qgtl_hip_set_lane(0));
    HIPCALL(qgtl_hip_lane_wait(p->ev_chain[src]));
    return run_chunks(p, bi, slot, K, p->chain[src], st, backward, 0, p->tmp);
}

/* in_order: issue the library stream's chunks as soon as the chain state they
 * read is issued (eager launches: each stream starts its work as early as it
 * can); otherwise the chain's stream first, then the library stream's (the
 * order a captured graph replays with the two streams' work side by side:
 * captured in the first order, the graph ran most of both on one queue). */
/* The library stream's chunks that read chain[bi + 1]: block bi's carried
 * forward, block bi + 1's carried back (unless the chain's stream takes it). */
static int chain_lane0_after(struct qgtl_plan *p, int bi, int slot, int K, int side, qgtl_stats_t *st)
{
    int rc;
    if (bi <= p->meet) {
        rc = chain_lane0_chunk(p, bi, slot, st);
        if (rc) return rc;
    }
    const int nb = bi + 1;
    if (nb < p->nblocks && nb > p->meet && !(side && nb > p->meet + 1)) {
        rc = chain_lane0_chunk(p, nb, slot + K, st);
        if (rc) return rc;
    }
    return 0;
}

static int run_mixed_chain(struct qgtl_plan *p, qgtl_stats_t *st, int in_order)
{
    int rc;
    /* with a second chunk buffer and a spare event, the chain's stream goes on
     * with the chunks carried backward from beyond the meeting point (their
     * primal states are the chain's, issued before them on the same stream);
     * the library stream takes the forward chunks and the meeting point's
     * injections, so the two halves of the chunk work overlap */
    const int side = p->tmp_side && p->nev_chain > p->nchain;
    /* the chain starts after what the library stream holds (the phase-table upload) */
    HIPCALL(qgtl_hip_lane_signal(p->ev_inj));
    HIPCALL(qgtl_hip_set_lane(1));
    HIPCALL(qgtl_hip_lane_wait(p->ev_inj));
    LAUNCH(K_OTHER, qgtl_hip_fill_zero_state_c64(p->chain[0], p->n));
    HIPCALL(qgtl_hip_lane_signal(p->ev_chain[0]));
    /* in_order: after block bi of the chain, the library stream's chunks that
     * read chain[bi + 1] */
    int slot = 0;
    for (int bi = 0; bi < p->nblocks; bi++) {
        const int K = p->blocks[bi].param_end - p->blocks[bi].param_begin;
        HIPCALL(qgtl_hip_set_lane(1));
        rc = apply_block_to(p, &p->blocks[bi], p->chain[bi], p->chain[bi + 1], p->chain_tmp, st);
        if (rc) return rc;
        if (p->psi_in_gram && bi == p->meet) {
            /* the primal state at the meeting point joins the Gram as state P */
            void *slotP = (char *)p->D + (size_t)p->P * p->N * 8;
            if (p->presplit) LAUNCH(K_OTHER, qgtl_hip_c64_to_split(p->chain[bi + 1], slotP, p->n));
            else LAUNCH(K_OTHER, qgtl_hip_copy_f32x4(slotP, p->chain[bi + 1], p->N / 2));
        }
        HIPCALL(qgtl_hip_lane_signal(p->ev_chain[bi + 1]));
        if (in_order) {
            rc = chain_lane0_after(p, bi, slot, K, side, st);
            if (rc) return rc;
        }
        slot += K;
    }
    if (side) {
        HIPCALL(qgtl_hip_set_lane(1));
        slot = 0;
        for (int bi = 0; bi < p->nblocks; bi++) {
            const int K = p->blocks[bi].param_end - p->blocks[bi].param_begin;
            if (bi > p->meet + 1 && K > 0) {
                rc = run_chunks(p, bi, slot, K, p->chain[bi], st, 1, 0, p->tmp_side);
                if (rc) return rc;
            }
            slot += K;
        }
        HIPCALL(qgtl_hip_lane_signal(p->ev_chain[p->nchain]));
    }
    if (!in_order) {
        slot = 0;
        for (int bi = 0; bi < p->nblocks; bi++) {
            const int K = p->blocks[bi].param_end - p->blocks[bi].param_begin;
            rc = chain_lane0_after(p, bi, slot, K, side, st);
            if (rc) return rc;
            slot += K;
        }
    }
    HIPCALL(qgtl_hip_set_lane(0));
    if (side) HIPCALL(qgtl_hip_lane_wait(p->ev_chain[p->nchain]));
    /* the library stream takes over psi_final and slot P */
    HIPCALL(qgtl_hip_lane_wait(p->ev_chain[p->nblocks]));
    p->psi_final = p->chain[p->nblocks];
    return 0;
}

static int run_mixed_lanes(struct qgtl_plan *p, qgtl_stats_t *st, int lanes)
{
    if (lanes && p->nchain == p->nblocks + 1 && p->nev_chain >= p->nchain) return run_mixed_chain(p, st, !p->capturing);
    int rc;
    void *cur = p->psi[0], *alt = p->psi[1];
    const int depth_first = any_flip(p);
    if (lanes) {
        /* the chain starts after what the library stream holds (the phase-table upload) */
        HIPCALL(qgtl_hip_lane_signal(p->ev_inj));
        HIPCALL(qgtl_hip_set_lane(1));
        HIPCALL(qgtl_hip_lane_wait(p->ev_inj));
    }
    LAUNCH(K_OTHER, qgtl_hip_fill_zero_state_c64(cur, p->n));
    int slot = 0;
    for (int bi = 0; bi < p->nblocks; bi++) {
        const qgtl_block_t *b = &p->blocks[bi];
        const int K = b->param_end - b->param_begin;
        if (depth_first && bi > p->meet && K > 0) {
            if (lanes) {                                     /* cur: the state before block bi */
                HIPCALL(qgtl_hip_set_lane(0));
                HIPCALL(qgtl_hip_lane_wait(p->ev_psi));
            }
            rc = run_chunks(p, bi, slot, K, cur, st, 1, lanes, p->tmp);
            if (rc) return rc;
        }
        if (!depth_first && slot > 0) {
            void *dc = p->D, *da = NULL;
            rc = apply_block(p, b, &dc, &da, slot, st, 0, NULL, NULL);
            if (rc) return rc;
        }
        if (lanes) {
            HIPCALL(qgtl_hip_set_lane(1));
            HIPCALL(qgtl_hip_lane_wait(p->ev_inj));
        }
        rc = apply_block(p, b, &cur, &alt, 1, st, 0, NULL, NULL);
        if (rc) return rc;
        if (p->psi_in_gram && bi == p->meet) {
            /* the primal state at the meeting point joins the Gram as state P */
            void *slotP = (char *)p->D + (size_t)p->P * p->N * 8;
            if (p->presplit) LAUNCH(K_OTHER, qgtl_hip_c64_to_split(cur, slotP, p->n));
            else LAUNCH(K_OTHER, qgtl_hip_copy_f32x4(slotP, cur, p->N / 2));
        }
        if (lanes) HIPCALL(qgtl_hip_lane_signal(p->ev_psi));
        if (K > 0 && !depth_first) {
            for (int j0 = 0; j0 < K; j0 += QGTL_INJECT_MAX) {
                const int cnt = (K - j0 < QGTL_INJECT_MAX) ? K - j0 : QGTL_INJECT_MAX;
                rc = inject(p, slot + j0, cnt, cur, (char *)p->D + (size_t)(slot + j0) * p->N * 8, st, 0, 0, 1);
                if (rc) return rc;
            }
        } else if (K > 0 && bi <= p->meet) {
            if (lanes) {
                HIPCALL(qgtl_hip_set_lane(0));
                HIPCALL(qgtl_hip_lane_wait(p->ev_psi));
            }
            rc = run_chunks(p, bi, slot, K, cur, st, 0, lanes, p->tmp);
            if (rc) return rc;
        }
        slot += K;
    }
    if (lanes) {
        /* the library stream takes over psi_final (and everything after it) */
        HIPCALL(qgtl_hip_set_lane(0));
        HIPCALL(qgtl_hip_lane_wait(p->ev_psi));
    }
    p->psi_final = cur;
    return 0;
}

/* Everything before the launches of a mixed execute: device buffers, phase
 * tables and injection slots uploaded, the schedule chosen, the events of the
 * two lanes made (none of it may happen under graph capture). */
static int prepare_mixed(struct qgtl_plan *p)
{
    int rc = alloc_mixed(p);
    if (rc) return rc;
    rc = upload_tabs(p);
    if (rc) return rc;
    const int depth_first = any_flip(p);
    p->meet = choose_meet(p, depth_first);
    p->presplit = presplit_ok(p, depth_first, gram_terms(p));
    p->psi_in_gram = depth_first && p->P > 0 && p->N >= 64;
    if (depth_first && !p->ev_psi) {
        HIPCALL(qgtl_hip_event_create_sync(&p->ev_psi));
        HIPCALL(qgtl_hip_event_create_sync(&p->ev_inj));
    }
    /* one event per block boundary of the chain, one for the end of the chunks on its stream */
    if (p->nchain && p->nev_chain < p->nchain + 1) {
        qgtl_hip_event_t *ev = (qgtl_hip_event_t *)realloc(p->ev_chain, (size_t)(p->nchain + 1) * sizeof *ev);
        if (!ev) return QGT_ERROR_NO_MEMORY;
        p->ev_chain = ev;
        while (p->nev_chain < p->nchain + 1) {
            HIPCALL(qgtl_hip_event_create_sync(&p->ev_chain[p->nev_chain]));
            p->nev_chain++;
        }
    }
    return 0;
}

/* The launches of a mixed execute (prepare_mixed done). */
static int run_mixed(struct qgtl_plan *p, qgtl_stats_t *st)
{
    const int lanes = any_flip(p) && !g_profiling;
    int rc = run_mixed_lanes(p, st, lanes);
    if (lanes) qgtl_hip_set_lane(0);
    if (rc) return rc;
    if (p->P == 0) return 0;
    const int Pg = p->P + (p->psi_in_gram ? 1 : 0);
    /* split-plane states (the primal state among them): the complex-form Gram */
    p->gram_cplx = p->presplit;
    if (p->gram_cplx) {
        /* stages of 32 amplitudes (two LDS buffers): at n = 20, P = 240 the Gram took
         * 0.95-0.97 ms, with stages of 16 (four buffers) 0.98-1.02 ms */
        LAUNCH(K_GRAM_MAIN, qgtl_hip_gram_cplx(p->D, Pg, p->N, p->gram_partials, p->gram_chunks_cplx,
                                               gram_products(p), 32));
        LAUNCH(K_GRAM_AUX, qgtl_hip_gram_cplx_reduce(p->gram_partials, Pg, p->gram_chunks_cplx, p->C));
        LAUNCH(K_FINAL, qgtl_hip_qgt_finalize_cplx(p->C, p->P, p->Qdev));
    } else {
        LAUNCH(K_GRAM_MAIN, qgtl_hip_gram_bf16x3(p->D, Pg, p->N, p->gram_partials, p->gram_chunks, gram_terms(p)));
        LAUNCH(K_GRAM_AUX, qgtl_hip_gram_reduce(p->gram_partials, Pg, p->gram_chunks, p->C));
        if (p->psi_in_gram) LAUNCH(K_FINAL, qgtl_hip_qgt_finalize_psi(p->C, p->P, p->Qdev));
        else LAUNCH(K_FINAL, qgtl_hip_qgt_finalize(p->C, p->inj_sums, p->P, p->Qdev));
    }
    if (st) {
        st->gram_launches += 2;
        st->gram_flops += 4.0 * (double)p->N * p->P * (p->P + 1);
    }
    return 0;
}

/* Carry one complex64 state from the end of the circuit back to the meeting
 * point of the last execution (the inverses of blocks nblocks-1 .. meet+1);
 * *state and *scratch may swap. */
int qgtl_carry_to_meet(struct qgtl_plan *p, void **state, void **scratch)
{
    qgtl_stats_t *st = NULL;
    for (int b = p->nblocks - 1; b > p->meet; b--) {
        if (!p->blocks[b].inv_ok) return qgtl_set_error(QGT_ERROR_INVALID_STATE, "block %d has no inverse", b);
        int rc = run_passes(p, &p->blocks[b], 1, state, scratch, 1, st, 0, NULL, NULL);
        if (rc) return rc;
    }
    return 0;
}

/* complex128 op application for the plain path */
static int plain_apply(struct qgtl_plan *p, void *psi, const qgtl_op_t *op, qgtl_stats_t *st)
{
    double m8[8];
    qgtl_mat2 m;
    qgtl_op_t base = {GATE_TYPE_X, op->q1, -1, -1, 0.0};
    switch (op->type) {
    case GATE_TYPE_CNOT:
    case GATE_TYPE_CZ:
    case GATE_TYPE_CY:
        base.type = op->type == GATE_TYPE_CNOT ? GATE_TYPE_X : op->type == GATE_TYPE_CZ ? GATE_TYPE_Z : GATE_TYPE_Y;
        qgtl_gate_matrix(&base, &m);
        for (int e = 0; e < 4; e++) { m8[2 * e] = m.m[e].re; m8[2 * e + 1] = m.m[e].im; }
        LAUNCH(K_SWEEP, qgtl_hip_cgate_c128(psi, p->n, op->q0, op->q1, m8));
        break;
    case GATE_TYPE_SWAP:
        LAUNCH(K_SWEEP, qgtl_hip_swap_c128(psi, p->n, op->q0, op->q1));
        break;
    default:
        if (qgtl_gate_matrix(op, &m) != 0) return qgtl_set_error(QGT_ERROR_NOT_IMPLEMENTED, "gate type %d", op->type);
        for (int e = 0; e < 4; e++) { m8[2 * e] = m.m[e].re; m8[2 * e + 1] = m.m[e].im; }
        LAUNCH(K_SWEEP, qgtl_hip_gate1q_c128(psi, p->n, op->q0, m8));
    }
    if (st) {
        st->sweep_launches++;
        st->sweep_state_passes++;
        st->sweep_bytes += (double)p->N * 32.0;
    }
    return 0;
}

static int run_plain(struct qgtl_plan *p, qgtl_stats_t *st)
{
    int rc = alloc_plain(p);
    if (rc) return rc;
    void *psi = p->psi[0];
    LAUNCH(K_OTHER, qgtl_hip_fill_zero_state_c128(psi, p->n));
    for (int g = 0; g < p->nops; g++) {
        rc = plain_apply(p, psi, &p->ops[g], st);
        if (rc) return rc;
        const int k = p->ops[g].param;
        if (k < 0) continue;
        void *d = (char *)p->D + (size_t)k * p->N * 16;
        LAUNCH(K_OTHER, qgtl_hip_d2d(d, psi, p->N * 16));
        double gen[8] = {0};
        switch (p->ops[g].type) {           /* (-i/2) G */
        case GATE_TYPE_RX: gen[3] = -0.5; gen[5] = -0.5; break;
        case GATE_TYPE_RY: gen[2] = -0.5; gen[4] = 0.5; break;
        default: gen[1] = -0.5; gen[7] = 0.5; break;
        }
        LAUNCH(K_SWEEP, qgtl_hip_gate1q_c128(d, p->n, p->ops[g].q0, gen));
        for (int g2 = g + 1; g2 < p->nops; g2++) {
            rc = plain_apply(p, d, &p->ops[g2], st);
            if (rc) return rc;
        }
    }
    if (p->P == 0) return 0;
    LAUNCH(K_GRAM_AUX, qgtl_hip_dotc_c128(p->D, p->P, p->N, psi, p->plain_partials, p->inj_sums));
    LAUNCH(K_GRAM_MAIN, qgtl_hip_gram_f64(p->D, p->P, p->N, p->plain_partials, p->C));
    LAUNCH(K_FINAL, qgtl_hip_qgt_finalize_plain(p->C, p->inj_sums, p->P, p->Qdev));
    if (st) {
        st->gram_launches += 1;
        st->gram_flops += 4.0 * (double)p->N * p->P * (p->P + 1);
    }
    return 0;
}

int qgtl_plan_execute(qgtl_plan_t *p, double *Q, qgtl_stats_t *stats)
{
    QGTL_LOCKED();
    if (!p) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null plan");
    int rc = ensure_init();
    if (rc) return rc;
    qgtl_stats_t local;
    qgtl_stats_t *st = stats ? stats : &local;
    memset(st, 0, sizeof *st);
    st->P = p->P;
    st->n = p->n;
    const double t0 = now_ms();
    prof_reset();
    if (p->mode == QGTL_MODE_MIXED) {
        rc = prepare_mixed(p);
        if (rc) return rc;
        const int graphable = !g_profiling && !p->tune.no_graph && p->P > 0;
        if (graphable && p->graph && p->graph_gen == p->build_gen) {
            /* unchanged plan: replay the captured launch sequence */
            rc = qgtl_hip_graph_launch(p->graph);
            if (rc) return hip_fail(rc, "graph launch");
            *st = p->graph_st;
        } else if (graphable && p->exec_gen == p->build_gen) {
            /* second execute of this build: capture the launches, then replay them */
            HIPCALL(qgtl_hip_capture_begin());
            p->capturing = 1;
            rc = run_mixed(p, st);
            p->capturing = 0;
            if (rc) { qgtl_hip_capture_abort(); qgtl_hip_set_lane(0); return rc; }
            rc = qgtl_hip_capture_end_update(&p->graph);
            if (rc) return hip_fail(rc, "graph capture");
            p->graph_gen = p->build_gen;
            p->graph_st = *st;
            p->graph_st.graph = 1;
            rc = qgtl_hip_graph_launch(p->graph);
            if (rc) return hip_fail(rc, "graph launch");
            *st = p->graph_st;
        } else {
            rc = run_mixed(p, st);
            p->exec_gen = p->build_gen;
        }
    } else {
        rc = run_plain(p, st);
    }
    if (rc) { qgtl_hip_sync(); prof_reset(); return rc; }
    st->presplit = p->mode == QGTL_MODE_MIXED ? p->presplit : 0;
    st->meet_block = p->mode == QGTL_MODE_MIXED ? p->meet : p->nblocks - 1;
    const size_t P = (size_t)p->P;
    /* Q comes back through the plan's pinned buffer (a direct DMA copy) */
    double *Qslot = (Q && P > 0) ? p->Qhost : NULL;
    if (Qslot) {
        rc = qgtl_hip_d2h(Qslot, p->Qdev, P * P * 2 * sizeof(double));
        if (rc) return hip_fail(rc, "download Q");
    }
    rc = qgtl_hip_sync();
    if (rc) return hip_fail(rc, "qgtl_plan_execute");
    st->wall_ms = now_ms() - t0;
    rc = prof_collect(st);
    if (rc) return rc;
    if (Qslot) {
        int identity = 1;
        for (size_t a = 0; a < P && identity; a++) identity = p->slot_param[a] == (int)a;
        if (p->mode == QGTL_MODE_MIXED && !identity) {
            /* slot order (creation order) -> parameter numbers */
            for (size_t a = 0; a < P; a++)
                for (size_t b = 0; b < P; b++) {
                    const size_t ka = (size_t)p->slot_param[a], kb = (size_t)p->slot_param[b];
                    Q[2 * (ka * P + kb)] = Qslot[2 * (a * P + b)];
                    Q[2 * (ka * P + kb) + 1] = Qslot[2 * (a * P + b) + 1];
                }
        } else {
            memcpy(Q, Qslot, P * P * 2 * sizeof(double));
        }
    }
    return 0;
}

/* --------------------------------------------------------- plan cache */
#define PLAN_CACHE 4
static qgtl_plan_t *g_cache[PLAN_CACHE];
static uint64_t g_cache_tick[PLAN_CACHE];
static uint64_t g_tick = 0;

static qgtl_plan_t *cached_plan(int n, int nops, const qgtl_op_t *ops, int mode)
{
    const uint64_t h = qgtl_structure_hash(n, nops, ops, mode);
    for (int i = 0; i < PLAN_CACHE; i++) {
        qgtl_plan_t *p = g_cache[i];
        if (p && p->structure_hash == h && p->n == n && p->nops == nops && p->mode == mode) {
            int same = 1;
            for (int j = 0; j < nops && same; j++)
                same = p->ops[j].type == ops[j].type && p->ops[j].q0 == ops[j].q0 && p->ops[j].q1 == ops[j].q1 &&
                       p->ops[j].param == ops[j].param;
            if (!same) continue;
            double *ang = (double *)malloc(sizeof(double) * (size_t)(nops > 0 ? nops : 1));
            if (!ang) return NULL;
            for (int j = 0; j < nops; j++) ang[j] = ops[j].angle;
            int rc = qgtl_plan_set_angles(p, nops, ang);
            free(ang);
            if (rc) return NULL;
            g_cache_tick[i] = ++g_tick;
            return p;
        }
    }
    qgtl_plan_t *p = qgtl_plan_create(n, nops, ops, mode, NULL);
    if (!p) return NULL;
    int victim = 0;
    for (int i = 0; i < PLAN_CACHE; i++) {
        if (!g_cache[i]) { victim = i; break; }
        if (g_cache_tick[i] < g_cache_tick[victim]) victim = i;
    }
    qgtl_plan_destroy(g_cache[victim]);
    g_cache[victim] = p;
    g_cache_tick[victim] = ++g_tick;
    return p;
}

void qgtl_release_cache(void)
{
    QGTL_LOCKED();
    for (int i = 0; i < PLAN_CACHE; i++) {
        qgtl_plan_destroy(g_cache[i]);
        g_cache[i] = NULL;
    }
}

int qgtl_qgt(int n, int nops, const qgtl_op_t *ops, int mode, double *Q, qgtl_stats_t *stats)
{
    QGTL_LOCKED();
    int rc = ensure_init();
    if (rc) return rc;
    qgtl_plan_t *p = cached_plan(n, nops, ops, mode);
    if (!p) return QGT_ERROR_INVALID_ARGUMENT;   /* message already set */
    return qgtl_plan_execute(p, Q, stats);
}

int qgtl_statevector(int n, int nops, const qgtl_op_t *ops, int mode, double *psi_out)
{
    QGTL_LOCKED();
    int rc = ensure_init();
    if (rc) return rc;
    qgtl_plan_t *p = qgtl_plan_create(n, nops, ops, mode, NULL);
    if (!p) return QGT_ERROR_INVALID_ARGUMENT;
    qgtl_stats_t *st = NULL;
    const size_t N = (size_t)1 << n;
    void *out = NULL;
    rc = qgtl_hip_malloc(&out, N * 16);
    if (rc) { qgtl_plan_destroy(p); return hip_fail(rc, "statevector buffer"); }
    if (mode == QGTL_MODE_MIXED) {
        rc = qgtl_hip_malloc(&p->psi[0], N * 8);
        if (!rc) rc = qgtl_hip_malloc(&p->psi[1], N * 8);
        if (rc) { qgtl_hip_free(out); qgtl_plan_destroy(p); return hip_fail(rc, "statevector buffer"); }
        void *cur = p->psi[0], *alt = p->psi[1];
        rc = qgtl_hip_fill_zero_state_c64(cur, n);
        for (int b = 0; !rc && b < p->nblocks; b++) rc = apply_block(p, &p->blocks[b], &cur, &alt, 1, st, 0, NULL, NULL);
        if (!rc) rc = qgtl_hip_c64_to_c128(cur, out, N);
    } else {
        rc = qgtl_hip_fill_zero_state_c128(out, n);
        for (int g = 0; !rc && g < nops; g++) rc = plain_apply(p, out, &ops[g], st);
    }
    if (!rc) rc = qgtl_hip_d2h(psi_out, out, N * 16);
    if (!rc) rc = qgtl_hip_sync();
    qgtl_hip_free(out);
    qgtl_plan_destroy(p);
    prof_reset();
    return rc ? (rc < -100 ? hip_fail(rc, "qgtl_statevector") : rc) : 0;
}

/* Diagnostics: run the MFMA lane-layout probe and copy its 32 x 32 output
 * to host_out (see mfma_layout_probe_kernel for the expected values). */


int qgtl_debug_mfma_layout(float *host_out)
{
    QGTL_LOCKED();
    int rc = ensure_init();
    if (rc) return rc;
    void *d = NULL;
    HIPCALL(qgtl_hip_malloc(&d, 32 * 32 * sizeof(float)));
    rc = qgtl_hip_debug_mfma_layout((float *)d);
    if (!rc) rc = qgtl_hip_d2h(host_out, d, 32 * 32 * sizeof(float));
    if (!rc) rc = qgtl_hip_sync();
    qgtl_hip_free(d);
    return rc ? hip_fail(rc, "qgtl_debug_mfma_layout") : 0;
}

int qgtl_debug_mfma_layout16(float *host_out)
{
    QGTL_LOCKED();
    int rc = ensure_init();
    if (rc) return rc;
    void *d = NULL;
    HIPCALL(qgtl_hip_malloc(&d, 16 * 16 * sizeof(float)));
    rc = qgtl_hip_debug_mfma16_layout((float *)d);
    if (!rc) rc = qgtl_hip_d2h(host_out, d, 16 * 16 * sizeof(float));
    if (!rc) rc = qgtl_hip_sync();
    qgtl_hip_free(d);
    return rc ? hip_fail(rc, "qgtl_debug_mfma_layout16") : 0;
}

/* Diagnostics: after qgtl_plan_execute in mixed mode, copy the FP64 Gram
 * C = Z Z^T (R x R, R = 128 ceil(2P/128), slot order) and the injection sums
 * (P x 4: Re/Im <d|psi>, |d|^2, |psi|^2, slot order). */
int qgtl_plan_debug_buffers(qgtl_plan_t *p, double *C_out, double *inj_out, int *R_out)
{
    QGTL_LOCKED();
    if (!p || p->mode != QGTL_MODE_MIXED || !p->C) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "no mixed-mode result");
    /* depth-first plans: C is the Gram of P + 1 states (the last one the primal
     * state) and the injection sums are not computed; after the complex-form
     * Gram C holds G (R x R complex128, R_out = -R), else the real form (R x R) */
    const int Pg = p->P + (p->psi_in_gram ? 1 : 0);
    const size_t R = p->gram_cplx ? (size_t)((Pg + 127) / 128) * 128 : (size_t)((2 * Pg + 127) / 128) * 128;
    if (R_out) *R_out = p->gram_cplx ? -(int)R : (int)R;
    if (C_out) HIPCALL(qgtl_hip_d2h(C_out, p->C, R * R * (p->gram_cplx ? 2 : 1) * sizeof(double)));
    if (inj_out) HIPCALL(qgtl_hip_d2h(inj_out, p->inj_sums, (size_t)p->P * 4 * sizeof(double)));
    HIPCALL(qgtl_hip_sync());
    return 0;
}

/* The overlaps a_k = <d_k psi|psi> (2P doubles: Re, Im, parameter order) that
 * the last mixed-mode execute subtracted (Q_kl = G_kl - a_k conj(a_l)), with
 * the same normalisation by ||d_k psi|| = 1/2 and ||psi|| = 1: from the FP64
 * reduction of the Gram's row of the primal state on depth-first plans, from
 * the FP64 injection sums otherwise. */
int qgtl_plan_overlaps(qgtl_plan_t *p, double *a)
{
    QGTL_LOCKED();
    if (!p || !a) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qgtl_plan_overlaps: null argument");
    if (p->mode != QGTL_MODE_MIXED || !p->C || p->P < 1)
        return qgtl_set_error(QGT_ERROR_INVALID_STATE, "qgtl_plan_overlaps: no mixed-mode result");
    const size_t P = (size_t)p->P;
    if (p->gram_cplx) {
        /* complex Gram of P + 1 states: a_k = G_kP / (2 sqrt(G_kk G_PP)) */
        const size_t R = ((P + 1 + 127) / 128) * 128;
        double *G = (double *)malloc(R * R * 2 * sizeof(double));
        if (!G) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        int rc = qgtl_hip_d2h(G, p->C, R * R * 2 * sizeof(double));
        if (!rc) rc = qgtl_hip_sync();
        if (rc) { free(G); return hip_fail(rc, "qgtl_plan_overlaps"); }
        const double gmm = G[2 * (P * R + P)];
        for (size_t s = 0; s < P; s++) {
            const double gkk = G[2 * (s * R + s)];
            const double sc = 0.5 / sqrt(gkk * gmm);
            const size_t k = (size_t)p->slot_param[s];
            a[2 * k] = G[2 * (s * R + P)] * sc;
            a[2 * k + 1] = G[2 * (s * R + P) + 1] * sc;
        }
        free(G);
        return 0;
    }
    if (p->psi_in_gram) {
        const size_t R = ((2 * (P + 1) + 127) / 128) * 128;
        double *C = (double *)malloc(R * R * sizeof(double));
        if (!C) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        int rc = qgtl_hip_d2h(C, p->C, R * R * sizeof(double));
        if (!rc) rc = qgtl_hip_sync();
        if (rc) { free(C); return hip_fail(rc, "qgtl_plan_overlaps"); }
        const size_t m = P;
        const double gmm = C[(2 * m) * R + 2 * m] + C[(2 * m + 1) * R + 2 * m + 1];
        for (size_t s = 0; s < P; s++) {
            const double gkk = C[(2 * s) * R + 2 * s] + C[(2 * s + 1) * R + 2 * s + 1];
            const double re = C[(2 * s) * R + 2 * m] + C[(2 * s + 1) * R + 2 * m + 1];
            const double im = C[(2 * s) * R + 2 * m + 1] - C[(2 * s + 1) * R + 2 * m];
            const double sc = 0.5 / sqrt(gkk * gmm);
            const size_t k = (size_t)p->slot_param[s];
            a[2 * k] = re * sc;
            a[2 * k + 1] = im * sc;
        }
        free(C);
        return 0;
    }
    double *inj = (double *)malloc(P * 4 * sizeof(double));
    if (!inj) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
    int rc = qgtl_hip_d2h(inj, p->inj_sums, P * 4 * sizeof(double));
    if (!rc) rc = qgtl_hip_sync();
    if (rc) { free(inj); return hip_fail(rc, "qgtl_plan_overlaps"); }
    for (size_t s = 0; s < P; s++) {
        const double sc = 0.5 / sqrt(inj[4 * s + 2] * inj[4 * s + 3]);
        const size_t k = (size_t)p->slot_param[s];
        a[2 * k] = inj[4 * s] * sc;
        a[2 * k + 1] = inj[4 * s + 1] * sc;
    }
    free(inj);
    return 0;
}

/* Diagnostics: one plain copy of `bytes` (a multiple of 16) between two
 * fresh buffers, the known-size reference for HBM traffic counters. */
int qgtl_debug_copy_kernel(size_t bytes, double *ms)
{
    QGTL_LOCKED();
    if (bytes == 0 || bytes % 16) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "copy size must be a multiple of 16");
    int rc = ensure_init();
    if (rc) return rc;
    void *a = NULL, *b = NULL;
    HIPCALL(qgtl_hip_malloc(&a, bytes));
    rc = qgtl_hip_malloc(&b, bytes);
    qgtl_hip_event_t e0 = NULL, e1 = NULL;
    if (!rc) rc = qgtl_hip_event_create(&e0);
    if (!rc) rc = qgtl_hip_event_create(&e1);
    if (!rc) rc = qgtl_hip_fill_f32(a, bytes / 4, 1.0f);
    if (!rc) rc = qgtl_hip_event_record(e0);
    if (!rc) rc = qgtl_hip_copy_f32x4(b, a, bytes / 16);
    if (!rc) rc = qgtl_hip_event_record(e1);
    if (!rc) rc = qgtl_hip_sync();
    float t = 0.f;
    if (!rc) rc = qgtl_hip_event_elapsed_ms(e0, e1, &t);
    qgtl_hip_event_destroy(e0);
    qgtl_hip_event_destroy(e1);
    qgtl_hip_free(a);
    qgtl_hip_free(b);
    if (rc) return hip_fail(rc, "qgtl_debug_copy_kernel");
    if (ms) *ms = t;
    return 0;
}

unsigned long long qgtl_debug_alloc_calls(void) { return qgtl_hip_alloc_calls(); }

int qgtl_init_if_needed(void)
{
    QGTL_LOCKED();
    return ensure_init();
}
