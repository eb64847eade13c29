/*
 * qgtl_batch.c -- many independent mixed-precision QGTs in one captured
 * launch sequence.
 *
 * Every instance is compiled by the single-instance plan builder (blocks,
 * passes and inverse passes, injection operators).  The batch then lays all
 * instances out in a few device arenas and advances them together,
 * breadth-first, on the two-sided schedule of the single plans: each instance
 * has a meeting block c (the one that minimises its state-passes).  Forward
 * phase, block step b: ONE grouped launch injects, for every instance with
 * b > c, the states of block b at the block's input (from the primal before
 * the block, with the block's inverse folded into the 2 x 2); for pass j one
 * grouped launch sweeps the primal and, for instances with b <= c, every
 * forward derivative state; one grouped launch injects the new forward states
 * of instances with b <= c.  Backward phase, block step b' descending: one
 * grouped launch per inverse pass carries every instance's states of blocks
 * > b' (b' > c) back through block b'.  All states then sit after their
 * instance's block c, in one arena, beside a copy of the primal after block c
 * (slot P + 1, made by an identity injection), and one grouped launch per Gram
 * precision class over the P + 1 states, one split-K reduction and one FP64
 * correction finish every Q: the overlaps <d_k|psi> and the norms are the
 * Gram's primal row and diagonal, so the injections write states only (the
 * inner products are the same after any common unitary).  The primal is slot
 * 0 of an instance's state array and derivative state s is slot 1 + s (forward
 * states first), so one pass item covers a contiguous range of an instance's
 * live states.  Launch grids are
 * persistent (a few workgroups per CU looping over units) and the whole
 * sequence is captured once in a HIP graph; set_angles only rewrites the
 * descriptors the graph reads.
 *
 * QGTL evaluates one circuit per call and one QGT element per call
 * (src/quantum_geometric/core/quantum_geometric_curvature.c:289-320); the
 * batch has no counterpart there and returns, per instance, the matrix
 * geometric_compute_full_qgt would.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qgtl_internal.h"

#define BTRY(call)                                                                              \
    do {                                                                                        \
        int _rc = (call);                                                                       \
        if (_rc) return qgtl_set_error(_rc, "%s: %s", #call, qgtl_hip_last_error());            \
    } while (0)

#define INJECT_TILE 2048          /* amplitudes per injection / fill unit */
#define GRAM_TILE 128
#define GRAM_KS QGTL_GRAM_KSTEP
#define XCDS 8

enum { BL_FILL, BL_PASS, BL_INJECT, BL_GRAM, BL_GRAM_REDUCE, BL_FINAL, BL_GRAM_CPLX, BL_GRAM_REDUCE_CPLX, BL_FINAL_CPLX };

#define CPLX_CHUNK 1024           /* amplitudes per split-K chunk of the complex-form Gram */

typedef struct {
    int kind;
    int terms;                 /* BL_GRAM: operand split; BL_PASS: kernel variant */
    int nitems;
    uint32_t units;
    size_t item_off;           /* bytes into the item arena */
} launch_t;

typedef struct {
    qgtl_plan_t *plan;
    int n, P, nops;
    size_t N;
    size_t op_off;             /* first op of the instance in the creation arrays */
    size_t x_off, y_off;       /* complex64 elements into the state arena; y_off == SIZE_MAX: no flips */
    size_t gp_off;             /* floats into the Gram partials */
    size_t c_off;              /* doubles into C */
    size_t q_off;              /* complex elements of Q */
    int tiles;                 /* injection tiles */
    int nt, T, nchunks, terms;
    size_t chunk_len;
    int flips;                 /* any pass of the instance stores to the other buffer */
    int meet;                  /* states meet after this block (the last one: all carried forward) */
    int nfwd;                  /* derivative states of blocks <= meet (slots 1 .. nfwd) */
    int final_arena;           /* 0: x, 1: y -- where all derivative states end */
    size_t pass_off, inj_off;  /* the instance's first pass / injection descriptor */
    int cplx;                  /* final states in split-plane form and the complex-form Gram (129..256 Gram
                                  states, register final passes); else complex64 and the real-form Gram */
} inst_t;

struct qgtl_batch {
    int count;
    inst_t *inst;
    qgtl_batch_tune_t tune;
    int cus;
    /* descriptors (host mirror, device copy) */
    qgtl_pass_desc_t *pass_h, *pass_d;
    void *tabs_d;                        /* phase tables of every pass descriptor, built on the device per execution */
    size_t npass;
    qgtl_inject_desc_t *inj_h, *inj_d;
    size_t ninj;
    int descs_dirty;
    /* arenas */
    void *states;
    double *C, *Qdev;
    float *gram_partials;
    size_t gp_len;             /* floats allocated */
    size_t q_len;              /* complex elements */
    size_t device_bytes;
    /* launch sequence */
    unsigned char *items_h, *items_d;
    size_t items_bytes, items_cap;
    launch_t *launches;
    int nlaunch, launch_cap;
    qgtl_hip_graph_t graph;
    double *qhost;
    double sweep_bytes, gram_flops;
    int64_t units;
};

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static size_t upper_tiles(int nt) { return (size_t)nt * (size_t)(nt + 1) / 2; }

static int regs_variant(const qgtl_pass_desc_t *d, int pass_variant)
{
    return pass_variant != 1 && d->m == QGTL_TILE_BITS_MAX && d->nphase >= 1 && d->nphase <= QGTL_MAX_PHASES &&
           d->store != QGTL_STORE_PERM;
}

/* ------------------------------------------------------------ descriptors */
static size_t count_pass_descs(const qgtl_plan_t *p)
{
    size_t c = 0;
    for (int b = 0; b < p->nblocks; b++) c += (size_t)p->blocks[b].npasses + (size_t)p->blocks[b].ninv;
    return c;
}

/* the injection chunks of every block, then one identity "injection" that
 * copies the primal state at the meeting block into slot P + 1 */
static size_t count_inj_descs(const qgtl_plan_t *p)
{
    size_t c = 0;
    for (int b = 0; b < p->nblocks; b++) {
        const int K = p->blocks[b].param_end - p->blocks[b].param_begin;
        c += (size_t)((K + QGTL_INJECT_MAX - 1) / QGTL_INJECT_MAX);
    }
    return c + 1;
}

/* Descriptor order per instance: the passes of every block, then the inverse
 * passes of every block (block order); the injection chunks in block order,
 * forward (A after the block) for blocks <= meet, backward (B at the block's
 * input, on the parameter's qubit) for later blocks.  The schedule walks the
 * same order. */
/* the descriptors of instance i from its plan, at its fixed offsets (so that
 * instances can be filled in parallel) */
static void fill_inst_descs(struct qgtl_batch *B, int i)
{
    const inst_t *in = &B->inst[i];
    const qgtl_plan_t *p = in->plan;
    const int meet = in->meet;
    size_t ip = in->pass_off, ii = in->inj_off;
    for (int b = 0; b < p->nblocks; b++)
        for (int j = 0; j < p->blocks[b].npasses; j++) B->pass_h[ip++] = p->blocks[b].pass[j];
    for (int b = 0; b < p->nblocks; b++)
        for (int j = 0; j < p->blocks[b].ninv; j++) B->pass_h[ip++] = p->blocks[b].ipass[j];
    for (int b = 0; b < p->nblocks; b++) {
        const qgtl_block_t *blk = &p->blocks[b];
        const int backward = b > meet;
        const int K = blk->param_end - blk->param_begin;
        for (int j0 = 0; j0 < K; j0 += QGTL_INJECT_MAX) {
            const int cnt = (K - j0 < QGTL_INJECT_MAX) ? K - j0 : QGTL_INJECT_MAX;
            qgtl_inject_desc_t *d = &B->inj_h[ii++];
            memset(d, 0, sizeof *d);
            d->n = p->n;
            d->K = cnt;
            for (int j = 0; j < cnt; j++) {
                const qgtl_param_t *pp = &p->params[p->slot_param[blk->param_begin + j0 + j]];
                const qgtl_mat2 *M = backward ? &pp->B : &pp->A;
                d->v[j] = backward ? 1u << pp->qubit : pp->v;
                d->w[j] = backward ? 1u << pp->qubit : pp->w;
                for (int e = 0; e < 4; e++) {
                    d->A[j][e].re = (float)M->m[e].re;
                    d->A[j][e].im = (float)M->m[e].im;
                }
            }
        }
    }
    qgtl_inject_desc_t *d = &B->inj_h[ii];          /* the primal copy: d[y] = psi[y] */
    memset(d, 0, sizeof *d);
    d->n = p->n;
    d->K = 1;
    d->A[0][0].re = 1.f;
    d->A[0][3].re = 1.f;
}

static void fill_descs(struct qgtl_batch *B)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < B->count; i++) fill_inst_descs(B, i);
    B->descs_dirty = 1;
}

/* ------------------------------------------------------------ schedule */
static int push_items(struct qgtl_batch *B, const void *items, size_t bytes, int kind, int terms, int nitems,
                      uint64_t units)
{
    if (nitems == 0) return 0;
    if (units >= 0x80000000ull)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "batch launch of %llu units is too large; split the batch",
                              (unsigned long long)units);
    const size_t off = (B->items_bytes + 63) & ~(size_t)63;
    if (off + bytes > B->items_cap) {
        size_t cap = B->items_cap ? B->items_cap : 4096;
        while (cap < off + bytes) cap *= 2;
        unsigned char *h = (unsigned char *)realloc(B->items_h, cap);
        if (!h) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        B->items_h = h;
        B->items_cap = cap;
    }
    memcpy(B->items_h + off, items, bytes);
    B->items_bytes = off + bytes;
    if (B->nlaunch == B->launch_cap) {
        const int cap = B->launch_cap ? 2 * B->launch_cap : 64;
        launch_t *l = (launch_t *)realloc(B->launches, sizeof(launch_t) * (size_t)cap);
        if (!l) return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
        B->launches = l;
        B->launch_cap = cap;
    }
    launch_t *L = &B->launches[B->nlaunch++];
    L->kind = kind;
    L->terms = terms;
    L->nitems = nitems;
    L->units = (uint32_t)units;
    L->item_off = off;
    B->units += (int64_t)units;
    return 0;
}

static uint64_t dev_addr(const void *base, size_t byte_off) { return (uint64_t)(uintptr_t)base + byte_off; }

/* Gram chunking of instance i for the current tune. */
/* The final pass of instance i's forward group (block meet's last pass) and of its
 * backward group (block meet + 1's last inverse pass) can store the split-plane form. */
static int split_final_ok(const struct qgtl_batch *B, const inst_t *in)
{
    const qgtl_plan_t *p = in->plan;
    const qgtl_block_t *bm = &p->blocks[in->meet];
    if (bm->npasses < 1 || !regs_variant(&bm->pass[bm->npasses - 1], B->tune.pass_variant)) return 0;
    if (in->meet < p->nblocks - 1) {
        const qgtl_block_t *bb = &p->blocks[in->meet + 1];
        if (bb->ninv < 1 || !regs_variant(&bb->ipass[bb->ninv - 1], B->tune.pass_variant)) return 0;
    }
    return 1;
}

static void gram_shape(const struct qgtl_batch *B, inst_t *in)
{
    const int states = in->P + 1;
    in->cplx = (in->terms == QGTL_GRAM_F16X3 || in->terms == QGTL_GRAM_F16X4) && states > GRAM_TILE &&
               states <= 2 * GRAM_TILE && in->n >= 14 && split_final_ok(B, in);
    if (in->cplx) {
        in->nchunks = (int)((in->N + CPLX_CHUNK - 1) / CPLX_CHUNK);
        in->chunk_len = CPLX_CHUNK;
        in->nt = 2;
        in->T = 3;
        return;
    }
    const size_t want = B->tune.gram_chunk_len > 0 ? (size_t)B->tune.gram_chunk_len : 16384;
    size_t nch = (in->N + want - 1) / want;
    nch = (nch + XCDS - 1) / XCDS * XCDS;
    if (nch < XCDS) nch = XCDS;
    in->nchunks = (int)nch;
    in->chunk_len = ((in->N + nch - 1) / nch + GRAM_KS - 1) / GRAM_KS * GRAM_KS;
    in->nt = (2 * (in->P + 1) + GRAM_TILE - 1) / GRAM_TILE;
    in->T = (int)upper_tiles(in->nt);
}

static int block_flips(const qgtl_block_t *b, int inv)
{
    int f = 0;
    const int np = inv ? b->ninv : b->npasses;
    for (int j = 0; j < np; j++) f += inv ? b->ipass_flips[j] : b->pass_flips[j];
    return f;
}

/* Arena (0: x, 1: y) of instance i's backward group before it is carried
 * through block b (meet <= b <= nblocks - 2), i.e. where the states of block
 * b + 1 are injected; final_arena for b = meet. */
static int bwd_arena(const inst_t *in, int b)
{
    const qgtl_plan_t *p = in->plan;
    int a = in->final_arena;
    for (int b2 = in->meet + 1; b2 <= b; b2++) a ^= block_flips(&p->blocks[b2], 1) & 1;
    return a;
}

static size_t arena_off(const inst_t *in, int a) { return a ? in->y_off : in->x_off; }

typedef struct {
    int psi;                   /* arena of the primal */
    int bwd;                   /* arena of the backward group during the backward phase */
    int fwd_slots;             /* forward derivative states created so far */
    size_t pass_desc, inv_desc, inj_desc;
} sched_t;

/* Pass items of one launch: (instance, pass descriptor, first state slot,
 * state count, source / destination arena); the arena flips when the pass
 * stores to the other buffer. */
static int add_pass_item(struct qgtl_batch *B, qgtl_pass_item_t *it, int *k, uint64_t *u, int i, size_t desc,
                         int flip, int slot0, int nstates, int *arena)
{
    const inst_t *in = &B->inst[i];
    if (flip && in->y_off == SIZE_MAX)
        return qgtl_set_error(QGT_ERROR_COMPUTATION_FAILED, "internal: flip without a second buffer");
    const qgtl_pass_desc_t *d = &B->pass_h[desc];
    const size_t c64 = 8;
    it[*k].desc = dev_addr(B->pass_d, desc * sizeof(qgtl_pass_desc_t));
    it[*k].src = dev_addr(B->states, (arena_off(in, *arena) + (size_t)slot0 * in->N) * c64);
    it[*k].dst = dev_addr(B->states, (arena_off(in, *arena ^ flip) + (size_t)slot0 * in->N) * c64);
    it[*k].stride = in->N;
    it[*k].ntiles = 1u << d->nrest;
    it[*k].nstates = (uint32_t)nstates;
    it[*k].unit_begin = (uint32_t)*u;
    it[*k].variant = (uint32_t)regs_variant(d, B->tune.pass_variant);
    it[*k].tabs = dev_addr(B->tabs_d, desc * 2 * QGTL_PHASE_TAB * sizeof(qgtl_c64_t));
    *u += (uint64_t)it[*k].ntiles * (uint64_t)nstates;
    B->sweep_bytes += (double)nstates * (double)in->N * 16.0;
    *arena ^= flip;
    (*k)++;
    return 0;
}

static int build_schedule(struct qgtl_batch *B)
{
    const int cnt = B->count;
    B->items_bytes = 0;
    B->nlaunch = 0;
    B->units = 0;
    B->sweep_bytes = 0.0;
    B->gram_flops = 0.0;
    sched_t *sc = (sched_t *)calloc((size_t)cnt, sizeof(sched_t));
    size_t bytes = sizeof(qgtl_pass_item_t) * (size_t)cnt * 2;
    if (bytes < sizeof(qgtl_inject_item_t) * (size_t)cnt * 9) bytes = sizeof(qgtl_inject_item_t) * (size_t)cnt * 9;
    void *tmp = malloc(bytes + sizeof(qgtl_gram_item_t) * (size_t)cnt * 2);
    if (!sc || !tmp) {
        free(sc); free(tmp);
        return qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory");
    }
    int rc = 0;
    size_t ip = 0, ii = 0;
    int maxb = 0, maxinv = 0;
    for (int i = 0; i < cnt; i++) {
        const qgtl_plan_t *p = B->inst[i].plan;
        sc[i].psi = 0;
        sc[i].pass_desc = ip;
        size_t nf = 0;
        for (int b = 0; b < p->nblocks; b++) {
            nf += (size_t)p->blocks[b].npasses;
            if (p->blocks[b].ninv > maxinv) maxinv = p->blocks[b].ninv;
        }
        sc[i].inv_desc = ip + nf;
        sc[i].inj_desc = ii;
        sc[i].bwd = B->inst[i].meet < p->nblocks - 1 ? bwd_arena(&B->inst[i], p->nblocks - 2) : B->inst[i].final_arena;
        ip += count_pass_descs(p);
        ii += count_inj_descs(p);
        if (p->nblocks > maxb) maxb = p->nblocks;
    }
    const size_t c64 = 8;

    /* |0...0> into every primal */
    {
        qgtl_fill_item_t *it = (qgtl_fill_item_t *)tmp;
        uint64_t u = 0;
        for (int i = 0; i < cnt; i++) {
            it[i].psi = dev_addr(B->states, B->inst[i].x_off * c64);
            it[i].N = (uint32_t)B->inst[i].N;
            it[i].unit_begin = (uint32_t)u;
            u += (B->inst[i].N + INJECT_TILE - 1) / INJECT_TILE;
        }
        rc = push_items(B, it, sizeof(*it) * (size_t)cnt, BL_FILL, 0, cnt, u);
        if (rc) goto out;
    }
    /* forward phase */
    for (int b = 0; b < maxb; b++) {
        for (int dir = 1; dir >= 0; dir--) {
            /* dir 1: backward injections at block b's input (before its passes);
             * dir 0: the passes, then forward injections after block b */
            if (dir == 0) {
                int maxp = 0;
                for (int i = 0; i < cnt; i++)
                    if (b < B->inst[i].plan->nblocks && B->inst[i].plan->blocks[b].npasses > maxp)
                        maxp = B->inst[i].plan->blocks[b].npasses;
                for (int j = 0; j < maxp; j++) {
                    /* one launch per kernel variant (2: register phases storing the split-plane
                     * form -- the forward group's final pass; 1: register phases; 0: LDS body) */
                    for (int var = 2; var >= 0; var--) {
                        qgtl_pass_item_t *it = (qgtl_pass_item_t *)tmp;
                        int k = 0;
                        uint64_t u = 0;
                        for (int i = 0; i < cnt; i++) {
                            const inst_t *in = &B->inst[i];
                            const qgtl_plan_t *p = in->plan;
                            if (b >= p->nblocks || j >= p->blocks[b].npasses) continue;
                            size_t o = 0;
                            for (int b2 = 0; b2 < b; b2++) o += (size_t)p->blocks[b2].npasses;
                            const size_t desc = sc[i].pass_desc + o + (size_t)j;
                            const int rv = regs_variant(&B->pass_h[desc], B->tune.pass_variant);
                            const int split = in->cplx && b == in->meet && j == p->blocks[b].npasses - 1 &&
                                              sc[i].fwd_slots > 0;
                            if (var == 2) {
                                if (!split) continue;
                                /* the forward states (slots 1..) in split-plane form; the primal's item
                                 * (slot 0, the launch below) moves the arena */
                                int a = sc[i].psi;
                                rc = add_pass_item(B, it, &k, &u, i, desc, p->blocks[b].pass_flips[j], 1,
                                                   sc[i].fwd_slots, &a);
                                if (rc) goto out;
                                it[k - 1].variant = 2;
                                continue;
                            }
                            if (rv != var) continue;
                            const int nstates = (b <= in->meet && !split) ? 1 + sc[i].fwd_slots : 1;
                            rc = add_pass_item(B, it, &k, &u, i, desc, p->blocks[b].pass_flips[j], 0, nstates,
                                               &sc[i].psi);
                            if (rc) goto out;
                        }
                        rc = push_items(B, it, sizeof(*it) * (size_t)k, BL_PASS, var, k, u);
                        if (rc) goto out;
                    }
                }
            }
            qgtl_inject_item_t *it = (qgtl_inject_item_t *)tmp;
            int k = 0;
            uint64_t u = 0;
            for (int i = 0; i < cnt; i++) {
                const inst_t *in = &B->inst[i];
                const qgtl_plan_t *p = in->plan;
                if (b >= p->nblocks || (dir == 1) != (b > in->meet)) continue;
                const qgtl_block_t *blk = &p->blocks[b];
                const int K = blk->param_end - blk->param_begin;
                const size_t N = in->N;
                /* forward states go where the primal is; the states of a later block where the
                 * backward group will be when they join it */
                const int dst_arena = dir ? bwd_arena(in, b - 1) : sc[i].psi;
                /* states injected right after the meeting block are final: split-plane form for the
                 * complex-form Gram */
                const uint32_t final_split =
                    (in->cplx && b == (dir ? in->meet + 1 : in->meet)) ? QGTL_INJECT_ITEM_SPLIT : 0u;
                for (int j0 = 0; j0 < K; j0 += QGTL_INJECT_MAX) {
                    const int c = (K - j0 < QGTL_INJECT_MAX) ? K - j0 : QGTL_INJECT_MAX;
                    const size_t s0 = (size_t)(blk->param_begin + j0);
                    if (k >= cnt * 8) {
                        rc = qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "more than %d injection chunks in one block",
                                            8 * cnt);
                        goto out;
                    }
                    it[k].desc = dev_addr(B->inj_d, sc[i].inj_desc++ * sizeof(qgtl_inject_desc_t));
                    it[k].psi = dev_addr(B->states, arena_off(in, sc[i].psi) * c64);
                    it[k].dst = dev_addr(B->states, (arena_off(in, dst_arena) + (1 + s0) * N) * c64);
                    it[k].partials = 0;                       /* no FP64 sums: the overlaps come from the Gram */
                    it[k].stride = N;
                    it[k].tiles = (uint32_t)in->tiles;
                    it[k].K = (uint32_t)c;
                    it[k].unit_begin = (uint32_t)u;
                    it[k].pad = final_split;
                    u += (uint64_t)it[k].tiles * (uint64_t)c;
                    k++;
                }
                if (!dir) sc[i].fwd_slots += K;
                if (!dir && b == in->meet) {
                    /* the primal after the meeting block joins the Gram as slot P + 1 */
                    it[k].desc = dev_addr(B->inj_d, (in->inj_off + count_inj_descs(p) - 1) * sizeof(qgtl_inject_desc_t));
                    it[k].psi = dev_addr(B->states, arena_off(in, sc[i].psi) * c64);
                    it[k].dst = dev_addr(B->states, (arena_off(in, sc[i].psi) + (size_t)(in->P + 1) * N) * c64);
                    it[k].partials = 0;
                    it[k].stride = N;
                    it[k].tiles = (uint32_t)in->tiles;
                    it[k].K = 1;
                    it[k].unit_begin = (uint32_t)u;
                    it[k].pad = in->cplx ? QGTL_INJECT_ITEM_SPLIT : 0u;
                    u += (uint64_t)it[k].tiles;
                    k++;
                }
            }
            rc = push_items(B, it, sizeof(*it) * (size_t)k, BL_INJECT, 0, k, u);
            if (rc) goto out;
        }
    }
    for (int i = 0; i < cnt; i++)
        if (sc[i].fwd_slots != B->inst[i].nfwd) {
            rc = qgtl_set_error(QGT_ERROR_COMPUTATION_FAILED, "internal: instance %d created %d of %d forward states",
                                i, sc[i].fwd_slots, B->inst[i].nfwd);
            goto out;
        }
    /* backward phase: block b' descending carries the states of blocks > b' through its inverse */
    for (int bb = maxb - 2; bb >= 0; bb--) {
        for (int j = 0; j < maxinv; j++) {
            for (int var = 2; var >= 0; var--) {
                qgtl_pass_item_t *it = (qgtl_pass_item_t *)tmp;
                int k = 0;
                uint64_t u = 0;
                for (int i = 0; i < cnt; i++) {
                    const inst_t *in = &B->inst[i];
                    const qgtl_plan_t *p = in->plan;
                    if (bb <= in->meet || bb > p->nblocks - 2 || j >= p->blocks[bb].ninv) continue;
                    size_t o = 0;
                    for (int b2 = 0; b2 < bb; b2++) o += (size_t)p->blocks[b2].ninv;
                    const size_t desc = sc[i].inv_desc + o + (size_t)j;
                    /* the backward group's final pass stores the split-plane form (variant 2) */
                    const int split = in->cplx && bb == in->meet + 1 && j == p->blocks[bb].ninv - 1;
                    const int v = split ? 2 : regs_variant(&B->pass_h[desc], B->tune.pass_variant);
                    if (v != var) continue;
                    const int first = p->blocks[bb + 1].param_begin;
                    rc = add_pass_item(B, it, &k, &u, i, desc, p->blocks[bb].ipass_flips[j], 1 + first, in->P - first,
                                       &sc[i].bwd);
                    if (rc) goto out;
                    if (split) it[k - 1].variant = 2;
                }
                rc = push_items(B, it, sizeof(*it) * (size_t)k, BL_PASS, var, k, u);
                if (rc) goto out;
            }
        }
    }
    for (int i = 0; i < cnt; i++)
        if (sc[i].bwd != B->inst[i].final_arena) {
            rc = qgtl_set_error(QGT_ERROR_COMPUTATION_FAILED, "internal: instance %d backward states ended off their arena",
                                i);
            goto out;
        }
    static const int term_classes[5] = {6, 4, 3, QGTL_GRAM_F16X3, QGTL_GRAM_F16X4};
    for (int tc = 0; tc < 5; tc++) {
        for (int form = 0; form < 2; form++) {          /* 0: real form (complex64 states); 1: complex form */
            qgtl_gram_item_t *it = (qgtl_gram_item_t *)tmp;
            int k = 0;
            uint64_t u = 0;
            for (int i = 0; i < cnt; i++) {
                const inst_t *in = &B->inst[i];
                if (in->terms != term_classes[tc] || in->cplx != form) continue;
                it[k].D = dev_addr(B->states, (arena_off(in, in->final_arena) + in->N) * c64);
                it[k].partials = dev_addr(B->gram_partials, in->gp_off * sizeof(float));
                it[k].N = in->N;
                it[k].chunk_len = in->chunk_len;
                it[k].P = in->P + 1;                       /* the derivative states and the primal */
                it[k].nt = in->nt;
                it[k].nchunks = in->nchunks;
                it[k].unit_begin = (uint32_t)u;
                /* real form: tile pairs x chunks (a multiple of 8, see gram_shape); complex: chunks */
                u += form ? (uint64_t)in->nchunks : (uint64_t)in->T * (uint64_t)in->nchunks;
                B->gram_flops += 4.0 * (double)in->N * in->P * (in->P + 1);
                k++;
            }
            rc = push_items(B, it, sizeof(*it) * (size_t)k, form ? BL_GRAM_CPLX : BL_GRAM,
                            form ? (term_classes[tc] == QGTL_GRAM_F16X4 ? 4 : 3) : term_classes[tc], k, u);
            if (rc) goto out;
        }
    }
    for (int form = 0; form < 2; form++) {
        qgtl_gram_reduce_item_t *it = (qgtl_gram_reduce_item_t *)tmp;
        int k = 0;
        uint64_t u = 0;
        for (int i = 0; i < cnt; i++) {
            const inst_t *in = &B->inst[i];
            if (in->cplx != form) continue;
            it[k].partials = dev_addr(B->gram_partials, in->gp_off * sizeof(float));
            it[k].C = dev_addr(B->C, in->c_off * sizeof(double));
            it[k].nt = in->nt;
            it[k].nchunks = in->nchunks;
            it[k].unit_begin = (uint32_t)u;
            it[k].pad = 0;
            /* units: 256-element groups (real form), 64-entry groups (complex form) */
            u += (uint64_t)in->T * (form ? GRAM_TILE * GRAM_TILE / 64 : GRAM_TILE * GRAM_TILE / 256);
            k++;
        }
        rc = push_items(B, it, sizeof(*it) * (size_t)k, form ? BL_GRAM_REDUCE_CPLX : BL_GRAM_REDUCE, 0, k, u);
        if (rc) goto out;
    }
    for (int form = 0; form < 2; form++) {
        qgtl_finalize_item_t *it = (qgtl_finalize_item_t *)tmp;
        int k = 0;
        uint64_t u = 0;
        const uint64_t per = form ? 512 : 256;           /* entries per unit (the kernels' workgroup sizes) */
        for (int i = 0; i < cnt; i++) {
            const inst_t *in = &B->inst[i];
            if (in->cplx != form) continue;
            it[k].C = dev_addr(B->C, in->c_off * sizeof(double));
            it[k].inj = 0;                             /* a_k and the norms from the Gram's primal row */
            it[k].Q = dev_addr(B->Qdev, in->q_off * 2 * sizeof(double));
            it[k].R = in->nt * GRAM_TILE;
            it[k].P = in->P;
            it[k].unit_begin = (uint32_t)u;
            it[k].pad = 0;
            u += ((uint64_t)in->P * in->P + per - 1) / per;
            k++;
        }
        rc = push_items(B, it, sizeof(*it) * (size_t)k, form ? BL_FINAL_CPLX : BL_FINAL, 0, k, u);
        if (rc) goto out;
    }
out:
    free(sc); free(tmp);
    return rc;
}

static int grid_for(const struct qgtl_batch *B, const launch_t *L)
{
    const int c = B->cus > 0 ? B->cus : 256;
    int g;
    switch (L->kind) {
    case BL_PASS: g = B->tune.pass_grid == 0 ? 2 * c : B->tune.pass_grid; break;
    case BL_INJECT: g = B->tune.inject_grid == 0 ? 8 * c : B->tune.inject_grid; break;
    case BL_GRAM: g = B->tune.gram_grid == 0 ? 2 * c : B->tune.gram_grid; break;
    case BL_GRAM_CPLX: g = B->tune.gram_grid == 0 ? c : B->tune.gram_grid; break;   /* 128 KB of LDS: one per CU */
    default: g = 8 * c;
    }
    return g < 0 ? 0 : g;    /* 0: one workgroup per unit */
}

static int issue(struct qgtl_batch *B)
{
    /* the descriptors' phase tables, once per execution, for every pass launch after it */
    BTRY(qgtl_hip_phase_tables_grouped(B->pass_d, (int)B->npass, B->tabs_d));
    for (int l = 0; l < B->nlaunch; l++) {
        const launch_t *L = &B->launches[l];
        const void *items = B->items_d + L->item_off;
        const int g = grid_for(B, L);
        switch (L->kind) {
        case BL_FILL: BTRY(qgtl_hip_fill_grouped((const qgtl_fill_item_t *)items, L->nitems, L->units, g)); break;
        case BL_PASS:
            BTRY(qgtl_hip_pass_grouped((const qgtl_pass_item_t *)items, L->nitems, L->units, g, L->terms));
            break;
        case BL_INJECT: BTRY(qgtl_hip_inject_grouped((const qgtl_inject_item_t *)items, L->nitems, L->units, g)); break;
        case BL_GRAM:
            BTRY(qgtl_hip_gram_grouped((const qgtl_gram_item_t *)items, L->nitems, L->units, g, L->terms));
            break;
        case BL_GRAM_REDUCE:
            BTRY(qgtl_hip_gram_reduce_grouped((const qgtl_gram_reduce_item_t *)items, L->nitems, L->units, g));
            break;
        case BL_GRAM_CPLX:
            BTRY(qgtl_hip_gram_cplx_grouped((const qgtl_gram_item_t *)items, L->nitems, L->units, g, L->terms));
            break;
        case BL_GRAM_REDUCE_CPLX:
            BTRY(qgtl_hip_gram_cplx_reduce_grouped((const qgtl_gram_reduce_item_t *)items, L->nitems, L->units, g));
            break;
        case BL_FINAL_CPLX:
            BTRY(qgtl_hip_finalize_cplx_grouped((const qgtl_finalize_item_t *)items, L->nitems, L->units, g));
            break;
        default: BTRY(qgtl_hip_finalize_grouped((const qgtl_finalize_item_t *)items, L->nitems, L->units, g));
        }
    }
    return 0;
}

/* (Re)size the Gram partials for the tune, rebuild items and the graph. */
static int rebuild(struct qgtl_batch *B)
{
    qgtl_hip_graph_destroy(B->graph);
    B->graph = NULL;
    size_t gp = 0;
    for (int i = 0; i < B->count; i++) {
        inst_t *in = &B->inst[i];
        gram_shape(B, in);
        in->gp_off = gp;
        gp += (size_t)in->nchunks * (size_t)in->T * GRAM_TILE * GRAM_TILE * (in->cplx ? 2 : 1);
    }
    if (gp > B->gp_len) {
        qgtl_hip_free(B->gram_partials);
        B->device_bytes -= B->gp_len * sizeof(float);
        B->gram_partials = NULL;
        B->gp_len = 0;
        BTRY(qgtl_hip_malloc((void **)&B->gram_partials, gp * sizeof(float)));
        B->gp_len = gp;
        B->device_bytes += gp * sizeof(float);
    }
    int rc = build_schedule(B);
    if (rc) return rc;
    qgtl_hip_free(B->items_d);
    B->items_d = NULL;
    BTRY(qgtl_hip_malloc((void **)&B->items_d, B->items_bytes));
    BTRY(qgtl_hip_h2d(B->items_d, B->items_h, B->items_bytes));
    BTRY(qgtl_hip_sync());
    if (!B->tune.no_graph) {
        BTRY(qgtl_hip_capture_begin());
        rc = issue(B);
        qgtl_hip_graph_t g = NULL;
        const int rc2 = qgtl_hip_capture_end(&g);
        if (rc) { qgtl_hip_graph_destroy(g); return rc; }
        if (rc2) return qgtl_set_error(rc2, "graph capture: %s", qgtl_hip_last_error());
        B->graph = g;
    }
    return 0;
}

/* ------------------------------------------------------------ public API */
void qgtl_batch_destroy(qgtl_batch_t *B)
{
    QGTL_LOCKED();
    if (!B) return;
    qgtl_hip_graph_destroy(B->graph);
    for (int i = 0; i < B->count; i++) qgtl_plan_destroy(B->inst[i].plan);
    qgtl_hip_free(B->states);
    qgtl_hip_free(B->gram_partials);
    qgtl_hip_free(B->C);
    qgtl_hip_free(B->Qdev);
    qgtl_hip_free(B->pass_d);
    qgtl_hip_free(B->tabs_d);
    qgtl_hip_free(B->inj_d);
    qgtl_hip_free(B->items_d);
    free(B->inst);
    qgtl_hip_host_free(B->pass_h);
    qgtl_hip_host_free(B->inj_h);
    free(B->items_h);
    free(B->launches);
    qgtl_hip_host_free(B->qhost);
    free(B);
}

static qgtl_batch_t *fail_create(struct qgtl_batch *B) { qgtl_batch_destroy(B); return NULL; }

#define CTRY(call)                                                                              \
    do {                                                                                        \
        int _rc = (call);                                                                       \
        if (_rc) {                                                                              \
            qgtl_set_error(_rc, "%s: %s", #call, qgtl_hip_last_error());                        \
            return fail_create(B);                                                              \
        }                                                                                       \
    } while (0)

qgtl_batch_t *qgtl_batch_create(int count, const int *n, const int *nops, const qgtl_op_t *ops,
                                const qgtl_batch_tune_t *tune)
{
    QGTL_LOCKED();
    if (count < 1 || !n || !nops || !ops) {
        qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "empty batch");
        return NULL;
    }
    struct qgtl_batch *B = (struct qgtl_batch *)calloc(1, sizeof *B);
    if (!B) { qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); return NULL; }
    B->count = count;
    if (tune) B->tune = *tune;
    B->inst = (inst_t *)calloc((size_t)count, sizeof(inst_t));
    if (!B->inst) { qgtl_set_error(QGT_ERROR_NO_MEMORY, "out of host memory"); return fail_create(B); }
    size_t op_off = 0, xo = 0, co = 0, qo = 0;
    for (int i = 0; i < count; i++) {
        inst_t *in = &B->inst[i];
        qgtl_tune_t pt;
        memset(&pt, 0, sizeof pt);
        pt.sweep_form = B->tune.sweep_form == 1 ? 1 : 0;
        pt.sweep_schedule = B->tune.sweep_schedule == 1 ? 1 : 0;   /* 0: with inverse blocks (two-sided) */
        in->plan = qgtl_plan_create(n[i], nops[i], ops + op_off, QGTL_MODE_MIXED, &pt);
        if (!in->plan) {
            char msg[512];
            snprintf(msg, sizeof msg, "%s", qgtl_last_error());
            qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "instance %d: %s", i, msg);
            return fail_create(B);
        }
        const qgtl_plan_t *p = in->plan;
        in->n = n[i];
        in->P = p->P;
        in->nops = nops[i];
        in->N = (size_t)1 << n[i];
        in->op_off = op_off;
        op_off += (size_t)nops[i];
        if (in->P < 1) {
            qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "instance %d has no parameters", i);
            return fail_create(B);
        }
        in->flips = 0;
        for (int b = 0; b < p->nblocks; b++) {
            if (p->blocks[b].mstage_kernels) {
                qgtl_set_error(QGT_ERROR_NOT_IMPLEMENTED,
                               "instance %d: a CNOT/SWAP stage that is not a ring on more than 13 qubits "
                               "(batched sweeps fold M-stages into passes)", i);
                return fail_create(B);
            }
            for (int j = 0; j < p->blocks[b].npasses; j++) in->flips |= p->blocks[b].pass_flips[j];
            for (int j = 0; j < p->blocks[b].ninv; j++) in->flips |= p->blocks[b].ipass_flips[j];
        }
        /* the meeting block of the two-sided schedule; the forward group ends where the
         * primal is after that block, and so must every backward-carried state */
        in->meet = qgtl_plan_meet_block(in->plan);
        in->nfwd = p->blocks[in->meet].param_end;
        int fa = 0;
        for (int b = 0; b <= in->meet; b++) fa ^= block_flips(&p->blocks[b], 0) & 1;
        in->final_arena = fa;
        const size_t states = (size_t)(in->P + 2) * in->N;    /* primal, P derivative states, primal at the meet */
        in->x_off = xo;
        xo += states;
        in->y_off = SIZE_MAX;
        if (in->flips) { in->y_off = xo; xo += states; }
        in->tiles = (int)((in->N + INJECT_TILE - 1) / INJECT_TILE);
        in->nt = (2 * (in->P + 1) + GRAM_TILE - 1) / GRAM_TILE;
        in->T = (int)upper_tiles(in->nt);
        in->c_off = co;
        co += (size_t)(in->nt * GRAM_TILE) * (size_t)(in->nt * GRAM_TILE);
        in->q_off = qo;
        qo += (size_t)in->P * (size_t)in->P;
        in->terms = qgtl_default_gram_terms(n[i]);
        in->pass_off = B->npass;
        in->inj_off = B->ninj;
        B->npass += count_pass_descs(p);
        B->ninj += count_inj_descs(p);
    }
    B->q_len = qo;
    if (qgtl_init_if_needed()) return fail_create(B);
    B->cus = qgtl_hip_cu_count();
    /* pinned host mirrors: the per-execution uploads of the angle-dependent descriptors are direct DMA */
    CTRY(qgtl_hip_host_alloc((void **)&B->pass_h, (B->npass ? B->npass : 1) * sizeof(qgtl_pass_desc_t)));
    CTRY(qgtl_hip_host_alloc((void **)&B->inj_h, (B->ninj ? B->ninj : 1) * sizeof(qgtl_inject_desc_t)));
    CTRY(qgtl_hip_host_alloc((void **)&B->qhost, qo * 2 * sizeof(double)));
    CTRY(qgtl_hip_malloc(&B->states, xo * 8));
    CTRY(qgtl_hip_malloc((void **)&B->C, co * sizeof(double)));
    CTRY(qgtl_hip_malloc((void **)&B->Qdev, qo * 2 * sizeof(double)));
    CTRY(qgtl_hip_malloc((void **)&B->pass_d, (B->npass ? B->npass : 1) * sizeof(qgtl_pass_desc_t)));
    CTRY(qgtl_hip_malloc(&B->tabs_d, (B->npass ? B->npass : 1) * 2 * QGTL_PHASE_TAB * sizeof(qgtl_c64_t)));
    CTRY(qgtl_hip_malloc((void **)&B->inj_d, (B->ninj ? B->ninj : 1) * sizeof(qgtl_inject_desc_t)));
    B->device_bytes = xo * 8 + (co + qo * 2) * sizeof(double) +
                      (B->npass ? B->npass : 1) * (sizeof(qgtl_pass_desc_t) + 2 * QGTL_PHASE_TAB * sizeof(qgtl_c64_t)) +
                      (B->ninj ? B->ninj : 1) * sizeof(qgtl_inject_desc_t);
    fill_descs(B);
    if (rebuild(B)) return fail_create(B);
    return B;
}

int qgtl_batch_count(const qgtl_batch_t *B) { return B ? B->count : -1; }

int qgtl_batch_num_params(const qgtl_batch_t *B, int i)
{
    if (!B || i < 0 || i >= B->count) return -1;
    return B->inst[i].P;
}

size_t qgtl_batch_q_offset(const qgtl_batch_t *B, int i)
{
    if (!B || i < 0 || i >= B->count) return 0;
    return B->inst[i].q_off;
}

size_t qgtl_batch_q_size(const qgtl_batch_t *B) { return B ? B->q_len : 0; }

size_t qgtl_batch_device_bytes(const qgtl_batch_t *B) { return B ? B->device_bytes : 0; }

int qgtl_batch_set_angles(qgtl_batch_t *B, const double *angles)
{
    QGTL_LOCKED();
    if (!B || !angles) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null batch or angles");
    /* the instances' plans are independent: rebuild them on the host cores (the
     * workers must not take the library lock this thread holds) */
    int first_bad = B->count;
#pragma omp parallel for schedule(dynamic, 4) reduction(min : first_bad)
    for (int i = 0; i < B->count; i++) {
        const inst_t *in = &B->inst[i];
        if (qgtl_plan_set_angles_nolock(in->plan, in->nops, angles + in->op_off) != 0) {
            if (i < first_bad) first_bad = i;
        } else {
            fill_inst_descs(B, i);
        }
    }
    B->descs_dirty = 1;
    if (first_bad < B->count) {
        /* again on this thread, for the error message */
        const inst_t *in = &B->inst[first_bad];
        return qgtl_plan_set_angles_nolock(in->plan, in->nops, angles + in->op_off);
    }
    return 0;
}

/* plans follow the tune's sweep form (descriptors rebuilt on the host) */
static int apply_sweep_form(struct qgtl_batch *B)
{
    const int form = B->tune.sweep_form == 1 ? 1 : 0;
    int changed = 0;
    for (int i = 0; i < B->count; i++) {
        qgtl_plan_t *p = B->inst[i].plan;
        if (p->tune.sweep_form == form) continue;
        p->tune.sweep_form = form;
        const int rc = qgtl_plan_build(p);
        if (rc) return rc;
        changed = 1;
    }
    if (changed) fill_descs(B);
    return 0;
}

int qgtl_batch_set_tune(qgtl_batch_t *B, const qgtl_batch_tune_t *tune)
{
    QGTL_LOCKED();
    if (!B || !tune) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null batch or tune");
    B->tune = *tune;
    const int rc = apply_sweep_form(B);
    if (rc) return rc;
    return rebuild(B);
}

int qgtl_batch_get_tune(const qgtl_batch_t *B, qgtl_batch_tune_t *tune)
{
    if (!B || !tune) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null batch or tune");
    *tune = B->tune;
    return 0;
}

/* Run the sequence; device time (HIP events) to *device_ms. */
static int run_device(struct qgtl_batch *B, double *device_ms)
{
    if (B->descs_dirty) {
        BTRY(qgtl_hip_h2d(B->pass_d, B->pass_h, B->npass * sizeof(qgtl_pass_desc_t)));
        BTRY(qgtl_hip_h2d(B->inj_d, B->inj_h, B->ninj * sizeof(qgtl_inject_desc_t)));
        BTRY(qgtl_hip_sync());          /* the host mirrors may change right after return */
        B->descs_dirty = 0;
    }
    qgtl_hip_event_t e0 = NULL, e1 = NULL;
    BTRY(qgtl_hip_event_create(&e0));
    int rc = qgtl_hip_event_create(&e1);
    if (!rc) rc = qgtl_hip_event_record(e0);
    if (!rc) rc = B->graph ? qgtl_hip_graph_launch(B->graph) : issue(B);
    if (!rc) rc = qgtl_hip_event_record(e1);
    if (!rc) rc = qgtl_hip_sync();
    float ms = 0.f;
    if (!rc) rc = qgtl_hip_event_elapsed_ms(e0, e1, &ms);
    qgtl_hip_event_destroy(e0);
    qgtl_hip_event_destroy(e1);
    if (rc) return qgtl_set_error(rc, "batch execution: %s", qgtl_hip_last_error());
    *device_ms = ms;
    return 0;
}

int qgtl_batch_execute(qgtl_batch_t *B, double *Q, qgtl_batch_stats_t *st)
{
    QGTL_LOCKED();
    if (!B) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null batch");
    const double t0 = now_ms();
    const int dirty = B->descs_dirty;
    double dev = 0.0;
    int rc = run_device(B, &dev);
    if (rc) return rc;
    const double t1 = now_ms();
    if (Q) {
        BTRY(qgtl_hip_d2h(B->qhost, B->Qdev, B->q_len * 2 * sizeof(double)));
        BTRY(qgtl_hip_sync());
        /* slot order (creation order) -> parameter numbers, per instance on the host cores
         * (a plain copy where the two orders agree) */
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = 0; i < B->count; i++) {
            const inst_t *in = &B->inst[i];
            const int *sp = in->plan->slot_param;
            const size_t P = (size_t)in->P;
            const double *src = B->qhost + 2 * in->q_off;
            double *dst = Q + 2 * in->q_off;
            int identity = 1;
            for (size_t a = 0; a < P && identity; a++) identity = sp[a] == (int)a;
            if (identity) {
                memcpy(dst, src, P * P * 2 * sizeof(double));
                continue;
            }
            for (size_t a = 0; a < P; a++)
                for (size_t c = 0; c < P; c++) {
                    const size_t o = 2 * ((size_t)sp[a] * P + (size_t)sp[c]);
                    dst[o] = src[2 * (a * P + c)];
                    dst[o + 1] = src[2 * (a * P + c) + 1];
                }
        }
    }
    const double t2 = now_ms();
    if (st) {
        memset(st, 0, sizeof *st);
        st->wall_ms = t2 - t0;
        st->device_ms = dev;
        st->upload_ms = dirty ? (t1 - t0) - dev : 0.0;
        if (st->upload_ms < 0) st->upload_ms = 0;
        st->download_ms = t2 - t1;
        st->launches = B->nlaunch;
        st->units = B->units;
        st->sweep_bytes = B->sweep_bytes;
        st->gram_flops = B->gram_flops;
        st->instances = B->count;
        st->graph = B->graph != NULL;
    }
    return 0;
}

/* ------------------------------------------------------------ autotuner */
static int time_tune(struct qgtl_batch *B, const qgtl_batch_tune_t *t, int reps, double *ms)
{
    B->tune = *t;
    int rc = apply_sweep_form(B);
    if (rc) return rc;
    rc = rebuild(B);
    if (rc) return rc;
    double v[16];
    if (reps > 16) reps = 16;
    if (reps < 1) reps = 1;
    double warm = 0.0;
    rc = run_device(B, &warm);
    if (rc) return rc;
    for (int r = 0; r < reps; r++) {
        rc = run_device(B, &v[r]);
        if (rc) return rc;
    }
    for (int a = 1; a < reps; a++)        /* median */
        for (int c = a; c > 0 && v[c] < v[c - 1]; c--) { const double x = v[c]; v[c] = v[c - 1]; v[c - 1] = x; }
    *ms = v[reps / 2];
    return 0;
}

int qgtl_batch_autotune(qgtl_batch_t *B, int reps, const char *json_path, qgtl_batch_tune_t *best_out)
{
    QGTL_LOCKED();
    if (!B) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null batch");
    const int c = B->cus > 0 ? B->cus : 256;
    qgtl_batch_tune_t best = B->tune;
    best.no_graph = 0;
    double best_ms = 0.0;
    int rc = time_tune(B, &best, reps, &best_ms);
    if (rc) return rc;
    const double first_ms = best_ms;
    /* coordinate descent over the knobs, one sweep */
    const int pass_grid[] = {2 * c, 4 * c, 8 * c, -1};
    const int variant[] = {0, 1};
    const int inj_grid[] = {4 * c, 8 * c, 16 * c, -1};
    const int chunk[] = {4096, 8192, 16384, 32768, 65536};
    const int gram_grid[] = {c, 2 * c, 4 * c, -1};
    const int form[] = {0, 1};
    struct { int *field; const int *vals; int nv; } knobs[] = {
        {&best.pass_grid, pass_grid, 4},
        {&best.pass_variant, variant, 2},
        {&best.inject_grid, inj_grid, 4},
        {&best.gram_chunk_len, chunk, 5},
        {&best.gram_grid, gram_grid, 4},
        {&best.sweep_form, form, 2},
    };
    for (size_t k = 0; k < sizeof knobs / sizeof knobs[0]; k++) {
        const int keep = *knobs[k].field;
        int chosen = keep;
        for (int v = 0; v < knobs[k].nv; v++) {
            if (knobs[k].vals[v] == keep) continue;
            qgtl_batch_tune_t cand = best;
            *(int *)((char *)&cand + ((char *)knobs[k].field - (char *)&best)) = knobs[k].vals[v];
            double ms = 0.0;
            rc = time_tune(B, &cand, reps, &ms);
            if (rc) return rc;
            if (ms < best_ms) { best_ms = ms; chosen = knobs[k].vals[v]; }
        }
        *knobs[k].field = chosen;
    }
    B->tune = best;
    rc = apply_sweep_form(B);
    if (!rc) rc = rebuild(B);
    if (rc) return rc;
    if (best_out) *best_out = best;
    if (json_path) {
        FILE *f = fopen(json_path, "w");
        if (!f) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "cannot write %s", json_path);
        fprintf(f,
                "{\"pass_grid\": %d, \"pass_variant\": %d, \"inject_grid\": %d, \"gram_chunk_len\": %d, "
                "\"gram_grid\": %d, \"sweep_form\": %d, \"cus\": %d, \"instances\": %d, \"device_ms\": %.4f, "
                "\"untuned_ms\": %.4f}\n",
                best.pass_grid, best.pass_variant, best.inject_grid, best.gram_chunk_len, best.gram_grid,
                best.sweep_form, c, B->count, best_ms, first_ms);
        fclose(f);
    }
    return 0;
}

static int json_int(const char *text, const char *key, int32_t *out)
{
    char pat[64];
    snprintf(pat, sizeof pat, "\"%s\"", key);
    const char *p = strstr(text, pat);
    if (!p) return 0;
    p = strchr(p + strlen(pat), ':');
    if (!p) return 0;
    int v = 0;
    if (sscanf(p + 1, "%d", &v) != 1) return 0;
    *out = v;
    return 1;
}

int qgtl_batch_tune_load(const char *json_path, qgtl_batch_tune_t *tune)
{
    if (!json_path || !tune) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null path or tune");
    FILE *f = fopen(json_path, "r");
    if (!f) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "cannot read %s", json_path);
    char text[1024];
    const size_t len = fread(text, 1, sizeof text - 1, f);
    fclose(f);
    text[len] = '\0';
    memset(tune, 0, sizeof *tune);
    json_int(text, "pass_grid", &tune->pass_grid);
    json_int(text, "pass_variant", &tune->pass_variant);
    json_int(text, "inject_grid", &tune->inject_grid);
    json_int(text, "gram_chunk_len", &tune->gram_chunk_len);
    json_int(text, "gram_grid", &tune->gram_grid);
    json_int(text, "sweep_form", &tune->sweep_form);
    return 0;
}
