/*
 * CPU check of the register-pass descriptors (no GPU work): every pass the plan
 * builder emits for the register kernel is applied here in FP64 by the
 * contract qgtl_hip_abi.h states for it -- tile gather, pre layer, the gates
 * as K(t) = [[1, -t], [t, 1]], post layer, the store at tile index j ^ xmask,
 * the ring permutation on store or (inverted) on load, the layers moved
 * between the two passes of a block (nolayer / tf) -- and the result of each
 * block's pass sequence is compared with the block applied densely: each
 * qubit's 2 x 2 matrix, then the CNOT ring (forward); the inverse ring, then
 * each qubit's adjoint (the inverse passes).  Circuits: the hardware-efficient
 * ansatz (RY, RZ layers, post phases only; random angles give rotations with
 * |c| >= |s| and |c| < |s|) and an RX, RY, RZ ansatz (pre and post phases, so
 * no layer moves), at 14, 15 and 17 qubits.  Prints "ok" and exits 0 on
 * success.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_internal.h"

typedef double complex cplx;

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static double urand(void)
{
    rng_state += 0x9e3779b97f4a7c15ull;
    uint64_t z = rng_state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static int bit(uint64_t x, int b) { return (int)((x >> b) & 1u); }
static int parity(uint64_t x) { return __builtin_popcountll(x) & 1; }

/* CNOT(c -> t) on the whole state */
static void cnot(cplx *psi, int n, int c, int t)
{
    const uint64_t N = 1ull << n;
    for (uint64_t i = 0; i < N; i++)
        if (bit(i, c) && !bit(i, t)) {
            const uint64_t k = i | (1ull << t);
            const cplx a = psi[i];
            psi[i] = psi[k];
            psi[k] = a;
        }
}

/* the M-stage of the ansatz: CNOT(i -> i+1 mod n), i ascending; inverse: descending */
static void ring(cplx *psi, int n, int inverse)
{
    for (int s = 0; s < n; s++) {
        const int i = inverse ? n - 1 - s : s;
        cnot(psi, n, i, (i + 1) % n);
    }
}

static void apply_1q(cplx *psi, int n, int q, const qgtl_mat2 *u, int adjoint)
{
    cplx m[4];
    for (int e = 0; e < 4; e++) m[e] = u->m[e].re + I * u->m[e].im;
    if (adjoint) {
        const cplx a = conj(m[0]), b = conj(m[2]), c = conj(m[1]), d = conj(m[3]);
        m[0] = a; m[1] = b; m[2] = c; m[3] = d;
    }
    const uint64_t N = 1ull << n;
    for (uint64_t i = 0; i < N; i++)
        if (!bit(i, q)) {
            const uint64_t k = i | (1ull << q);
            const cplx a = psi[i], b = psi[k];
            psi[i] = m[0] * a + m[1] * b;
            psi[k] = m[2] * a + m[3] * b;
        }
}

static cplx c64c(qgtl_c64_t z) { return (double)z.re + I * (double)z.im; }

/* state index of tile element j of tile r */
static uint64_t place(const qgtl_pass_desc_t *d, uint32_t j, uint64_t r)
{
    uint64_t x = j & ((1u << d->nlow) - 1u);
    for (int i = 0; i < d->m - d->nlow; i++) x |= (uint64_t)bit(j >> d->nlow, i) << d->hi_pos[i];
    for (int i = 0; i < d->nrest; i++) x |= (uint64_t)bit(r, i) << d->rest_pos[i];
    return x;
}

/* one register pass by its descriptor contract; psi -> out (both 2^n) */
static int apply_pass(const qgtl_pass_desc_t *d, int n, cplx *psi, cplx *out)
{
    if (d->form != 1 || d->m != QGTL_TILE_BITS_MAX || d->nphase < 1 || d->store == QGTL_STORE_PERM) {
        fprintf(stderr, "not a rotation-form register pass\n");
        return 1;
    }
    const uint64_t N = 1ull << n;
    const uint32_t M = 1u << d->m;
    if (d->load == QGTL_LOAD_RING) ring(psi, n, 1);
    cplx *w = malloc(sizeof(cplx) * M);
    if (!w) return 1;
    const int layers = d->nolayer == 0;
    const int post_on = layers && (d->has_post || !d->has_pre), pre_on = layers && d->has_pre;
    for (uint64_t r = 0; r < (1ull << d->nrest); r++) {
        for (uint32_t j = 0; j < M; j++) w[j] = psi[place(d, j, r)];
        cplx tf = 1.0;
        if (layers)
            for (int i = 0; i < d->tf_n; i++) tf *= c64c(d->tf[i][bit(r, d->tf_bit[i])]);
        if (pre_on)
            for (uint32_t j = 0; j < M; j++) {
                cplx f = post_on ? 1.0 : (double)d->scale * tf;
                for (int g = 0; g < d->ngates; g++) f *= c64c(d->pre[g][bit(j, d->gate_bit[g])]);
                if (d->xfold == 1 && parity(j & d->xmask)) f = -f;
                w[j] *= f;
            }
        for (int ph = 0; ph < d->nphase; ph++)
            for (int g = d->ph_first[ph]; g < d->ph_first[ph + 1]; g++) {
                const int reg = d->ph_greg[g], b = d->ph_rbits[ph][reg];
                if (b != d->gate_bit[g]) {
                    fprintf(stderr, "gate %d: register bit maps to tile bit %d, gate bit %d\n", g, b, d->gate_bit[g]);
                    free(w);
                    return 1;
                }
                const double t = d->ph_tan[ph][reg];
                for (uint32_t j = 0; j < M; j++)
                    if (!bit(j, b)) {
                        const uint32_t k = j | (1u << b);
                        const cplx a = w[j], c = w[k];
                        w[j] = a - t * c;
                        w[k] = c + t * a;
                    }
            }
        if (post_on)
            for (uint32_t j = 0; j < M; j++) {
                cplx f = (double)d->scale * tf;
                for (int g = 0; g < d->ngates; g++) f *= c64c(d->post[g][bit(j ^ d->xmask, d->gate_bit[g])]);
                if (d->xfold == 0 && parity(j & d->xmask)) f = -f;
                w[j] *= f;
            }
        for (uint32_t j = 0; j < M; j++) out[place(d, j ^ d->xmask, r)] = w[j];
    }
    free(w);
    if (d->store == QGTL_STORE_RING) ring(out, n, 0);
    (void)N;
    return 0;
}

static double max_diff(const cplx *a, const cplx *b, uint64_t N)
{
    double e = 0.0;
    for (uint64_t i = 0; i < N; i++) {
        const double d = cabs(a[i] - b[i]);
        if (d > e) e = d;
    }
    return e;
}

/* 0: ok */
static int check_circuit(int n, int layers, int with_rx, int *folds, int *swaps)
{
    const int per = (with_rx ? 3 : 2) * n + n;
    const int nops = layers * per;
    qgtl_op_t *ops = calloc((size_t)nops, sizeof *ops);
    int k = 0, p = 0;
    for (int l = 0; l < layers; l++) {
        const int types[3] = {GATE_TYPE_RX, GATE_TYPE_RY, GATE_TYPE_RZ};
        for (int t = with_rx ? 0 : 1; t < 3; t++)
            for (int q = 0; q < n; q++) {
                qgtl_op_t op = {types[t], q, -1, p++, 2.0 * M_PI * urand()};
                ops[k++] = op;
            }
        for (int q = 0; q < n; q++) {
            qgtl_op_t op = {GATE_TYPE_CNOT, q, (q + 1) % n, -1, 0.0};
            ops[k++] = op;
        }
    }
    struct qgtl_plan *plan = calloc(1, sizeof *plan);
    plan->n = n;
    plan->nops = k;
    plan->ops = ops;
    plan->mode = QGTL_MODE_MIXED;
    if (qgtl_plan_build(plan) != 0) {
        fprintf(stderr, "plan build failed: %s\n", qgtl_last_error());
        return 1;
    }
    const uint64_t N = 1ull << n;
    cplx *psi = malloc(sizeof(cplx) * N), *ref = malloc(sizeof(cplx) * N), *a = malloc(sizeof(cplx) * N),
         *b = malloc(sizeof(cplx) * N);
    int fail = 0;
    for (int bi = 0; bi < plan->nblocks && !fail; bi++) {
        const qgtl_block_t *blk = &plan->blocks[bi];
        for (int inv = 0; inv < 2 && !fail; inv++) {
            const int np = inv ? blk->ninv : blk->npasses;
            const qgtl_pass_desc_t *passes = inv ? blk->ipass : blk->pass;
            if (np < 1 || !blk->is_ring) continue;
            double nrm = 0.0;
            for (uint64_t i = 0; i < N; i++) {
                psi[i] = (urand() - 0.5) + I * (urand() - 0.5);
                nrm += creal(psi[i] * conj(psi[i]));
            }
            for (uint64_t i = 0; i < N; i++) psi[i] /= sqrt(nrm);
            memcpy(ref, psi, sizeof(cplx) * N);
            if (!inv) {
                for (int q = 0; q < n; q++)
                    if (blk->has_u[q]) apply_1q(ref, n, q, &blk->u[q], 0);
                ring(ref, n, 0);
            } else {
                ring(ref, n, 1);
                for (int q = 0; q < n; q++)
                    if (blk->has_u[q]) apply_1q(ref, n, q, &blk->u[q], 1);
            }
            memcpy(a, psi, sizeof(cplx) * N);
            for (int i = 0; i < np && !fail; i++) {
                *folds += passes[i].nolayer;
                *swaps += __builtin_popcount(passes[i].xmask);
                fail = apply_pass(&passes[i], n, a, b);
                cplx *t = a; a = b; b = t;
            }
            const double e = max_diff(a, ref, N);
            printf("n=%d rx=%d block %d %s: %d passes, max |diff| %.3e\n", n, with_rx, bi, inv ? "inverse" : "forward", np, e);
            if (!(e < 2e-5)) fail = 1;
        }
    }
    free(psi); free(ref); free(a); free(b); free(ops);
    return fail;
}

int main(void)
{
    int folds = 0, swaps = 0;
    const int cases[][3] = {{14, 2, 0}, {15, 2, 1}, {17, 2, 0}, {14, 3, 1}};
    for (size_t c = 0; c < sizeof cases / sizeof cases[0]; c++)
        if (check_circuit(cases[c][0], cases[c][1], cases[c][2], &folds, &swaps)) {
            printf("FAIL\n");
            return 1;
        }
    printf("moved layers %d, swapped gates %d\n", folds, swaps);
    if (folds == 0 || swaps == 0) {
        printf("FAIL: the cases did not exercise moved layers and swapped gates\n");
        return 1;
    }
    printf("ok\n");
    return 0;
}
