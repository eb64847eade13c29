/*
 * qgtl_plan.c -- compile a flat gate list into blocks and device passes.
 *
 * Gate conventions follow QGTL: RX = [[c, -is], [-is, c]], RY = [[c, -s], [s, c]],
 * RZ = diag(e^{-i t/2}, e^{i t/2}) with c = cos(t/2), s = sin(t/2)
 * (src/quantum_geometric/core/quantum_gate_operations.c:11-47); qubit q is
 * bit 1 << q (src/quantum_geometric/hybrid/quantum_classical_algorithms.c:174-192);
 * CNOT flips the target where the control bit is set (:195-210).
 *
 * Block rule: a single-qubit gate on qubit q joins the current block's
 * U-layer unless the block's M-stage has already touched q; a CNOT or SWAP
 * joins the M-stage.  CZ and CY are rewritten as H.CNOT.H and S.CNOT.Sdg.
 *
 * Derivative of parameter k carried by gate g on qubit q: with the block's
 * U_q = S g R (S = gates after g in the block, R = before), the derivative
 * state before the M-stage is S (-i/2 G) g R = A_k U_q, A_k = S (-i/2 G) S^H,
 * applied to the block's output.  Through the M-stage permutation pi,
 * d_k[y] = A[s][s] psi[y] + A[s][1-s] psi[y ^ pi(e_q)], s = (pi^-1 y)_q.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_internal.h"

#define CMIN_LOW_BITS 5   /* contiguous low bits kept in every high-qubit pass (256-B runs) */

static qgtl_cd cd(double re, double im) { qgtl_cd z = {re, im}; return z; }
static qgtl_cd cmul(qgtl_cd a, qgtl_cd b) { return cd(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static qgtl_cd cadd(qgtl_cd a, qgtl_cd b) { return cd(a.re + b.re, a.im + b.im); }

static qgtl_mat2 mat_mul(const qgtl_mat2 *a, const qgtl_mat2 *b)
{
    qgtl_mat2 r;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            r.m[2 * i + j] = cadd(cmul(a->m[2 * i], b->m[j]), cmul(a->m[2 * i + 1], b->m[2 + j]));
    return r;
}

static qgtl_mat2 mat_adj(const qgtl_mat2 *a)
{
    qgtl_mat2 r;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) r.m[2 * i + j] = cd(a->m[2 * j + i].re, -a->m[2 * j + i].im);
    return r;
}

static qgtl_mat2 mat_identity(void)
{
    qgtl_mat2 r;
    memset(&r, 0, sizeof r);
    r.m[0].re = 1.0;
    r.m[3].re = 1.0;
    return r;
}

static void set4(qgtl_mat2 *m, qgtl_cd a, qgtl_cd b, qgtl_cd c, qgtl_cd d)
{
    m->m[0] = a; m->m[1] = b; m->m[2] = c; m->m[3] = d;
}

int qgtl_is_param_gate(int type)
{
    return type == GATE_TYPE_RX || type == GATE_TYPE_RY || type == GATE_TYPE_RZ;
}

int qgtl_gate_matrix(const qgtl_op_t *op, qgtl_mat2 *m)
{
    const double r = M_SQRT1_2;
    const double c = cos(op->angle / 2.0), s = sin(op->angle / 2.0);
    const qgtl_cd z = cd(0, 0), one = cd(1, 0);
    switch (op->type) {
    case GATE_TYPE_I: set4(m, one, z, z, one); return 0;
    case GATE_TYPE_X: set4(m, z, one, one, z); return 0;
    case GATE_TYPE_Y: set4(m, z, cd(0, -1), cd(0, 1), z); return 0;
    case GATE_TYPE_Z: set4(m, one, z, z, cd(-1, 0)); return 0;
    case GATE_TYPE_H: set4(m, cd(r, 0), cd(r, 0), cd(r, 0), cd(-r, 0)); return 0;
    case GATE_TYPE_S: set4(m, one, z, z, cd(0, 1)); return 0;
    case GATE_TYPE_SDG: set4(m, one, z, z, cd(0, -1)); return 0;
    case GATE_TYPE_T: set4(m, one, z, z, cd(r, r)); return 0;
    case GATE_TYPE_TDG: set4(m, one, z, z, cd(r, -r)); return 0;
    case GATE_TYPE_SX: set4(m, cd(0.5, 0.5), cd(0.5, -0.5), cd(0.5, -0.5), cd(0.5, 0.5)); return 0;
    case GATE_TYPE_RX: set4(m, cd(c, 0), cd(0, -s), cd(0, -s), cd(c, 0)); return 0;
    case GATE_TYPE_RY: set4(m, cd(c, 0), cd(-s, 0), cd(s, 0), cd(c, 0)); return 0;
    case GATE_TYPE_RZ: set4(m, cd(c, -s), z, z, cd(c, s)); return 0;
    case GATE_TYPE_U1:
    case GATE_TYPE_PHASE: set4(m, one, z, z, cd(cos(op->angle), sin(op->angle))); return 0;
    default: return -1;
    }
}

/* (-i/2) G for the generator G of RX/RY/RZ. */
static qgtl_mat2 generator(int type)
{
    qgtl_mat2 g;
    const qgtl_cd z = cd(0, 0);
    if (type == GATE_TYPE_RX) set4(&g, z, cd(0, -0.5), cd(0, -0.5), z);
    else if (type == GATE_TYPE_RY) set4(&g, z, cd(-0.5, 0), cd(0.5, 0), z);
    else set4(&g, cd(0, -0.5), z, z, cd(0, 0.5));
    return g;
}

static int is_single_qubit(int type)
{
    qgtl_op_t probe = {type, 0, -1, -1, 0.0};
    qgtl_mat2 m;
    return qgtl_gate_matrix(&probe, &m) == 0;
}

uint64_t qgtl_structure_hash(int n, int nops, const qgtl_op_t *ops, int mode)
{
    uint64_t h = 1469598103934665603ull;
#define MIX(x) do { h ^= (uint64_t)(uint32_t)(x); h *= 1099511628211ull; } while (0)
    MIX(n); MIX(mode); MIX(nops);
    for (int i = 0; i < nops; i++) { MIX(ops[i].type); MIX(ops[i].q0); MIX(ops[i].q1); MIX(ops[i].param); }
#undef MIX
    return h;
}

static uint32_t perm_apply(const qgtl_block_t *b, uint32_t x)
{
    for (int i = 0; i < b->nperm; i++) {
        const uint32_t ab = (x >> b->perm_a[i]) & 1u, bb = (x >> b->perm_b[i]) & 1u;
        if (b->perm_type[i] == 0) x ^= ab << b->perm_b[i];
        else x ^= (ab ^ bb) * ((1u << b->perm_a[i]) | (1u << b->perm_b[i]));
    }
    return x;
}

static uint32_t perm_apply_inverse(const qgtl_block_t *b, uint32_t y)
{
    for (int i = b->nperm - 1; i >= 0; i--) {
        const uint32_t ab = (y >> b->perm_a[i]) & 1u, bb = (y >> b->perm_b[i]) & 1u;
        if (b->perm_type[i] == 0) y ^= ab << b->perm_b[i];
        else y ^= (ab ^ bb) * ((1u << b->perm_a[i]) | (1u << b->perm_b[i]));
    }
    return y;
}

static void to_c64(const qgtl_mat2 *m, qgtl_c64_t out[4])
{
    for (int e = 0; e < 4; e++) {
        out[e].re = (float)m->m[e].re;
        out[e].im = (float)m->m[e].im;
    }
}

static double cabs2(qgtl_cd z) { return sqrt(z.re * z.re + z.im * z.im); }
static qgtl_cd cdiv_real(qgtl_cd z, double r) { return cd(z.re / r, z.im / r); }
static qgtl_cd cconj(qgtl_cd z) { return cd(z.re, -z.im); }
static qgtl_c64_t c64(qgtl_cd z) { qgtl_c64_t o = {(float)z.re, (float)z.im}; return o; }

/* U = diag(post) [[c, -s], [s, c]] diag(pre), FP64, for a unitary 2 x 2 U.
 * When U00 conj(U01) and U10 conj(U11) are real (U = RZ RY, the ansatz's
 * product) pre = (1, 1) and c, s carry the signs; otherwise pre0 = 1. */
static void factor_rotation(const qgtl_mat2 *U, double *c, double *s, qgtl_cd pre[2], qgtl_cd post[2])
{
    const qgtl_cd a = U->m[0], b = U->m[1], cc = U->m[2], d = U->m[3];
    const double na = cabs2(a), nb = cabs2(b), nc = cabs2(cc), nd = cabs2(d);
    const qgtl_cd r0 = cmul(a, cconj(b)), r1 = cmul(cc, cconj(d));
    const double tol = 1e-13;
    pre[0] = pre[1] = cd(1.0, 0.0);
    if (fabs(r0.im) <= tol && fabs(r1.im) <= tol) {
        /* rows are a phase times a real row */
        post[0] = na >= nb ? cdiv_real(a, na) : cdiv_real(cd(-b.re, -b.im), nb);
        post[1] = nd >= nc ? cdiv_real(d, nd) : cdiv_real(cc, nc);
        const qgtl_cd p0c = cconj(post[0]), p1c = cconj(post[1]);
        /* c = Re(U00 / post0), s = -Re(U01 / post0); check row 1 agrees */
        *c = cmul(a, p0c).re;
        *s = -cmul(b, p0c).re;
        const double c1 = cmul(d, p1c).re, s1 = cmul(cc, p1c).re;
        if (fabs(c1 - *c) <= 1e-12 && fabs(s1 - *s) <= 1e-12) return;
        /* row 1 differs in sign: fold it into post1 */
        post[1] = cd(-post[1].re, -post[1].im);
        if (fabs(-c1 - *c) <= 1e-12 && fabs(-s1 - *s) <= 1e-12) return;
        post[1] = cd(-post[1].re, -post[1].im);
    }
    /* columns are a phase times a real column: U = R diag(pre), post = (1, 1)
     * (the adjoint of the row case, as in the inverse of U = RZ RY) */
    const qgtl_cd q0 = cmul(cconj(a), cc), q1 = cmul(cconj(b), d);
    if (fabs(q0.im) <= tol && fabs(q1.im) <= tol) {
        post[0] = post[1] = cd(1.0, 0.0);
        pre[0] = na >= nc ? cdiv_real(a, na) : cdiv_real(cc, nc);
        const qgtl_cd p0c = cconj(pre[0]);
        *c = cmul(a, p0c).re;
        *s = cmul(cc, p0c).re;
        /* U11 = c pre1, U01 = -s pre1 */
        pre[1] = fabs(*c) >= fabs(*s) ? cdiv_real(d, *c) : cdiv_real(cd(-b.re, -b.im), *s);
        const qgtl_cd e01 = cmul(cd(-*s, 0.0), pre[1]), e11 = cmul(cd(*c, 0.0), pre[1]);
        if (fabs(e01.re - b.re) <= 1e-12 && fabs(e01.im - b.im) <= 1e-12 && fabs(e11.re - d.re) <= 1e-12 &&
            fabs(e11.im - d.im) <= 1e-12)
            return;
        pre[0] = pre[1] = cd(1.0, 0.0);
    }
    /* general: pre0 = 1, c = |U00| >= 0, s = |U10| >= 0 */
    *c = na;
    *s = nc;
    if (na > 0.0 && nc > 0.0) {
        post[0] = cdiv_real(a, na);
        post[1] = cdiv_real(cc, nc);
        pre[1] = cdiv_real(cmul(cconj(post[1]), d), na);     /* U11 = post1 c pre1 */
    } else if (nc == 0.0) {                                  /* diagonal */
        *c = 1.0; *s = 0.0;
        post[0] = a;
        post[1] = cd(1.0, 0.0);
        pre[1] = d;
    } else {                                                 /* anti-diagonal: U01 = -post0 pre1, U10 = post1 */
        *c = 0.0; *s = 1.0;
        post[1] = cc;
        post[0] = cd(1.0, 0.0);
        pre[1] = cd(-b.re, -b.im);
    }
}

/* matrix g of pass p: general form always, rotation form alongside */
static void set_gate(qgtl_pass_desc_t *p, int g, const qgtl_mat2 *U)
{
    to_c64(U, p->mat[g]);
    double c, s;
    qgtl_cd pre[2], post[2];
    factor_rotation(U, &c, &s, pre, post);
    p->rot[g].re = (float)c;
    p->rot[g].im = (float)s;
    for (int b = 0; b < 2; b++) {
        p->pre[g][b] = c64(pre[b]);
        p->post[g][b] = c64(post[b]);
    }
}

/* switch a finished pass to the rotation form */
static void finish_form(qgtl_pass_desc_t *p, int general)
{
    p->form = 0;
    p->has_pre = 0;
    if (general) return;
    for (int g = 0; g < p->ngates; g++) {
        if (p->pre[g][0].re != 1.f || p->pre[g][0].im != 0.f || p->pre[g][1].re != 1.f || p->pre[g][1].im != 0.f)
            p->has_pre = 1;
    }
    p->form = 1;
}

/* Register-kernel form of a finished rotation-form pass (see
 * qgtl_pass_desc_t): per gate the scaled-out rotation coefficient and the
 * factor taken out, the pair swaps left to the store (xmask) and where their
 * signs go (xfold), and the register parts of the phase layers, from the
 * block's matrices u (FP64, rounded once). */
static void finish_rotation(qgtl_pass_desc_t *p, const qgtl_mat2 *u)
{
    p->has_post = 0;
    p->scale = 1.f;
    p->xmask = 0;
    p->xfold = 0;
    p->nolayer = 0;
    p->tf_n = 0;
    for (int ph = 0; ph < QGTL_MAX_PHASES; ph++)
        for (int j = 0; j < 5; j++) p->ph_tan[ph][j] = 0.f;
    for (int r = 0; r < 32; r++) p->reg_post[r] = p->reg_pre[r] = c64(cd(1.0, 0.0));
    if (p->form != 1) return;
    double cs[QGTL_TILE_BITS_MAX], sns[QGTL_TILE_BITS_MAX], coef[QGTL_TILE_BITS_MAX], scale = 1.0;
    int swap[QGTL_TILE_BITS_MAX], gate_of_bit[QGTL_TILE_BITS_MAX];
    qgtl_cd rel_pre[QGTL_TILE_BITS_MAX], rel_post[QGTL_TILE_BITS_MAX];
    for (int b = 0; b < QGTL_TILE_BITS_MAX; b++) gate_of_bit[b] = -1;
    for (int g = 0; g < p->ngates; g++) {
        const int tb = p->gate_bit[g];
        const int q = tb < p->nlow ? tb : p->hi_pos[tb - p->nlow];
        qgtl_cd pre[2], post[2];
        factor_rotation(&u[q], &cs[g], &sns[g], pre, post);
        rel_pre[g] = cmul(pre[1], cconj(pre[0]));
        rel_post[g] = cmul(post[1], cconj(post[0]));
        if (post[0].re != 1.0 || post[0].im != 0.0 || post[1].re != 1.0 || post[1].im != 0.0) p->has_post = 1;
        gate_of_bit[tb] = g;
    }
    /* the sign of a swapped gate goes into the layer the kernel applies:
     * the post layer unless the pass has a pre layer and no post layer */
    p->xfold = (p->has_pre && !p->has_post) ? 1 : 0;
    for (int g = 0; g < p->ngates; g++) {
        const double c = cs[g], sn = sns[g];
        swap[g] = fabs(c) < fabs(sn);
        if (swap[g]) {
            p->xmask |= 1u << p->gate_bit[g];
            coef[g] = p->xfold ? c / sn : -c / sn;
            scale *= sn;
        } else {
            coef[g] = sn / c;
            scale *= c;
        }
    }
    p->scale = (float)scale;
    if (p->nphase < 1) return;
    for (int ph = 0; ph < p->nphase; ph++)
        for (int g = p->ph_first[ph]; g < p->ph_first[ph + 1]; g++) p->ph_tan[ph][p->ph_greg[g]] = (float)coef[g];
    const int last = p->nphase - 1;
    for (int r = 0; r < 32; r++) {
        qgtl_cd a = cd(1.0, 0.0), b = cd(1.0, 0.0);
        for (int i = 0; i < 5; i++) {
            if (!((r >> i) & 1)) continue;
            const int gl = gate_of_bit[p->ph_rbits[last][i]], g0 = gate_of_bit[p->ph_rbits[0][i]];
            if (gl >= 0) {
                /* post layer: a swapped bit reads the other phase, signed under xfold = 0 */
                qgtl_cd f = swap[gl] ? cconj(rel_post[gl]) : rel_post[gl];
                if (swap[gl] && !p->xfold) f = cd(-f.re, -f.im);
                a = cmul(a, f);
            }
            if (g0 >= 0) {
                qgtl_cd f = rel_pre[g0];
                if (swap[g0] && p->xfold) f = cd(-f.re, -f.im);
                b = cmul(b, f);
            }
        }
        p->reg_post[r] = c64(a);
        p->reg_pre[r] = c64(b);
    }
}

/* Register passes that the register kernel runs (the LDS kernel applies its
 * own tables and ignores the moved layers). */
static int reg_pass_ok(const qgtl_pass_desc_t *p)
{
    return p->form == 1 && p->m == QGTL_TILE_BITS_MAX && p->nphase >= 1 && p->nphase <= QGTL_MAX_PHASES &&
           p->store != QGTL_STORE_PERM;
}

/* Move the phase layer of donor pass d into receiver pass r as a per-tile
 * factor (see qgtl_pass_desc_t): every gate of d must act on a tile-index bit
 * of r (r's tile is bits 0..12).  fwd = 1: d runs just before r and applies a
 * post layer only (xfold = 0); its layer in d's output index y is, per gate
 * bit, c diag(p0, p1) or, for a swapped gate, s diag(-p0, p1) (the post layer
 * read at y with the sign of y ^ xmask).  fwd = 0: d runs just after r and
 * applies a pre layer only (xfold = 1); its layer in d's input index is
 * c diag(q0, q1), or s diag(q0, -q1) for a swapped gate.  u: the matrices d
 * was finished with. */
static void fold_layer(qgtl_pass_desc_t *d, qgtl_pass_desc_t *r, const qgtl_mat2 *u, int fwd)
{
    if (!reg_pass_ok(d) || !reg_pass_ok(r) || r->nlow != QGTL_TILE_BITS_MAX || r->tf_n || r->nolayer || d->tf_n)
        return;
    if (fwd ? (d->has_pre || d->xfold != 0 || r->load != QGTL_LOAD_NATURAL)
            : (!d->has_pre || d->has_post || d->xfold != 1 || r->store != QGTL_STORE_INPLACE))
        return;
    if (d->ngates < 1 || d->ngates > QGTL_TILE_BITS_MAX) return;
    int bit[QGTL_TILE_BITS_MAX];
    for (int g = 0; g < d->ngates; g++) {
        const int tb = d->gate_bit[g];
        const int q = tb < d->nlow ? tb : d->hi_pos[tb - d->nlow];
        bit[g] = -1;
        for (int i = 0; i < r->nrest; i++)
            if (r->rest_pos[i] == q) bit[g] = i;
        if (bit[g] < 0) return;                 /* a gate inside r's tile: no fold */
    }
    for (int g = 0; g < d->ngates; g++) {
        const int tb = d->gate_bit[g];
        const int q = tb < d->nlow ? tb : d->hi_pos[tb - d->nlow];
        double c, sn;
        qgtl_cd pre[2], post[2];
        factor_rotation(&u[q], &c, &sn, pre, post);
        const int swapped = fabs(c) < fabs(sn);
        const double f = swapped ? sn : c;
        const qgtl_cd *ph = fwd ? post : pre;
        qgtl_cd f0 = cd(f * ph[0].re, f * ph[0].im), f1 = cd(f * ph[1].re, f * ph[1].im);
        if (swapped && fwd) f0 = cd(-f0.re, -f0.im);
        if (swapped && !fwd) f1 = cd(-f1.re, -f1.im);
        r->tf_bit[g] = bit[g];
        r->tf[g][0] = c64(f0);
        r->tf[g][1] = c64(f1);
    }
    r->tf_n = d->ngates;
    d->nolayer = 1;
}

/* LDS bank of tile index j in the register kernel's swizzled tile:
 * ((j ^ (j >> 4) ^ (j >> 8) ^ (j >> 12)) & 15), a GF(2)-linear map, so tile
 * bit b lands on bank bit (b & 3).  For the ring store the written index is
 * the linear ring image of j. */
static uint32_t bank_vec(uint32_t j)
{
    return (j ^ (j >> 4) ^ (j >> 8) ^ (j >> 12)) & 15u;
}

static uint32_t ring_image(uint32_t p)
{
    p ^= p << 1; p ^= p << 2; p ^= p << 4; p ^= p << 8;
    p &= (1u << QGTL_TILE_BITS_MAX) - 1u;
    return p ^ ((p >> (QGTL_TILE_BITS_MAX - 1)) & 1u);
}

static int independent4(const uint32_t v[4])
{
    uint32_t r[4] = {v[0], v[1], v[2], v[3]};
    for (int c = 0; c < 4; c++) {
        int piv = -1;
        for (int i = c; i < 4; i++)
            if ((r[i] >> c) & 1u) { piv = i; break; }
        if (piv < 0) {
            /* no pivot in column c among the remaining rows: look for any set bit */
            return 0;
        }
        const uint32_t t = r[c]; r[c] = r[piv]; r[piv] = t;
        for (int i = 0; i < 4; i++)
            if (i != c && ((r[i] >> c) & 1u)) r[i] ^= r[c];
    }
    return 1;
}

/* Thread-bit order of a layout with register bits rbits: the first four
 * thread bits (lanes 0..15) map to tile bits whose bank vectors are
 * independent, also through the ring map when the layout feeds a ring
 * store; phase 0 keeps ascending order (coalesced global access). */
static void choose_thread_bits(const int32_t rbits[5], int ascending, int ring, int32_t tbits[8])
{
    int others[QGTL_TILE_BITS_MAX], no = 0;
    for (int b = 0; b < QGTL_TILE_BITS_MAX; b++) {
        int isr = 0;
        for (int i = 0; i < 5; i++) isr |= rbits[i] == b;
        if (!isr) others[no++] = b;
    }
    for (int i = 0; i < 8; i++) tbits[i] = others[i];
    if (ascending) return;
    for (int a = 0; a < no; a++)
        for (int b = a + 1; b < no; b++)
            for (int c = b + 1; c < no; c++)
                for (int d = c + 1; d < no; d++) {
                    const int pick[4] = {others[a], others[b], others[c], others[d]};
                    uint32_t v[4], w[4];
                    for (int i = 0; i < 4; i++) {
                        v[i] = bank_vec(1u << pick[i]);
                        w[i] = bank_vec(ring_image(1u << pick[i]));
                    }
                    if (!independent4(v) || (ring && !independent4(w))) continue;
                    int k = 0;
                    for (int i = 0; i < 4; i++) tbits[k++] = pick[i];
                    for (int i = 0; i < no; i++)
                        if (others[i] != pick[0] && others[i] != pick[1] && others[i] != pick[2] && others[i] != pick[3])
                            tbits[k++] = others[i];
                    return;
                }
}

/* Register-phase plan for a full 2^13 tile (see qgtl_pass_desc_t): phase 0
 * holds tile bits {0, 7, 8, 9, 10}; the remaining gate bits go five per phase
 * (topped up with other tile bits).  Gates are reordered by phase (they
 * commute: one matrix per qubit). */
static void plan_phases(qgtl_pass_desc_t *p)
{
    static const int8_t ph0[5] = {0, 7, 8, 9, 10};
    p->nphase = 0;
    if (p->m != QGTL_TILE_BITS_MAX || p->store == QGTL_STORE_PERM) return;
    int in0[QGTL_TILE_BITS_MAX] = {0};
    for (int i = 0; i < 5; i++) in0[ph0[i]] = 1;
    int order[QGTL_TILE_BITS_MAX], cnt = 0;
    int rest[QGTL_TILE_BITS_MAX], nrest = 0;
    for (int i = 0; i < 5; i++) p->ph_rbits[0][i] = ph0[i];
    p->ph_first[0] = 0;
    for (int g = 0; g < p->ngates; g++)
        if (in0[p->gate_bit[g]]) {
            for (int i = 0; i < 5; i++)
                if (ph0[i] == p->gate_bit[g]) p->ph_greg[cnt] = (int8_t)i;
            order[cnt++] = g;
        } else {
            rest[nrest++] = g;
        }
    int nph = 1;
    for (int r0 = 0; r0 < nrest; r0 += 5, nph++) {
        p->ph_first[nph] = (int8_t)cnt;
        const int len = (nrest - r0 < 5) ? nrest - r0 : 5;
        int used[QGTL_TILE_BITS_MAX] = {0};
        for (int i = 0; i < len; i++) {
            const int g = rest[r0 + i];
            p->ph_rbits[nph][i] = p->gate_bit[g];
            used[p->gate_bit[g]] = 1;
            p->ph_greg[cnt] = (int8_t)i;
            order[cnt++] = g;
        }
        for (int i = len, b = 0; i < 5; i++) {
            while (used[b]) b++;
            p->ph_rbits[nph][i] = (int8_t)b;
            used[b] = 1;
        }
    }
    p->ph_first[nph] = (int8_t)cnt;
    p->nphase = nph;
    /* reorder matrices and their tile bits into phase order */
    qgtl_c64_t mats[QGTL_TILE_BITS_MAX][4], rot[QGTL_TILE_BITS_MAX], pre[QGTL_TILE_BITS_MAX][2],
        post[QGTL_TILE_BITS_MAX][2];
    int32_t bits[QGTL_TILE_BITS_MAX];
    for (int i = 0; i < cnt; i++) {
        memcpy(mats[i], p->mat[order[i]], sizeof mats[i]);
        rot[i] = p->rot[order[i]];
        memcpy(pre[i], p->pre[order[i]], sizeof pre[i]);
        memcpy(post[i], p->post[order[i]], sizeof post[i]);
        bits[i] = p->gate_bit[order[i]];
    }
    memcpy(p->mat, mats, sizeof(qgtl_c64_t) * 4 * (size_t)cnt);
    memcpy(p->rot, rot, sizeof(qgtl_c64_t) * (size_t)cnt);
    memcpy(p->pre, pre, sizeof(qgtl_c64_t) * 2 * (size_t)cnt);
    memcpy(p->post, post, sizeof(qgtl_c64_t) * 2 * (size_t)cnt);
    memcpy(p->gate_bit, bits, sizeof(int32_t) * (size_t)cnt);
    /* thread-bit orders (the last layout also writes the ring store) */
    for (int ph = 0; ph < nph; ph++)
        choose_thread_bits(p->ph_rbits[ph], ph == 0, ph == nph - 1 && p->store == QGTL_STORE_RING, p->ph_tbits[ph]);
    /* by register bit, for the kernels */
    for (int ph = 0; ph < nph; ph++) {
        p->ph_mask[ph] = 0;
        for (int g = p->ph_first[ph]; g < p->ph_first[ph + 1]; g++) {
            const int j = p->ph_greg[g];
            p->ph_mask[ph] |= 1 << j;
            p->ph_rot[ph][j] = p->rot[g];
            memcpy(p->ph_mat[ph][j], p->mat[g], sizeof(qgtl_c64_t) * 4);
        }
    }
}

/* High-qubit passes of a block with n > 13 (matrices u[q]): the qubits >= 13
 * that carry a gate, in groups that leave CMIN_LOW_BITS contiguous low bits
 * in the tile; in place. */
static void high_passes(const qgtl_block_t *b, int n, const qgtl_mat2 *u, qgtl_pass_desc_t *out, int *flips,
                        int *count)
{
    const int TB = QGTL_TILE_BITS_MAX;
    int hq[QGTL_MAX_QUBITS], nh = 0;
    for (int q = TB; q < n; q++)
        if (b->has_u[q]) hq[nh++] = q;
    const int group = TB - CMIN_LOW_BITS;
    for (int g0 = 0; g0 < nh; g0 += group) {
        const int len = (nh - g0 < group) ? nh - g0 : group;
        const int c = TB - len;
        qgtl_pass_desc_t *p = &out[*count];
        memset(p, 0, sizeof *p);
        p->n = n; p->m = TB; p->nlow = c;
        int in_tile[QGTL_MAX_QUBITS] = {0};
        for (int i = 0; i < c; i++) in_tile[i] = 1;
        for (int i = 0; i < len; i++) {
            p->hi_pos[i] = (int8_t)hq[g0 + i];
            in_tile[hq[g0 + i]] = 1;
            p->gate_bit[i] = (int8_t)(c + i);
            set_gate(p, i, &u[hq[g0 + i]]);
        }
        p->ngates = len;
        for (int q = 0; q < n; q++)
            if (!in_tile[q]) p->rest_pos[p->nrest++] = (int8_t)q;
        p->store = QGTL_STORE_INPLACE;
        flips[*count] = 0;
        (*count)++;
    }
}

/* The pass over tile bits 0..12 (matrices u[q], q < 13) with the given load
 * and store; returns 0 when it has nothing to do. */
static int low_pass(const qgtl_block_t *b, int n, const qgtl_mat2 *u, qgtl_pass_desc_t *p, int store, int load)
{
    const int TB = QGTL_TILE_BITS_MAX;
    memset(p, 0, sizeof *p);
    p->n = n; p->m = TB; p->nlow = TB;
    for (int q = TB; q < n; q++) p->rest_pos[p->nrest++] = (int8_t)q;
    for (int q = 0; q < TB; q++)
        if (b->has_u[q]) {
            p->gate_bit[p->ngates] = (int8_t)q;
            set_gate(p, p->ngates, &u[q]);
            p->ngates++;
        }
    p->store = store;
    p->load = load;
    return p->ngates > 0 || store != QGTL_STORE_INPLACE || load != QGTL_LOAD_NATURAL;
}

static void build_passes(qgtl_block_t *b, int n)
{
    const int TB = QGTL_TILE_BITS_MAX;
    b->npasses = 0;
    b->mstage_kernels = 0;
    if (n <= TB) {
        qgtl_pass_desc_t *p = &b->pass[0];
        memset(p, 0, sizeof *p);
        p->n = n; p->m = n; p->nlow = n; p->nrest = 0;
        for (int q = 0; q < n; q++)
            if (b->has_u[q]) {
                p->gate_bit[p->ngates] = (int8_t)q;
                set_gate(p, p->ngates, &b->u[q]);
                p->ngates++;
            }
        p->store = QGTL_STORE_INPLACE;
        if (b->nperm > 0) {
            if (b->nperm <= QGTL_PERM_MAX) {
                p->store = QGTL_STORE_PERM;
                p->nperm = b->nperm;
                for (int i = 0; i < b->nperm; i++) {
                    p->perm_type[i] = b->perm_type[i];
                    p->perm_a[i] = b->perm_a[i];
                    p->perm_b[i] = b->perm_b[i];
                }
            } else {
                b->mstage_kernels = 1;
            }
        }
        if (p->ngates > 0 || p->store != QGTL_STORE_INPLACE) {
            b->pass_flips[0] = 0;
            b->npasses = 1;
        }
        return;
    }
    high_passes(b, n, b->u, b->pass, b->pass_flips, &b->npasses);
    if (low_pass(b, n, b->u, &b->pass[b->npasses], b->is_ring ? QGTL_STORE_RING : QGTL_STORE_INPLACE,
                 QGTL_LOAD_NATURAL)) {
        b->pass_flips[b->npasses] = b->is_ring;
        b->npasses++;
    }
    if (!b->is_ring && b->nperm > 0) b->mstage_kernels = 1;
    for (int i = 0; i < b->npasses; i++) plan_phases(&b->pass[i]);
}

/* The inverse of block b (n > 13): the inverse ring as a gather into the low
 * pass, which also applies u_q^dagger on bits 0..12, then the high passes with
 * u_q^dagger.  Blocks whose M-stage is neither the ring nor empty have none. */
static void build_inverse_passes(qgtl_block_t *b, int n)
{
    b->ninv = 0;
    b->inv_ok = 0;
    if (n <= QGTL_TILE_BITS_MAX || (!b->is_ring && b->nperm > 0)) return;
    qgtl_mat2 uadj[QGTL_MAX_QUBITS];
    for (int q = 0; q < n; q++) uadj[q] = mat_adj(&b->u[q]);
    if (low_pass(b, n, uadj, &b->ipass[0], QGTL_STORE_INPLACE, b->is_ring ? QGTL_LOAD_RING : QGTL_LOAD_NATURAL)) {
        b->ipass_flips[0] = b->is_ring;
        b->ninv = 1;
    }
    int nh = 0;
    high_passes(b, n, uadj, &b->ipass[b->ninv], &b->ipass_flips[b->ninv], &nh);
    b->ninv += nh;
    for (int i = 0; i < b->ninv; i++) {
        qgtl_pass_desc_t *p = &b->ipass[i];
        plan_phases(p);
        if (p->load == QGTL_LOAD_RING) choose_thread_bits(p->ph_rbits[0], 0, 1, p->gather_tbits);
    }
    b->inv_ok = 1;
}

static void detect_ring(qgtl_block_t *b, int n)
{
    b->is_ring = 0;
    if (n < 2 || b->nperm != n) return;
    for (int i = 0; i < n; i++)
        if (b->perm_type[i] != 0 || b->perm_a[i] != i || b->perm_b[i] != (i + 1) % n) return;
    b->is_ring = 1;
}

typedef struct {
    int type, q0, q1, param, op_index;
    double angle;
} flat_gate_t;

/* Expand CZ / CY into single-qubit gates and CNOTs. */
static int expand(const struct qgtl_plan *plan, flat_gate_t **out, int *count)
{
    flat_gate_t *g = (flat_gate_t *)malloc(sizeof(flat_gate_t) * (size_t)(3 * plan->nops + 1));
    if (!g) return QGT_ERROR_NO_MEMORY;
    int k = 0;
    for (int i = 0; i < plan->nops; i++) {
        const qgtl_op_t *op = &plan->ops[i];
        flat_gate_t base = {op->type, op->q0, op->q1, op->param, i, op->angle};
        if (op->type == GATE_TYPE_CZ) {
            flat_gate_t h = {GATE_TYPE_H, op->q1, -1, -1, i, 0.0}, cx = {GATE_TYPE_CNOT, op->q0, op->q1, -1, i, 0.0};
            g[k++] = h; g[k++] = cx; g[k++] = h;
        } else if (op->type == GATE_TYPE_CY) {
            flat_gate_t sd = {GATE_TYPE_SDG, op->q1, -1, -1, i, 0.0}, cx = {GATE_TYPE_CNOT, op->q0, op->q1, -1, i, 0.0};
            flat_gate_t s = {GATE_TYPE_S, op->q1, -1, -1, i, 0.0};
            g[k++] = sd; g[k++] = cx; g[k++] = s;
        } else {
            g[k++] = base;
        }
    }
    *out = g;
    *count = k;
    return 0;
}

static qgtl_block_t *open_block(qgtl_block_t *blocks, int i, int n)
{
    qgtl_block_t *b = &blocks[i];
    memset(b, 0, sizeof *b);
    for (int q = 0; q < n; q++) b->u[q] = mat_identity();
    return b;
}

int qgtl_plan_build(struct qgtl_plan *plan)
{
    const int n = plan->n;
    plan->build_gen++;                   /* any captured launch sequence is stale */
    if (n < 1 || n > QGTL_MAX_QUBITS)
        return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "qubit count %d outside 1..%d", n, QGTL_MAX_QUBITS);
    /* validate ops and count parameters */
    int P = 0;
    for (int i = 0; i < plan->nops; i++) {
        const qgtl_op_t *op = &plan->ops[i];
        if (op->q0 < 0 || op->q0 >= n)
            return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "gate %d: qubit %d out of range", i, op->q0);
        const int two = (op->type == GATE_TYPE_CNOT || op->type == GATE_TYPE_CZ || op->type == GATE_TYPE_CY ||
                         op->type == GATE_TYPE_SWAP);
        if (two && (op->q1 < 0 || op->q1 >= n || op->q1 == op->q0))
            return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "gate %d: bad second qubit %d", i, op->q1);
        if (!two && !is_single_qubit(op->type))
            return qgtl_set_error(QGT_ERROR_NOT_IMPLEMENTED, "gate %d: gate type %d not supported on this path", i, op->type);
        if (op->param >= 0) {
            if (!qgtl_is_param_gate(op->type))
                return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "gate %d: only RX/RY/RZ carry parameters", i);
            P++;
        }
    }
    int *seen = (int *)calloc((size_t)(P > 0 ? P : 1), sizeof(int));
    if (!seen) return QGT_ERROR_NO_MEMORY;
    for (int i = 0; i < plan->nops; i++) {
        const int k = plan->ops[i].param;
        if (k < 0) continue;
        if (k >= P || seen[k]) {
            free(seen);
            return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "parameter indices must be 0..P-1, each once (gate %d)", i);
        }
        seen[k] = 1;
    }
    free(seen);
    plan->P = P;

    flat_gate_t *g = NULL;
    int ng = 0;
    int rc = expand(plan, &g, &ng);
    if (rc) return rc;

    /* the block array is kept across rebuilds (set_angles): each block is
     * cleared when it is opened, so only the blocks in use are touched */
    if (!plan->blocks || plan->blocks_cap < ng + 1) {
        free(plan->blocks);
        plan->blocks = (qgtl_block_t *)malloc((size_t)(ng + 1) * sizeof(qgtl_block_t));
        plan->blocks_cap = plan->blocks ? ng + 1 : 0;
    }
    free(plan->params);
    plan->params = (qgtl_param_t *)calloc((size_t)(P > 0 ? P : 1), sizeof(qgtl_param_t));
    qgtl_mat2 *u_at = (qgtl_mat2 *)calloc((size_t)(P > 0 ? P : 1), sizeof(qgtl_mat2));
    int *order = (int *)calloc((size_t)(P > 0 ? P : 1), sizeof(int));   /* params in creation order */
    if (!plan->blocks || !plan->params || !u_at || !order) {
        free(g); free(u_at); free(order);
        return QGT_ERROR_NO_MEMORY;
    }
    int nb = 0, npar = 0;
    int touched[QGTL_MAX_QUBITS];
    qgtl_block_t *cur = open_block(plan->blocks, 0, n);
    memset(touched, 0, sizeof touched);
    cur->param_begin = 0;
    int block_open = 0;

    for (int i = 0; i < ng; i++) {
        const flat_gate_t *op = &g[i];
        const int twoq = (op->type == GATE_TYPE_CNOT || op->type == GATE_TYPE_SWAP);
        if (!twoq && touched[op->q0]) {
            /* close the block */
            cur->param_end = npar;
            nb++;
            cur = open_block(plan->blocks, nb, n);
            memset(touched, 0, sizeof touched);
            cur->param_begin = npar;
            block_open = 0;
        }
        block_open = 1;
        if (twoq) {
            if (cur->nperm >= QGTL_PERM_MAX) {
                /* start a fresh block rather than overflow the M-stage list */
                cur->param_end = npar;
                nb++;
                cur = open_block(plan->blocks, nb, n);
                memset(touched, 0, sizeof touched);
                cur->param_begin = npar;
            }
            cur->perm_type[cur->nperm] = (int8_t)(op->type == GATE_TYPE_SWAP ? 1 : 0);
            cur->perm_a[cur->nperm] = (int8_t)op->q0;
            cur->perm_b[cur->nperm] = (int8_t)op->q1;
            cur->nperm++;
            touched[op->q0] = touched[op->q1] = 1;
        } else {
            qgtl_op_t tmp = {op->type, op->q0, -1, -1, op->angle};
            qgtl_mat2 m;
            qgtl_gate_matrix(&tmp, &m);
            cur->u[op->q0] = mat_mul(&m, &cur->u[op->q0]);
            cur->has_u[op->q0] = 1;
            if (op->param >= 0) {
                qgtl_param_t *pp = &plan->params[op->param];
                pp->block = nb;
                pp->qubit = op->q0;
                pp->op_index = op->op_index;
                u_at[op->param] = cur->u[op->q0];
                /* remember the generator through A (finished at block close) */
                pp->A = generator(op->type);
                order[npar++] = op->param;
            }
        }
    }
    if (block_open || nb == 0) {
        cur->param_end = npar;
        nb++;
    }
    plan->nblocks = nb;

    /* finish blocks: ring detection, passes, injection operators */
    for (int bi = 0; bi < nb; bi++) {
        qgtl_block_t *b = &plan->blocks[bi];
        detect_ring(b, n);
        build_passes(b, n);
        if (plan->mode == QGTL_MODE_MIXED && plan->tune.sweep_schedule != 1) build_inverse_passes(b, n);
        for (int i = 0; i < b->npasses; i++) {
            finish_form(&b->pass[i], plan->tune.sweep_form == 1);
            finish_rotation(&b->pass[i], b->u);
        }
        if (b->npasses == 2) fold_layer(&b->pass[0], &b->pass[1], b->u, 1);
        if (b->ninv > 0) {
            qgtl_mat2 uadj[QGTL_MAX_QUBITS];
            for (int q = 0; q < n; q++) uadj[q] = mat_adj(&b->u[q]);
            for (int i = 0; i < b->ninv; i++) {
                finish_form(&b->ipass[i], plan->tune.sweep_form == 1);
                finish_rotation(&b->ipass[i], uadj);
            }
            if (b->ninv == 2) fold_layer(&b->ipass[1], &b->ipass[0], uadj, 0);
        }
        /* images of the basis vectors under the inverse M-stage (the Z-frame masks below) */
        uint32_t inv_img[QGTL_MAX_QUBITS];
        if (b->param_end > b->param_begin)
            for (int j2 = 0; j2 < n; j2++) inv_img[j2] = perm_apply_inverse(b, 1u << j2);
        for (int j = b->param_begin; j < b->param_end; j++) {
            qgtl_param_t *pp = &plan->params[order[j]];
            const int q = pp->qubit;
            qgtl_mat2 at_adj = mat_adj(&u_at[order[j]]);
            qgtl_mat2 S = mat_mul(&b->u[q], &at_adj);           /* gates after the parameter */
            qgtl_mat2 Sadj = mat_adj(&S);
            qgtl_mat2 t = mat_mul(&S, &pp->A);
            pp->A = mat_mul(&t, &Sadj);
            const qgtl_mat2 uadj = mat_adj(&b->u[q]);
            t = mat_mul(&uadj, &pp->A);
            pp->B = mat_mul(&t, &b->u[q]);
            pp->v = perm_apply(b, 1u << q);
            uint32_t w = 0;
            for (int j2 = 0; j2 < n; j2++)
                if ((inv_img[j2] >> q) & 1u) w |= 1u << j2;
            pp->w = w;
        }
    }
    /* D slot j holds the j-th parameter created (gate order); slot_param maps
     * slots back to the caller's parameter numbers. */
    free(plan->slot_param);
    plan->slot_param = order;
    order = NULL;
    free(order);
    free(u_at);
    free(g);
    plan->tabs_dirty = 1;                                     /* new angles: phase tables to rebuild */
    plan->meet = nb - 1;
    return 0;
}

/* The rotation-form phase tables of pass p (the kernel's build_phase_tables,
 * here in FP64 and rounded once): out[t], t < QGTL_PHASE_TAB, the post phases,
 * out[QGTL_PHASE_TAB + t] the pre phases, as interleaved float pairs. */
void qgtl_pass_phase_tables(const qgtl_pass_desc_t *p, float *out)
{
    for (int t = 0; t < QGTL_PHASE_TAB; t++) {
        const int lo = t < 128;
        const int i = lo ? t : t - 128;
        double ar = 1.0, ai = 0.0, br = 1.0, bi = 0.0;
        for (int g = 0; g < p->ngates; g++) {
            const int bit = p->gate_bit[g];
            if (lo ? bit >= 7 : bit < 7) continue;
            const int v = (i >> (lo ? bit : bit - 7)) & 1;
            const double pr = p->post[g][v].re, pi = p->post[g][v].im;
            const double qr = p->pre[g][v].re, qi = p->pre[g][v].im;
            const double a2 = ar * pr - ai * pi, b2 = br * qr - bi * qi;
            ai = ar * pi + ai * pr;
            ar = a2;
            bi = br * qi + bi * qr;
            br = b2;
        }
        out[2 * t] = (float)ar;
        out[2 * t + 1] = (float)ai;
        out[2 * (QGTL_PHASE_TAB + t)] = (float)br;
        out[2 * (QGTL_PHASE_TAB + t) + 1] = (float)bi;
    }
}

/* Diagnostics: max |diag(post) R diag(pre) - U| over every gate of every pass
 * (FP32 descriptor values, FP64 arithmetic) and the number of passes that
 * need a pre-phase layer. */
int qgtl_plan_debug_form(const qgtl_plan_t *plan, double *max_err, int *passes_with_pre)
{
    if (!plan) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null plan");
    double err = 0.0;
    int npre = 0;
    for (int bi = 0; bi < plan->nblocks; bi++)
        for (int i = 0; i < plan->blocks[bi].npasses; i++) {
            const qgtl_pass_desc_t *p = &plan->blocks[bi].pass[i];
            npre += p->has_pre;
            for (int g = 0; g < p->ngates; g++) {
                const double c = p->rot[g].re, sn = p->rot[g].im;
                const double R[2][2] = {{c, -sn}, {sn, c}};
                for (int r = 0; r < 2; r++)
                    for (int k = 0; k < 2; k++) {
                        const qgtl_cd po = cd(p->post[g][r].re, p->post[g][r].im);
                        const qgtl_cd pr = cd(p->pre[g][k].re, p->pre[g][k].im);
                        const qgtl_cd v = cmul(po, cmul(cd(R[r][k], 0.0), pr));
                        const double dr = v.re - p->mat[g][2 * r + k].re, di = v.im - p->mat[g][2 * r + k].im;
                        const double e = sqrt(dr * dr + di * di);
                        if (e > err) err = e;
                    }
            }
        }
    if (max_err) *max_err = err;
    if (passes_with_pre) *passes_with_pre = npre;
    return 0;
}
