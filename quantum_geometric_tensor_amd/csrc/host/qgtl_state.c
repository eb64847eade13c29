/*
 * qgtl_state.c -- QGTL's quantum_state_* API (include/quantum_geometric/core/quantum_state.h:15-98)
 * with QGTL's semantics: FP64 accumulation, FP32 amplitudes
 * (src/quantum_geometric/physics/quantum_state_operations.c:313-439).
 * inner_product, norm, clone and initialize_zero are declared but never
 * defined in QGTL; they are defined here with the obvious meaning.
 *
 * Device residency (quantum_state_to_device / _from_device / _is_on_device,
 * quantum_state.h:93-98, declared but never defined in QGTL): a state moved to
 * the GPU (HARDWARE_TYPE_GPU) gets a complex64 copy in device memory, kept in
 * a side table of this library (the struct's own fields keep their QGTL
 * meaning; `hardware` becomes HARDWARE_TYPE_GPU).  While a state is on the
 * device its device copy is the authoritative one: norm, normalize,
 * is_normalized, inner_product and fidelity run there (FP64 reductions in a
 * fixed order, FP32 scale as on the host), initialize/initialize_basis write
 * both copies, clone copies the device values, and quantum_state_from_device
 * brings `coordinates` up to date and frees the device copy.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_api.h"
#include "qgtl_internal.h"

static size_t log2_exact(size_t d)
{
    size_t n = 0;
    while (((size_t)1 << n) < d) n++;
    return ((size_t)1 << n) == d ? n : 0;
}

/* ------------------------------------------------------- device copies */
typedef struct resident {
    const quantum_state_t *state;
    void *dev;                        /* dimension complex64 values */
    struct resident *next;
} resident_t;

static resident_t *g_resident;        /* guarded by the library lock */
static double *g_partials, *g_sums;   /* reduction scratch on the device */

static resident_t *find_resident(const quantum_state_t *s)
{
    for (resident_t *r = g_resident; r; r = r->next)
        if (r->state == s) return r;
    return NULL;
}

static void *device_copy(const quantum_state_t *s)
{
    resident_t *r = find_resident(s);
    return r ? r->dev : NULL;
}

/* a failed device call: its QGT_ERROR_GPU_* code (the message in qgtl_last_error) */
static qgt_error_t hip_error(int rc)
{
    return rc < 0 ? (qgt_error_t)rc : QGT_ERROR_HARDWARE_FAILURE;
}

static int ensure_scratch(void)
{
    int rc = qgtl_init_if_needed();
    if (rc) return rc;
    if (!g_partials) {
        rc = qgtl_hip_malloc((void **)&g_partials, sizeof(double) * (size_t)qgtl_hip_vec_partials_len());
        if (rc) return rc;
        rc = qgtl_hip_malloc((void **)&g_sums, sizeof(double) * 4);
        if (rc) { qgtl_hip_free(g_partials); g_partials = NULL; return rc; }
    }
    return 0;
}

/* out4 = Re <a|b>, Im <a|b>, |a|^2, |b|^2 over n amplitudes, each operand a
 * device copy or else uploaded for the call */
static int device_dotc(const ComplexFloat *ha, void *da, const ComplexFloat *hb, void *db, size_t n, double out4[4])
{
    int rc = ensure_scratch();
    if (rc) return rc;
    void *ta = NULL, *tb = NULL;
    if (!da) {
        rc = qgtl_hip_malloc(&ta, n * sizeof(ComplexFloat));
        if (!rc) rc = qgtl_hip_h2d(ta, ha, n * sizeof(ComplexFloat));
        da = ta;
    }
    if (!rc && !db) {
        rc = qgtl_hip_malloc(&tb, n * sizeof(ComplexFloat));
        if (!rc) rc = qgtl_hip_h2d(tb, hb, n * sizeof(ComplexFloat));
        db = tb;
    }
    if (!rc) rc = qgtl_hip_vec_dotc_c64(da, db, n, g_partials, g_sums);
    if (!rc) rc = qgtl_hip_d2h(out4, g_sums, 4 * sizeof(double));
    if (!rc) rc = qgtl_hip_sync();
    qgtl_hip_sync();
    qgtl_hip_free(ta);
    qgtl_hip_free(tb);
    return rc;
}

/* |state|^2 in FP64, on the device when the state is there */
static int state_norm_sq(const quantum_state_t *s, double *ns)
{
    void *d = device_copy(s);
    if (!d) {
        double acc = 0.0;
        for (size_t i = 0; i < s->dimension; i++)
            acc += (double)s->coordinates[i].real * s->coordinates[i].real +
                   (double)s->coordinates[i].imag * s->coordinates[i].imag;
        *ns = acc;
        return 0;
    }
    double out[4];
    int rc = device_dotc(NULL, d, NULL, d, s->dimension, out);
    if (rc) return rc;
    *ns = out[2];
    return 0;
}

/* <a|b> in FP64 (on the device when either state is there) */
static int state_dotc(const quantum_state_t *a, const quantum_state_t *b, double *re, double *im)
{
    void *da = device_copy(a), *db = device_copy(b);
    if (!da && !db) {
        double sr = 0.0, si = 0.0;
        for (size_t i = 0; i < a->dimension; i++) {
            const double ar = a->coordinates[i].real, ai = a->coordinates[i].imag;
            const double br = b->coordinates[i].real, bi = b->coordinates[i].imag;
            sr += ar * br + ai * bi;
            si += ar * bi - ai * br;
        }
        *re = sr;
        *im = si;
        return 0;
    }
    double out[4];
    int rc = device_dotc(a->coordinates, da, b->coordinates, db, a->dimension, out);
    if (rc) return rc;
    *re = out[0];
    *im = out[1];
    return 0;
}

/* multiply every amplitude by s, on the device copy when there is one */
static int state_scale(quantum_state_t *st, float s)
{
    void *d = device_copy(st);
    if (!d) {
        for (size_t i = 0; i < st->dimension; i++) {
            st->coordinates[i].real *= s;
            st->coordinates[i].imag *= s;
        }
        return 0;
    }
    int rc = qgtl_hip_vec_scale_c64(d, st->dimension, s);
    if (!rc) rc = qgtl_hip_sync();
    return rc;
}

/* after the host coordinates changed: refresh the device copy */
static int state_push(const quantum_state_t *st)
{
    void *d = device_copy(st);
    if (!d) return 0;
    int rc = qgtl_hip_h2d(d, st->coordinates, st->dimension * sizeof(ComplexFloat));
    if (!rc) rc = qgtl_hip_sync();
    return rc;
}

static void drop_resident(const quantum_state_t *s)
{
    for (resident_t **pr = &g_resident; *pr; pr = &(*pr)->next)
        if ((*pr)->state == s) {
            resident_t *r = *pr;
            *pr = r->next;
            qgtl_hip_free(r->dev);
            free(r);
            return;
        }
}

static qgt_error_t hip_status(int rc) { return rc ? hip_error(rc) : QGT_SUCCESS; }

/* ------------------------------------------------------------ the API */
qgt_error_t quantum_state_create(quantum_state_t **state, quantum_state_type_t type, size_t dimension)
{
    if (!state || dimension == 0) return QGT_ERROR_INVALID_ARGUMENT;
    quantum_state_t *s = (quantum_state_t *)calloc(1, sizeof *s);
    if (!s) return QGT_ERROR_ALLOCATION_FAILED;
    s->coordinates = (ComplexFloat *)calloc(dimension, sizeof(ComplexFloat));
    if (!s->coordinates) { free(s); return QGT_ERROR_ALLOCATION_FAILED; }
    s->type = type;
    s->dimension = dimension;
    s->num_qubits = log2_exact(dimension);
    s->coordinates[0].real = 1.0f;
    s->is_normalized = true;
    s->hardware = HARDWARE_TYPE_CPU;
    *state = s;
    return QGT_SUCCESS;
}

void quantum_state_destroy(quantum_state_t *state)
{
    if (!state) return;
    {
        QGTL_LOCKED();
        drop_resident(state);
    }
    free(state->coordinates);
    free(state->auxiliary_data);
    free(state);
}

qgt_error_t quantum_state_clone(quantum_state_t **dest, const quantum_state_t *src)
{
    QGTL_LOCKED();
    if (!dest || !src || !src->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    qgt_error_t rc = quantum_state_create(dest, src->type, src->dimension);
    if (rc != QGT_SUCCESS) return rc;
    void *d = device_copy(src);
    if (d) {
        int hr = qgtl_hip_d2h((*dest)->coordinates, d, src->dimension * sizeof(ComplexFloat));
        if (!hr) hr = qgtl_hip_sync();
        if (hr) { quantum_state_destroy(*dest); *dest = NULL; return hip_error(hr); }
    } else {
        memcpy((*dest)->coordinates, src->coordinates, src->dimension * sizeof(ComplexFloat));
    }
    (*dest)->num_qubits = src->num_qubits;
    (*dest)->manifold_dim = src->manifold_dim;
    (*dest)->is_normalized = src->is_normalized;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_initialize(quantum_state_t *state, const ComplexFloat *amplitudes)
{
    QGTL_LOCKED();
    if (!state || !state->coordinates || !amplitudes) return QGT_ERROR_INVALID_ARGUMENT;
    memcpy(state->coordinates, amplitudes, state->dimension * sizeof(ComplexFloat));
    double ns = 0.0;
    for (size_t i = 0; i < state->dimension; i++)
        ns += (double)amplitudes[i].real * amplitudes[i].real + (double)amplitudes[i].imag * amplitudes[i].imag;
    if (ns < 1e-15) return QGT_ERROR_INVALID_STATE;
    if (fabs(ns - 1.0) >= 1e-6) {
        const float inv = 1.0f / sqrtf((float)ns);
        for (size_t i = 0; i < state->dimension; i++) {
            state->coordinates[i].real *= inv;
            state->coordinates[i].imag *= inv;
        }
    }
    state->is_normalized = true;
    return hip_status(state_push(state));
}

qgt_error_t quantum_state_initialize_basis(quantum_state_t *state, size_t basis_index)
{
    QGTL_LOCKED();
    if (!state || !state->coordinates || basis_index >= state->dimension) return QGT_ERROR_INVALID_ARGUMENT;
    memset(state->coordinates, 0, state->dimension * sizeof(ComplexFloat));
    state->coordinates[basis_index].real = 1.0f;
    state->is_normalized = true;
    return hip_status(state_push(state));
}

qgt_error_t quantum_state_initialize_zero(quantum_state_t *state) { return quantum_state_initialize_basis(state, 0); }

qgt_error_t quantum_state_inner_product(ComplexFloat *result, const quantum_state_t *a, const quantum_state_t *b)
{
    QGTL_LOCKED();
    if (!result || !a || !b || !a->coordinates || !b->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    if (a->dimension != b->dimension) return QGT_ERROR_INCOMPATIBLE;
    double re = 0.0, im = 0.0;
    const int rc = state_dotc(a, b, &re, &im);
    if (rc) return hip_error(rc);
    result->real = (float)re;
    result->imag = (float)im;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_norm(float *norm, const quantum_state_t *state)
{
    QGTL_LOCKED();
    if (!norm || !state || !state->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    double ns = 0.0;
    const int rc = state_norm_sq(state, &ns);
    if (rc) return hip_error(rc);
    *norm = (float)sqrt(ns);
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_normalize(quantum_state_t *state)
{
    QGTL_LOCKED();
    if (!state || !state->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    double ns = 0.0;
    int rc = state_norm_sq(state, &ns);
    if (rc) return hip_error(rc);
    if (ns < 1e-15) return QGT_ERROR_INVALID_STATE;
    rc = state_scale(state, 1.0f / sqrtf((float)ns));
    if (rc) return hip_error(rc);
    state->is_normalized = true;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_fidelity(float *fidelity, const quantum_state_t *a, const quantum_state_t *b)
{
    QGTL_LOCKED();
    if (!fidelity || !a || !b || !a->coordinates || !b->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    if (a->dimension != b->dimension) return QGT_ERROR_INCOMPATIBLE;
    /* F = |<a|b>|^2 */
    double re = 0.0, im = 0.0;
    const int rc = state_dotc(a, b, &re, &im);
    if (rc) return hip_error(rc);
    *fidelity = (float)(re * re + im * im);
    return QGT_SUCCESS;
}

bool quantum_state_is_normalized(const quantum_state_t *state)
{
    QGTL_LOCKED();
    if (!state || !state->coordinates) return false;
    double ns = 0.0;
    if (state_norm_sq(state, &ns)) return false;
    return fabs(ns - 1.0) < 1e-6;
}

qgt_error_t quantum_state_to_device(quantum_state_t *state, quantum_hardware_t hardware)
{
    QGTL_LOCKED();
    if (!state || !state->coordinates || state->dimension == 0) return QGT_ERROR_INVALID_ARGUMENT;
    if (hardware.type == HARDWARE_TYPE_CPU) return quantum_state_from_device(state, hardware);
    if (hardware.type != HARDWARE_TYPE_GPU) return QGT_ERROR_NOT_IMPLEMENTED;
    if (find_resident(state)) return QGT_SUCCESS;
    int rc = qgtl_init_if_needed();
    if (rc) return hip_error(rc);
    resident_t *r = (resident_t *)calloc(1, sizeof *r);
    if (!r) return QGT_ERROR_ALLOCATION_FAILED;
    rc = qgtl_hip_malloc(&r->dev, state->dimension * sizeof(ComplexFloat));
    if (!rc) rc = qgtl_hip_h2d(r->dev, state->coordinates, state->dimension * sizeof(ComplexFloat));
    if (!rc) rc = qgtl_hip_sync();
    if (rc) {
        qgtl_hip_free(r->dev);
        free(r);
        return hip_error(rc);
    }
    r->state = state;
    r->next = g_resident;
    g_resident = r;
    state->hardware = HARDWARE_TYPE_GPU;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_from_device(quantum_state_t *state, quantum_hardware_t hardware)
{
    QGTL_LOCKED();
    (void)hardware;
    if (!state || !state->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    resident_t *r = find_resident(state);
    if (!r) return QGT_SUCCESS;                    /* already on the host */
    int rc = qgtl_hip_d2h(state->coordinates, r->dev, state->dimension * sizeof(ComplexFloat));
    if (!rc) rc = qgtl_hip_sync();
    if (rc) return hip_error(rc);
    drop_resident(state);
    state->hardware = HARDWARE_TYPE_CPU;
    return QGT_SUCCESS;
}

bool quantum_state_is_on_device(const quantum_state_t *state, quantum_hardware_t hardware)
{
    QGTL_LOCKED();
    return state && hardware.type == HARDWARE_TYPE_GPU && find_resident(state) != NULL;
}

/* Additive: the device copy of a state on the GPU (complex64, `dimension`
 * values), or NULL. */
void *qgtl_state_device_ptr(const quantum_state_t *state)
{
    QGTL_LOCKED();
    return state ? device_copy(state) : NULL;
}
