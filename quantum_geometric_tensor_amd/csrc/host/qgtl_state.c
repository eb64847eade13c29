/*
 * qgtl_state.c -- QGTL's quantum_state_* API (include/quantum_geometric/core/quantum_state.h:15-98)
 * for host-resident amplitude vectors, with QGTL's semantics: FP64
 * accumulation, FP32 amplitudes (src/quantum_geometric/physics/quantum_state_operations.c:313-439).
 * inner_product, norm, clone and initialize_zero are declared but never
 * defined in QGTL; they are defined here with the obvious meaning.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_api.h"

static size_t log2_exact(size_t d)
{
    size_t n = 0;
    while (((size_t)1 << n) < d) n++;
    return ((size_t)1 << n) == d ? n : 0;
}

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
    free(state->coordinates);
    free(state->auxiliary_data);
    free(state);
}

qgt_error_t quantum_state_clone(quantum_state_t **dest, const quantum_state_t *src)
{
    if (!dest || !src || !src->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    qgt_error_t rc = quantum_state_create(dest, src->type, src->dimension);
    if (rc != QGT_SUCCESS) return rc;
    memcpy((*dest)->coordinates, src->coordinates, src->dimension * sizeof(ComplexFloat));
    (*dest)->num_qubits = src->num_qubits;
    (*dest)->manifold_dim = src->manifold_dim;
    (*dest)->is_normalized = src->is_normalized;
    return QGT_SUCCESS;
}

static double norm_sq(const ComplexFloat *a, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += (double)a[i].real * a[i].real + (double)a[i].imag * a[i].imag;
    return s;
}

qgt_error_t quantum_state_initialize(quantum_state_t *state, const ComplexFloat *amplitudes)
{
    if (!state || !state->coordinates || !amplitudes) return QGT_ERROR_INVALID_ARGUMENT;
    memcpy(state->coordinates, amplitudes, state->dimension * sizeof(ComplexFloat));
    const double ns = norm_sq(amplitudes, state->dimension);
    if (ns < 1e-15) return QGT_ERROR_INVALID_STATE;
    if (fabs(ns - 1.0) >= 1e-6) {
        const float inv = 1.0f / sqrtf((float)ns);
        for (size_t i = 0; i < state->dimension; i++) {
            state->coordinates[i].real *= inv;
            state->coordinates[i].imag *= inv;
        }
    }
    state->is_normalized = true;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_initialize_basis(quantum_state_t *state, size_t basis_index)
{
    if (!state || !state->coordinates || basis_index >= state->dimension) return QGT_ERROR_INVALID_ARGUMENT;
    memset(state->coordinates, 0, state->dimension * sizeof(ComplexFloat));
    state->coordinates[basis_index].real = 1.0f;
    state->is_normalized = true;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_initialize_zero(quantum_state_t *state) { return quantum_state_initialize_basis(state, 0); }

qgt_error_t quantum_state_inner_product(ComplexFloat *result, const quantum_state_t *a, const quantum_state_t *b)
{
    if (!result || !a || !b || !a->coordinates || !b->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    if (a->dimension != b->dimension) return QGT_ERROR_INCOMPATIBLE;
    double re = 0.0, im = 0.0;
    for (size_t i = 0; i < a->dimension; i++) {
        const double ar = a->coordinates[i].real, ai = a->coordinates[i].imag;
        const double br = b->coordinates[i].real, bi = b->coordinates[i].imag;
        re += ar * br + ai * bi;
        im += ar * bi - ai * br;
    }
    result->real = (float)re;
    result->imag = (float)im;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_norm(float *norm, const quantum_state_t *state)
{
    if (!norm || !state || !state->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    *norm = (float)sqrt(norm_sq(state->coordinates, state->dimension));
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_normalize(quantum_state_t *state)
{
    if (!state || !state->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    const double ns = norm_sq(state->coordinates, state->dimension);
    if (ns < 1e-15) return QGT_ERROR_INVALID_STATE;
    const float inv = 1.0f / sqrtf((float)ns);
    for (size_t i = 0; i < state->dimension; i++) {
        state->coordinates[i].real *= inv;
        state->coordinates[i].imag *= inv;
    }
    state->is_normalized = true;
    return QGT_SUCCESS;
}

qgt_error_t quantum_state_fidelity(float *fidelity, const quantum_state_t *a, const quantum_state_t *b)
{
    if (!fidelity || !a || !b || !a->coordinates || !b->coordinates) return QGT_ERROR_INVALID_ARGUMENT;
    if (a->dimension != b->dimension) return QGT_ERROR_INCOMPATIBLE;
    /* F = |<a|b>|^2 */
    double re = 0.0, im = 0.0;
    for (size_t i = 0; i < a->dimension; i++) {
        const double ar = a->coordinates[i].real, ai = a->coordinates[i].imag;
        const double br = b->coordinates[i].real, bi = b->coordinates[i].imag;
        re += ar * br + ai * bi;
        im += ar * bi - ai * br;
    }
    *fidelity = (float)(re * re + im * im);
    return QGT_SUCCESS;
}

bool quantum_state_is_normalized(const quantum_state_t *state)
{
    if (!state || !state->coordinates) return false;
    return fabs(norm_sq(state->coordinates, state->dimension) - 1.0) < 1e-6;
}
