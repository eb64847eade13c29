/*
 * Drop-in check of libqgtl_mi355x through QGTL's own C API, in the style of
 * QGTL's tests/test_quantum_geometric_minimal.c: build a network gate by gate
 * with create_quantum_gate / apply_quantum_gate, then call the full-QGT,
 * metric, curvature, element, state and natural-gradient entry points.
 *
 * usage: test_dropin_api n layers seed out.bin
 * Builds the hardware-efficient ansatz (RY, RZ on every qubit, CNOT ring),
 * theta from the splitmix64 stream of the oracle, checks the API's internal
 * consistency (exit code != 0 on failure) and writes, as doubles:
 *   P, Q (P x P complex from geometric_compute_full_qgt), psi (2^n complex
 *   from get_quantum_state), x (P complex from compute_regularized_natural_gradient
 *   on Re Q with gradient g_k = 1 + k/P, lambda = 1e-3, adaptive off).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_api.h"

static double uniform(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

#define CHECK(cond, msg) do { if (!(cond)) { fprintf(stderr, "FAIL: %s\n", msg); return 1; } } while (0)

int main(int argc, char **argv)
{
    if (argc < 5) { fprintf(stderr, "usage: %s n layers seed out.bin\n", argv[0]); return 2; }
    const size_t n = (size_t)atoi(argv[1]), L = (size_t)atoi(argv[2]);
    unsigned long long seed = strtoull(argv[3], NULL, 10);
    const size_t P = 2 * n * L;
    double *theta = malloc(P * sizeof(double));
    for (size_t i = 0; i < P; i++) theta[i] = 2.0 * M_PI * uniform(&seed);

    quantum_geometric_tensor_network_t *qgtn = create_quantum_geometric_tensor_network(n, 1, false, true);
    CHECK(qgtn != NULL, "create network");
    for (size_t l = 0; l < L; l++) {
        for (int kind = 0; kind < 2; kind++)
            for (size_t q = 0; q < n; q++) {
                const double a = theta[2 * n * l + kind * n + q];
                quantum_gate_t *g = create_quantum_gate(kind ? GATE_TYPE_RZ : GATE_TYPE_RY, &q, 1, &a, 1);
                CHECK(g && g->is_parameterized, "create rotation");
                CHECK(apply_quantum_gate(qgtn, g, &q, 1), "apply rotation");
                destroy_quantum_gate(g);
            }
        if (n >= 2)
            for (size_t q = 0; q < n; q++) {
                size_t qs[2] = {q, (q + 1) % n};
                quantum_gate_t *g = create_quantum_gate(GATE_TYPE_CNOT, qs, 2, NULL, 0);
                CHECK(g && g->is_controlled, "create cnot");
                CHECK(apply_quantum_gate(qgtn, g, qs, 2), "apply cnot");
                destroy_quantum_gate(g);
            }
    }

    ComplexFloat *qgt = malloc(P * P * sizeof(ComplexFloat));
    CHECK(geometric_compute_full_qgt(qgt, qgtn, P) == QGT_SUCCESS, "full qgt");
    CHECK(geometric_compute_full_qgt(qgt, qgtn, P + 1) == QGT_ERROR_COMPUTATION_FAILED, "too many params");
    CHECK(geometric_compute_full_qgt(NULL, qgtn, P) == QGT_ERROR_INVALID_PARAMETER, "null output");

    quantum_geometric_metric_t *metric = NULL;
    CHECK(geometric_create_metric(&metric, GEOMETRIC_METRIC_FUBINI_STUDY, P) == QGT_SUCCESS, "create metric");
    CHECK(geometric_compute_fubini_study_metric(metric, qgtn, P) == QGT_SUCCESS, "fs metric");
    quantum_geometric_curvature_t *curv = NULL;
    CHECK(geometric_create_curvature(&curv, GEOMETRIC_CURVATURE_BERRY, P) == QGT_SUCCESS, "create curvature");
    CHECK(geometric_compute_berry_curvature(curv, qgtn, P) == QGT_SUCCESS, "berry");
    ComplexFloat *composed = malloc(P * P * sizeof(ComplexFloat));
    CHECK(geometric_compose_qgt(composed, metric, curv, P) == QGT_SUCCESS, "compose");
    for (size_t i = 0; i < P; i++)
        for (size_t j = 0; j < P; j++) {
            const ComplexFloat z = qgt[i * P + j];
            CHECK(metric->components[i * P + j].real == z.real, "metric = Re Q");
            CHECK(i == j ? curv->components[i * P + j].real == 0.f : curv->components[i * P + j].real == z.imag,
                  "berry = Im Q");
            CHECK(composed[i * P + j].real == z.real && composed[i * P + j].imag == (i == j ? 0.f : z.imag), "compose");
            CHECK(z.real == qgt[j * P + i].real && z.imag == -qgt[j * P + i].imag, "hermitian");
        }
    /* element API agrees with the full matrix */
    for (size_t t = 0; t < 16; t++) {
        const size_t i = (t * 7919) % P, j = (t * 104729 + 3) % P;
        ComplexFloat z;
        double g, w;
        float gf, wf;
        CHECK(compute_quantum_geometric_tensor(qgtn, i, j, &z), "element");
        CHECK(z.real == qgt[i * P + j].real && z.imag == qgt[i * P + j].imag, "element value");
        CHECK(compute_quantum_metric(qgtn, i, j, &g) && g == z.real, "metric element");
        CHECK(compute_berry_curvature(qgtn, i, j, &w) && w == z.imag, "berry element");
        CHECK(geometric_compute_fubini_study_element(qgtn, i, j, &gf) == QGT_SUCCESS && gf == z.real, "fs element");
        CHECK(geometric_compute_berry_curvature_element(qgtn, i, j, &wf) == QGT_SUCCESS && wf == z.imag, "berry el.");
    }
    ComplexFloat z;
    CHECK(!compute_quantum_geometric_tensor(qgtn, P, 0, &z), "out-of-range element fails");
    CHECK(strlen(get_quantum_geometric_tensor_network_error()) > 0, "error message set");

    ComplexFloat *psi = NULL;
    size_t dim = 0;
    CHECK(get_quantum_state(qgtn, &psi, &dim) && dim == ((size_t)1 << n), "state");

    ComplexFloat *grad = malloc(P * sizeof(ComplexFloat)), *x = malloc(P * sizeof(ComplexFloat));
    for (size_t k = 0; k < P; k++) grad[k] = (ComplexFloat){1.0f + (float)k / (float)P, 0.f};
    natural_gradient_config_t cfg = get_default_natural_gradient_config();
    CHECK(cfg.regularization_param == 1e-4f && cfg.use_adaptive_regularization, "default config");
    cfg.regularization_param = 1e-3f;
    cfg.use_adaptive_regularization = false;
    CHECK(compute_regularized_natural_gradient(grad, metric->components, x, P, &cfg), "natural gradient");

    FILE *f = fopen(argv[4], "wb");
    CHECK(f != NULL, "open output");
    const double Pd = (double)P;
    fwrite(&Pd, sizeof Pd, 1, f);
    for (size_t i = 0; i < P * P; i++) { double v[2] = {qgt[i].real, qgt[i].imag}; fwrite(v, sizeof v, 1, f); }
    for (size_t i = 0; i < dim; i++) { double v[2] = {psi[i].real, psi[i].imag}; fwrite(v, sizeof v, 1, f); }
    for (size_t i = 0; i < P; i++) { double v[2] = {x[i].real, x[i].imag}; fwrite(v, sizeof v, 1, f); }
    fclose(f);

    free(psi); free(grad); free(x); free(qgt); free(composed); free(theta);
    geometric_destroy_metric(metric);
    geometric_destroy_curvature(curv);
    destroy_quantum_geometric_tensor_network(qgtn);
    printf("ok P=%zu\n", P);
    return 0;
}
