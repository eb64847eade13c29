/*
 * More drop-in checks of libqgtl_mi355x through QGTL's C API (the program a
 * QGTL user would write), one scenario per sub-command.  Every scenario
 * checks what it can by itself (exit code != 0 on failure) and writes its
 * inputs and outputs as doubles for tests/test_dropin_more_gpu.py, which
 * compares them with the FP64 oracle or numpy.
 *
 *   flat OUT      a flat circuit built the way QGTL's circuit builders build
 *                 one (quantum_circuit_operations.c:786-1135: malloc'd gates,
 *                 only type / num_qubits / qubits / parameters set), applied
 *                 with apply_quantum_circuit; writes P, Q (complex128), psi
 *   threads N L SEED_A SEED_B REPS OUT
 *                 two host threads, each building its own network of the
 *                 same structure at its own angles and computing the full QGT
 *                 REPS times; every repetition must be bit-identical to the
 *                 thread's first; writes P, Q_A, Q_B
 *   state OUT     the quantum_state_* set, on the host and on the device
 *   tensor DIR    geometric_tensor_save / geometric_tensor_load round trip
 *   backend OUT   the HIP ComputeBackendOps slots
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_api.h"
#include "qgtl_compute_backend.h"
#include "qgtl_mi355x.h"

#define CHECK(cond, msg)                                                        \
    do {                                                                        \
        if (!(cond)) {                                                          \
            fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, msg);       \
            return 1;                                                           \
        }                                                                       \
    } while (0)

static double uniform(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void put(FILE *f, double v) { fwrite(&v, sizeof v, 1, f); }

/* ------------------------------------------------------------ flat */
/* A gate as QGTL's builders make it: every field the builder leaves alone
 * holds garbage, here a 0xA5 fill. */
static quantum_gate_t *builder_gate(gate_type_t type, size_t q0, long q1, const double *angle)
{
    quantum_gate_t *g = (quantum_gate_t *)malloc(sizeof *g);
    memset(g, 0xA5, sizeof *g);
    g->type = type;
    g->num_qubits = q1 >= 0 ? 2 : 1;
    g->qubits = (size_t *)malloc(g->num_qubits * sizeof(size_t));
    g->qubits[0] = q0;
    if (q1 >= 0) g->qubits[1] = (size_t)q1;
    if (angle) {
        g->parameters = (double *)malloc(sizeof(double));
        g->parameters[0] = *angle;
        g->num_parameters = 1;
    } else {
        g->parameters = NULL;
        g->num_parameters = 0;
    }
    g->custom_data = NULL;
    return g;
}

static void free_builder_gate(quantum_gate_t *g)
{
    free(g->qubits);
    free(g->parameters);
    free(g);
}

static int run_flat(const char *out)
{
    /* the gate list tests/test_dropin_more_gpu.py rebuilds for the oracle */
    const size_t n = 5;
    const double a[8] = {0.3, 1.1, -0.7, 2.9, 0.45, -1.8, 0.05, 3.7};
    quantum_circuit_t c;
    memset(&c, 0, sizeof c);
    c.num_qubits = n;
    quantum_gate_t *gates[32];
    size_t k = 0;
    gates[k++] = builder_gate(GATE_TYPE_H, 0, -1, NULL);
    gates[k++] = builder_gate(GATE_TYPE_RY, 1, -1, &a[0]);
    gates[k++] = builder_gate(GATE_TYPE_CNOT, 0, 1, NULL);
    gates[k++] = builder_gate(GATE_TYPE_RX, 2, -1, &a[1]);
    gates[k++] = builder_gate(GATE_TYPE_X, 3, -1, NULL);
    gates[k++] = builder_gate(GATE_TYPE_RZ, 3, -1, &a[2]);
    gates[k++] = builder_gate(GATE_TYPE_CZ, 2, 4, NULL);
    gates[k++] = builder_gate(GATE_TYPE_RY, 4, -1, &a[3]);
    gates[k++] = builder_gate(GATE_TYPE_S, 0, -1, &a[4]);     /* quantum_circuit_phase: an S, angle unused */
    gates[k++] = builder_gate(GATE_TYPE_SWAP, 1, 3, NULL);
    gates[k++] = builder_gate(GATE_TYPE_Y, 2, -1, NULL);
    gates[k++] = builder_gate(GATE_TYPE_RX, 0, -1, &a[5]);
    gates[k++] = builder_gate(GATE_TYPE_CNOT, 4, 0, NULL);
    gates[k++] = builder_gate(GATE_TYPE_Z, 1, -1, NULL);
    gates[k++] = builder_gate(GATE_TYPE_RZ, 4, -1, &a[6]);
    gates[k++] = builder_gate(GATE_TYPE_RY, 2, -1, &a[7]);
    c.gates = gates;
    c.num_gates = k;
    c.max_gates = 32;

    quantum_geometric_tensor_network_t *q = create_quantum_geometric_tensor_network(n, 1, false, true);
    CHECK(q != NULL, "create network");
    CHECK(apply_quantum_circuit(q, &c), get_quantum_geometric_tensor_network_error());
    int P = 0;
    CHECK(qgtl_network_qgt_f64(q, NULL, &P) == 0 && P == 7, "seven parameterised gates");
    double *Q = (double *)malloc(sizeof(double) * 2 * P * P);
    CHECK(qgtl_network_qgt_f64(q, Q, &P) == 0, "network QGT");
    ComplexFloat *psi = NULL;
    size_t dim = 0;
    CHECK(get_quantum_state(q, &psi, &dim) && dim == 32, "state");
    /* a builder gate with an out-of-range qubit is refused */
    quantum_gate_t *bad = builder_gate(GATE_TYPE_RY, 7, -1, &a[0]);
    quantum_circuit_t cb = c;
    cb.gates = &bad;
    cb.num_gates = 1;
    quantum_geometric_tensor_network_t *q2 = create_quantum_geometric_tensor_network(n, 1, false, true);
    CHECK(!apply_quantum_circuit(q2, &cb), "out-of-range qubit refused");
    destroy_quantum_geometric_tensor_network(q2);
    free_builder_gate(bad);

    FILE *f = fopen(out, "wb");
    CHECK(f != NULL, "open output");
    put(f, P);
    for (int i = 0; i < 2 * P * P; i++) put(f, Q[i]);
    for (size_t i = 0; i < dim; i++) { put(f, psi[i].real); put(f, psi[i].imag); }
    fclose(f);
    free(psi);
    free(Q);
    for (size_t i = 0; i < k; i++) free_builder_gate(gates[i]);
    destroy_quantum_geometric_tensor_network(q);
    printf("ok flat P=%d\n", P);
    return 0;
}

/* ---------------------------------------------------------- threads */
typedef struct {
    size_t n, L;
    unsigned long long seed;
    int reps;
    ComplexFloat *first;
    int failed;
    char msg[256];
} worker_t;

static quantum_geometric_tensor_network_t *hea_network(size_t n, size_t L, const double *theta)
{
    quantum_geometric_tensor_network_t *q = create_quantum_geometric_tensor_network(n, 1, false, true);
    if (!q) return NULL;
    for (size_t l = 0; l < L; l++) {
        for (int kind = 0; kind < 2; kind++)
            for (size_t qb = 0; qb < n; qb++) {
                const double a = theta[2 * n * l + kind * n + qb];
                quantum_gate_t *g = create_quantum_gate(kind ? GATE_TYPE_RZ : GATE_TYPE_RY, &qb, 1, &a, 1);
                const bool ok = g && apply_quantum_gate(q, g, &qb, 1);
                destroy_quantum_gate(g);
                if (!ok) { destroy_quantum_geometric_tensor_network(q); return NULL; }
            }
        for (size_t qb = 0; n >= 2 && qb < n; qb++) {
            size_t qs[2] = {qb, (qb + 1) % n};
            quantum_gate_t *g = create_quantum_gate(GATE_TYPE_CNOT, qs, 2, NULL, 0);
            const bool ok = g && apply_quantum_gate(q, g, qs, 2);
            destroy_quantum_gate(g);
            if (!ok) { destroy_quantum_geometric_tensor_network(q); return NULL; }
        }
    }
    return q;
}

static void *worker(void *arg)
{
    worker_t *w = (worker_t *)arg;
    const size_t P = 2 * w->n * w->L;
    double *theta = (double *)malloc(P * sizeof(double));
    unsigned long long s = w->seed;
    for (size_t i = 0; i < P; i++) theta[i] = 2.0 * M_PI * uniform(&s);
    ComplexFloat *Q = (ComplexFloat *)malloc(P * P * sizeof(ComplexFloat));
    w->first = (ComplexFloat *)malloc(P * P * sizeof(ComplexFloat));
    for (int r = 0; r < w->reps && !w->failed; r++) {
        /* a fresh network every repetition: each one is computed on the GPU */
        quantum_geometric_tensor_network_t *q = hea_network(w->n, w->L, theta);
        if (!q) { w->failed = 1; snprintf(w->msg, sizeof w->msg, "network"); break; }
        if (geometric_compute_full_qgt(Q, q, P) != QGT_SUCCESS) {
            w->failed = 1;
            snprintf(w->msg, sizeof w->msg, "full qgt: %s", qgtl_last_error());
        } else if (r == 0) {
            memcpy(w->first, Q, P * P * sizeof(ComplexFloat));
        } else if (memcmp(w->first, Q, P * P * sizeof(ComplexFloat)) != 0) {
            w->failed = 1;
            snprintf(w->msg, sizeof w->msg, "repetition %d differs from the first", r);
        }
        destroy_quantum_geometric_tensor_network(q);
    }
    free(Q);
    free(theta);
    return NULL;
}

static int run_threads(size_t n, size_t L, unsigned long long sa, unsigned long long sb, int reps, const char *out)
{
    worker_t w[2] = {{n, L, sa, reps, NULL, 0, ""}, {n, L, sb, reps, NULL, 0, ""}};
    pthread_t t[2];
    for (int i = 0; i < 2; i++) CHECK(pthread_create(&t[i], NULL, worker, &w[i]) == 0, "pthread_create");
    for (int i = 0; i < 2; i++) pthread_join(t[i], NULL);
    for (int i = 0; i < 2; i++)
        if (w[i].failed) { fprintf(stderr, "FAIL thread %d: %s\n", i, w[i].msg); return 1; }
    const size_t P = 2 * n * L;
    FILE *f = fopen(out, "wb");
    CHECK(f != NULL, "open output");
    put(f, (double)P);
    for (int i = 0; i < 2; i++)
        for (size_t e = 0; e < P * P; e++) { put(f, w[i].first[e].real); put(f, w[i].first[e].imag); }
    fclose(f);
    free(w[0].first);
    free(w[1].first);
    printf("ok threads P=%zu reps=%d\n", P, reps);
    return 0;
}

/* ------------------------------------------------------------ state */
static int run_state(const char *out)
{
    const size_t n = 12, N = (size_t)1 << n;
    unsigned long long s = 77;
    ComplexFloat *amp = (ComplexFloat *)malloc(N * sizeof(ComplexFloat));
    ComplexFloat *amp2 = (ComplexFloat *)malloc(N * sizeof(ComplexFloat));
    for (size_t i = 0; i < N; i++) {
        amp[i] = (ComplexFloat){(float)(uniform(&s) - 0.5), (float)(uniform(&s) - 0.5)};
        amp2[i] = (ComplexFloat){(float)(uniform(&s) - 0.5), (float)(uniform(&s) - 0.5)};
    }
    quantum_hardware_t gpu, cpu;
    memset(&gpu, 0, sizeof gpu);
    memset(&cpu, 0, sizeof cpu);
    gpu.type = HARDWARE_TYPE_GPU;
    cpu.type = HARDWARE_TYPE_CPU;

    quantum_state_t *a = NULL, *b = NULL, *c = NULL;
    CHECK(quantum_state_create(&a, QUANTUM_STATE_PURE, N) == QGT_SUCCESS && a->num_qubits == n, "create a");
    CHECK(quantum_state_create(&b, QUANTUM_STATE_PURE, N) == QGT_SUCCESS, "create b");
    CHECK(a->coordinates[0].real == 1.f && quantum_state_is_normalized(a), "create gives |0>");
    CHECK(quantum_state_initialize(a, amp) == QGT_SUCCESS && quantum_state_is_normalized(a), "initialize a");
    CHECK(quantum_state_initialize(b, amp2) == QGT_SUCCESS, "initialize b");
    /* host values */
    ComplexFloat ip_h;
    float fid_h, norm_h;
    CHECK(quantum_state_inner_product(&ip_h, a, b) == QGT_SUCCESS, "host inner product");
    CHECK(quantum_state_fidelity(&fid_h, a, b) == QGT_SUCCESS, "host fidelity");
    CHECK(quantum_state_norm(&norm_h, a) == QGT_SUCCESS, "host norm");
    /* the same on the device */
    CHECK(!quantum_state_is_on_device(a, gpu), "not on device yet");
    CHECK(quantum_state_to_device(a, gpu) == QGT_SUCCESS, "to device");
    CHECK(quantum_state_is_on_device(a, gpu) && a->hardware == HARDWARE_TYPE_GPU, "on device");
    CHECK(!quantum_state_is_on_device(a, cpu), "not on the CPU kind");
    ComplexFloat ip_d, ip_dd;
    float fid_d, norm_d;
    CHECK(quantum_state_inner_product(&ip_d, a, b) == QGT_SUCCESS, "device inner product (b on host)");
    CHECK(quantum_state_to_device(b, gpu) == QGT_SUCCESS, "b to device");
    CHECK(quantum_state_inner_product(&ip_dd, a, b) == QGT_SUCCESS, "device inner product (both)");
    CHECK(quantum_state_fidelity(&fid_d, a, b) == QGT_SUCCESS, "device fidelity");
    CHECK(quantum_state_norm(&norm_d, a) == QGT_SUCCESS, "device norm");
    CHECK(quantum_state_is_normalized(a), "device is_normalized");
    /* unnormalised values written straight into the device copy, then normalize there */
    void *dev = qgtl_state_device_ptr(a);
    CHECK(dev != NULL, "device pointer");
    const ComputeBackendOps *ops = qgtl_hip_compute_backend_ops();
    ComputeBackend *be = ops->init(NULL);
    CHECK(be != NULL, "backend init");
    CHECK(ops->memcpy(be, dev, COMPUTE_MEM_DEVICE, amp2, COMPUTE_MEM_HOST, N * sizeof(ComplexFloat), NULL) ==
              COMPUTE_SUCCESS, "write device copy");
    CHECK(!quantum_state_is_normalized(a), "device copy not normalised");
    CHECK(quantum_state_normalize(a) == QGT_SUCCESS, "device normalize");
    CHECK(quantum_state_is_normalized(a), "normalised on device");
    /* clone of a device state takes the device values */
    CHECK(quantum_state_clone(&c, a) == QGT_SUCCESS && !quantum_state_is_on_device(c, gpu), "clone");
    CHECK(quantum_state_from_device(a, gpu) == QGT_SUCCESS && !quantum_state_is_on_device(a, gpu) &&
              a->hardware == HARDWARE_TYPE_CPU, "from device");
    CHECK(memcmp(a->coordinates, c->coordinates, N * sizeof(ComplexFloat)) == 0, "clone = device values");
    CHECK(quantum_state_from_device(a, gpu) == QGT_SUCCESS, "from device twice is harmless");
    /* initialize_basis on a device state writes both copies */
    CHECK(quantum_state_initialize_basis(b, 5) == QGT_SUCCESS, "basis on device");
    ComplexFloat ipb;
    CHECK(quantum_state_inner_product(&ipb, c, b) == QGT_SUCCESS, "inner product with basis state");
    CHECK(quantum_state_to_device(b, cpu) == QGT_SUCCESS && !quantum_state_is_on_device(b, gpu), "to CPU kind");
    CHECK(b->coordinates[5].real == 1.f, "basis values back");
    quantum_hardware_t qpu = cpu;
    qpu.type = HARDWARE_TYPE_QPU;
    CHECK(quantum_state_to_device(b, qpu) == QGT_ERROR_NOT_IMPLEMENTED, "other hardware refused");
    ops->cleanup(be);

    FILE *f = fopen(out, "wb");
    CHECK(f != NULL, "open output");
    put(f, (double)N);
    for (size_t i = 0; i < N; i++) { put(f, amp[i].real); put(f, amp[i].imag); }
    for (size_t i = 0; i < N; i++) { put(f, amp2[i].real); put(f, amp2[i].imag); }
    put(f, ip_h.real); put(f, ip_h.imag); put(f, fid_h); put(f, norm_h);
    put(f, ip_d.real); put(f, ip_d.imag); put(f, ip_dd.real); put(f, ip_dd.imag); put(f, fid_d); put(f, norm_d);
    for (size_t i = 0; i < N; i++) { put(f, a->coordinates[i].real); put(f, a->coordinates[i].imag); }
    put(f, ipb.real); put(f, ipb.imag);
    fclose(f);
    quantum_state_destroy(a);
    quantum_state_destroy(b);
    quantum_state_destroy(c);
    free(amp);
    free(amp2);
    printf("ok state\n");
    return 0;
}

/* ----------------------------------------------------------- tensor */
/* A tensor allocated the way QGTL's geometric_tensor_create allocates one
 * (quantum_geometric_tensor.c:158-240). */
static quantum_geometric_tensor_t *qgtl_style_tensor(size_t P)
{
    quantum_geometric_tensor_t *t = (quantum_geometric_tensor_t *)aligned_alloc(64, (sizeof *t + 63) / 64 * 64);
    memset(t, 0, sizeof *t);
    t->dimensions = (size_t *)malloc(2 * sizeof(size_t));
    t->dimensions[0] = t->dimensions[1] = P;
    t->rank = 2;
    t->total_elements = P * P;
    t->aligned_elements = (P * P + 7) & ~(size_t)7;
    t->components = (ComplexFloat *)aligned_alloc(64, (t->aligned_elements * sizeof(ComplexFloat) + 63) / 64 * 64);
    t->type = GEOMETRIC_TENSOR_HERMITIAN;
    t->is_hermitian = true;
    t->hardware = HARDWARE_TYPE_SIMULATOR;
    t->dimension = P;
    return t;
}

static void destroy_tensor(quantum_geometric_tensor_t *t)
{
    free(t->dimensions);
    free(t->components);
    free(t->auxiliary_data);
    free(t);
}

static int run_tensor(const char *dir)
{
    const size_t n = 6, L = 3, P = 2 * n * L;
    double theta[36];
    unsigned long long s = 5;
    for (size_t i = 0; i < P; i++) theta[i] = 2.0 * M_PI * uniform(&s);
    quantum_geometric_tensor_network_t *q = hea_network(n, L, theta);
    CHECK(q != NULL, "network");
    quantum_geometric_tensor_t *t = qgtl_style_tensor(P);
    CHECK(geometric_compute_full_qgt(t->components, q, P) == QGT_SUCCESS, "full qgt");
    char path[1024], bad[1024];
    snprintf(path, sizeof path, "%s/qgt.tensor", dir);
    snprintf(bad, sizeof bad, "%s/bad.tensor", dir);
    CHECK(geometric_tensor_save(t, path) == QGT_SUCCESS, "save");
    quantum_geometric_tensor_t *u = NULL;
    CHECK(geometric_tensor_load(&u, path) == QGT_SUCCESS && u, "load");
    CHECK(u->rank == 2 && u->dimensions[0] == P && u->dimensions[1] == P && u->total_elements == P * P &&
              u->aligned_elements == t->aligned_elements && u->type == GEOMETRIC_TENSOR_HERMITIAN &&
              u->is_hermitian && !u->is_unitary && !u->is_symmetric && u->dimension == P &&
              u->hardware == HARDWARE_TYPE_SIMULATOR && u->auxiliary_data == NULL,
          "loaded header fields");
    CHECK(memcmp(u->components, t->components, P * P * sizeof(ComplexFloat)) == 0, "loaded components");
    CHECK(((uintptr_t)u->components & 63) == 0 && ((uintptr_t)u & 63) == 0, "64-byte aligned like QGTL's create");
    /* a flipped byte in the payload fails the checksum */
    FILE *fi = fopen(path, "rb"), *fo = fopen(bad, "wb");
    CHECK(fi && fo, "copy file");
    int ch;
    long pos = 0;
    while ((ch = fgetc(fi)) != EOF) { fputc(pos == 200 ? (ch ^ 0x10) : ch, fo); pos++; }
    fclose(fi);
    fclose(fo);
    quantum_geometric_tensor_t *v = NULL;
    CHECK(geometric_tensor_load(&v, bad) == QGT_ERROR_VALIDATION_FAILED && v == NULL, "corrupt file refused");
    CHECK(geometric_tensor_load(&v, "/nonexistent/x.tensor") == QGT_ERROR_IO_ERROR, "missing file");
    t->total_elements = 3;
    CHECK(geometric_tensor_save(t, bad) == QGT_ERROR_INVALID_STATE, "inconsistent tensor refused");
    destroy_tensor(t);
    destroy_tensor(u);
    destroy_quantum_geometric_tensor_network(q);
    printf("ok tensor P=%zu\n", P);
    return 0;
}

/* ---------------------------------------------------------- backend */
static int run_backend(const char *out)
{
    const ComputeBackendOps *ops = qgtl_hip_compute_backend_ops();
    const ComputeBackendInfo *info = qgtl_hip_compute_backend_info();
    CHECK(ops && info && info->ops == ops && info->type == COMPUTE_BACKEND_CUDA && info->priority == 100,
          "registration record");
    CHECK(ops->probe(), "probe");
    ComputeDistributedConfig two;
    memset(&two, 0, sizeof two);
    two.num_nodes = 2;
    CHECK(ops->init(&two) == NULL, "more than one node refused");
    ComputeBackend *be = ops->init(NULL);
    CHECK(be != NULL, "init");
    int ndev = 0;
    size_t mem = 0;
    CHECK(ops->get_capabilities(be, &ndev, &mem) == COMPUTE_SUCCESS && ndev == 1 && mem > ((size_t)1 << 30),
          "capabilities");

    const size_t N = 256, m = 5, nn = 7, kk = 3;
    unsigned long long s = 99;
    float *psi = (float *)malloc(2 * N * sizeof(float)), *phi = (float *)malloc(2 * N * sizeof(float));
    float *U = (float *)malloc(2 * N * N * sizeof(float)), *obs = (float *)malloc(N * sizeof(float));
    float *A = (float *)malloc(2 * m * nn * sizeof(float)), *B = (float *)malloc(2 * nn * kk * sizeof(float));
    for (size_t i = 0; i < 2 * N; i++) { psi[i] = (float)(uniform(&s) - 0.5); phi[i] = (float)(uniform(&s) - 0.5); }
    for (size_t i = 0; i < 2 * N * N; i++) U[i] = (float)(uniform(&s) - 0.5) * 0.1f;
    for (size_t i = 0; i < N; i++) obs[i] = (float)(uniform(&s) * 2.0 - 1.0);
    for (size_t i = 0; i < 2 * m * nn; i++) A[i] = (float)(uniform(&s) - 0.5);
    for (size_t i = 0; i < 2 * nn * kk; i++) B[i] = (float)(uniform(&s) - 0.5);

    FILE *f = fopen(out, "wb");
    CHECK(f != NULL, "open output");
    put(f, (double)N);
    for (size_t i = 0; i < 2 * N; i++) put(f, psi[i]);
    for (size_t i = 0; i < 2 * N; i++) put(f, phi[i]);
    for (size_t i = 0; i < 2 * N * N; i++) put(f, U[i]);
    for (size_t i = 0; i < N; i++) put(f, obs[i]);
    for (size_t i = 0; i < 2 * m * nn; i++) put(f, A[i]);
    for (size_t i = 0; i < 2 * nn * kk; i++) put(f, B[i]);

    /* host operands */
    float r[2], g[2], e;
    CHECK(ops->quantum_inner_product(be, r, psi, phi, N, NULL) == COMPUTE_SUCCESS, "inner product");
    CHECK(ops->quantum_gradient(be, g, psi, phi, N, NULL) == COMPUTE_SUCCESS, "gradient");
    CHECK(ops->quantum_expectation(be, &e, psi, obs, N, NULL) == COMPUTE_SUCCESS, "expectation");
    float *Upsi = (float *)malloc(2 * N * sizeof(float));
    memcpy(Upsi, psi, 2 * N * sizeof(float));
    CHECK(ops->quantum_unitary(be, Upsi, N, U, N, NULL) == COMPUTE_SUCCESS, "unitary");
    CHECK(ops->quantum_unitary(be, Upsi, N, U, N - 1, NULL) == COMPUTE_ERROR_INVALID_ARGUMENT, "unitary size");
    float *nrm = (float *)malloc(2 * N * sizeof(float));
    memcpy(nrm, phi, 2 * N * sizeof(float));
    CHECK(ops->quantum_normalize(be, nrm, N, NULL) == COMPUTE_SUCCESS, "normalize");
    float C[2 * 5 * 3];
    CHECK(ops->quantum_tensor_contract(be, C, A, B, m, nn, kk, NULL) == COMPUTE_SUCCESS, "contract");
    for (int i = 0; i < 2; i++) put(f, r[i]);
    for (int i = 0; i < 2; i++) put(f, g[i]);
    put(f, e);
    for (size_t i = 0; i < 2 * N; i++) put(f, Upsi[i]);
    for (size_t i = 0; i < 2 * N; i++) put(f, nrm[i]);
    for (size_t i = 0; i < 2 * m * kk; i++) put(f, C[i]);

    /* device operands through alloc / memcpy, and execute() dispatching the same slots */
    float *dpsi = (float *)ops->alloc(be, 2 * N * sizeof(float), COMPUTE_MEM_DEVICE);
    float *dU = (float *)ops->alloc(be, 2 * N * N * sizeof(float), COMPUTE_MEM_DEVICE);
    float *pin = (float *)ops->alloc(be, 2 * N * sizeof(float), COMPUTE_MEM_PINNED);
    CHECK(dpsi && dU && pin, "alloc");
    ComputeStream *st = ops->create_stream(be);
    ComputeEvent *ev = ops->create_event(be);
    CHECK(st && ev, "stream / event");
    CHECK(ops->memcpy(be, dpsi, COMPUTE_MEM_DEVICE, psi, COMPUTE_MEM_HOST, 2 * N * sizeof(float), st) ==
              COMPUTE_SUCCESS, "h2d");
    CHECK(ops->memcpy(be, dU, COMPUTE_MEM_DEVICE, U, COMPUTE_MEM_HOST, 2 * N * N * sizeof(float), st) ==
              COMPUTE_SUCCESS, "h2d U");
    CHECK(ops->record_event(be, ev, st) == COMPUTE_SUCCESS && ops->wait_event(be, st, ev) == COMPUTE_SUCCESS,
          "event");
    ComputeQuantumOp op;
    memset(&op, 0, sizeof op);
    op.type = QUANTUM_OP_UNITARY;
    op.output_data = dpsi;
    op.output_size = N;
    op.parameters = dU;
    op.param_size = N;
    ComputeExecutionPlan *plan = ops->create_plan(be, &op);
    CHECK(plan && plan->num_partitions == 1, "plan");
    CHECK(ops->execute(be, &op, plan, st) == COMPUTE_SUCCESS, "execute unitary on device operands");
    ops->destroy_plan(be, plan);
    CHECK(ops->synchronize_stream(be, st) == COMPUTE_SUCCESS, "sync");
    CHECK(ops->memcpy(be, pin, COMPUTE_MEM_PINNED, dpsi, COMPUTE_MEM_DEVICE, 2 * N * sizeof(float), NULL) ==
              COMPUTE_SUCCESS, "d2h");
    CHECK(memcmp(pin, Upsi, 2 * N * sizeof(float)) == 0, "device-operand unitary = host-operand unitary");
    float rd[2];
    op.type = QUANTUM_OP_INNER_PRODUCT;
    op.input_data = dpsi;
    op.input_size = N;
    op.parameters = dpsi;
    op.output_data = rd;
    CHECK(ops->execute(be, &op, NULL, NULL) == COMPUTE_SUCCESS, "execute inner product");
    op.type = QUANTUM_OP_MEASUREMENT;
    CHECK(ops->execute(be, &op, NULL, NULL) == COMPUTE_ERROR_NOT_IMPLEMENTED, "unsupported op");
    put(f, rd[0]);
    put(f, rd[1]);
    /* single-node collectives copy */
    double sd[3] = {1.5, -2.0, 7.25}, rv[3] = {0, 0, 0};
    CHECK(ops->allreduce(be, sd, rv, 3, COMPUTE_DTYPE_FLOAT64, COMPUTE_REDUCE_SUM) == COMPUTE_SUCCESS &&
              memcmp(sd, rv, sizeof sd) == 0, "allreduce");
    CHECK(ops->memset(be, dpsi, 0, 2 * N * sizeof(float), NULL) == COMPUTE_SUCCESS, "memset");
    CHECK(ops->memcpy(be, pin, COMPUTE_MEM_PINNED, dpsi, COMPUTE_MEM_DEVICE, 2 * N * sizeof(float), NULL) ==
              COMPUTE_SUCCESS && pin[0] == 0.f && pin[2 * N - 1] == 0.f, "memset values");
    CHECK(ops->barrier(be) == COMPUTE_SUCCESS, "barrier");
    ComputeMetrics met;
    CHECK(ops->get_metrics(be, &met) == COMPUTE_SUCCESS && met.flops > 0 && met.current_memory_bytes > 0, "metrics");
    CHECK(ops->reset_metrics(be) == COMPUTE_SUCCESS, "reset metrics");
    ops->destroy_event(be, ev);
    ops->destroy_stream(be, st);
    ops->free(be, dpsi, COMPUTE_MEM_DEVICE);
    ops->free(be, dU, COMPUTE_MEM_DEVICE);
    ops->free(be, pin, COMPUTE_MEM_PINNED);
    ops->cleanup(be);
    fclose(f);
    free(psi); free(phi); free(U); free(obs); free(A); free(B); free(Upsi); free(nrm);
    printf("ok backend\n");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 3 && strcmp(argv[1], "flat") == 0) return run_flat(argv[2]);
    if (argc >= 8 && strcmp(argv[1], "threads") == 0)
        return run_threads((size_t)atoi(argv[2]), (size_t)atoi(argv[3]), strtoull(argv[4], NULL, 10),
                           strtoull(argv[5], NULL, 10), atoi(argv[6]), argv[7]);
    if (argc >= 3 && strcmp(argv[1], "state") == 0) return run_state(argv[2]);
    if (argc >= 3 && strcmp(argv[1], "tensor") == 0) return run_tensor(argv[2]);
    if (argc >= 3 && strcmp(argv[1], "backend") == 0) return run_backend(argv[2]);
    fprintf(stderr, "usage: %s flat|threads|state|tensor|backend ...\n", argv[0]);
    return 2;
}
