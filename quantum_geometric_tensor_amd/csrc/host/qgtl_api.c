/*
 * qgtl_api.c -- QGTL's network, gate, metric, curvature and full-QGT API on
 * top of the GPU engine (qgtl_engine.c).
 *
 * A network's circuit is read as QGTL's find_parameterized_gate reads it
 * (src/quantum_geometric/core/quantum_parameter_shift.c:278-337): layers in
 * order, gates in order, each is_parameterized gate carrying the next
 * parameter index (angle = parameters[0]).  The full P x P QGT is computed
 * once per circuit content on the GPU and kept on the network
 * (backend_state); element calls read it.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qgtl_api.h"
#include "qgtl_internal.h"

/* ------------------------------------------------------------------ errors */
static _Thread_local char g_qgtn_err[256] = "";

static void set_qgtn_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_qgtn_err, sizeof g_qgtn_err, fmt, ap);
    va_end(ap);
}

const char *get_quantum_geometric_tensor_network_error(void) { return g_qgtn_err; }

/* ------------------------------------------------------------------ gates */
static void fill_gate_matrix(quantum_gate_t *g)
{
    const size_t dim = (size_t)1 << g->num_qubits;
    memset(g->matrix, 0, dim * dim * sizeof(ComplexFloat));
    qgtl_mat2 m;
    if (g->num_qubits == 1) {
        qgtl_op_t op = {g->type, 0, -1, -1, g->parameters ? g->parameters[0] : 0.0};
        if (qgtl_gate_matrix(&op, &m) == 0)
            for (int e = 0; e < 4; e++) g->matrix[e] = (ComplexFloat){(float)m.m[e].re, (float)m.m[e].im};
        else
            for (int e = 0; e < 2; e++) g->matrix[3 * e] = (ComplexFloat){1.f, 0.f};
        return;
    }
    if (g->num_qubits == 2 && (g->type == GATE_TYPE_CNOT || g->type == GATE_TYPE_CZ || g->type == GATE_TYPE_CY)) {
        /* |0><0| (x) I + |1><1| (x) U, control = high bit of the 2-qubit index
         * (quantum_gate_operations.c:130-176) */
        qgtl_op_t op = {g->type == GATE_TYPE_CNOT ? GATE_TYPE_X : g->type == GATE_TYPE_CZ ? GATE_TYPE_Z : GATE_TYPE_Y,
                        0, -1, -1, 0.0};
        qgtl_gate_matrix(&op, &m);
        g->matrix[0] = (ComplexFloat){1.f, 0.f};
        g->matrix[5] = (ComplexFloat){1.f, 0.f};
        for (int r = 0; r < 2; r++)
            for (int c = 0; c < 2; c++)
                g->matrix[(2 + r) * 4 + 2 + c] = (ComplexFloat){(float)m.m[2 * r + c].re, (float)m.m[2 * r + c].im};
        return;
    }
    if (g->num_qubits == 2 && g->type == GATE_TYPE_SWAP) {
        g->matrix[0] = g->matrix[6] = g->matrix[9] = g->matrix[15] = (ComplexFloat){1.f, 0.f};
        return;
    }
    for (size_t i = 0; i < dim; i++) g->matrix[i * dim + i] = (ComplexFloat){1.f, 0.f};
}

quantum_gate_t *create_quantum_gate(gate_type_t type, const size_t *qubits, size_t num_qubits,
                                    const double *parameters, size_t num_parameters)
{
    if (!qubits || num_qubits == 0 || num_qubits > 16 || (parameters == NULL && num_parameters > 0)) return NULL;
    const int rot = (type == GATE_TYPE_RX || type == GATE_TYPE_RY || type == GATE_TYPE_RZ);
    if (rot && num_parameters != 1) return NULL;
    if ((type == GATE_TYPE_CNOT || type == GATE_TYPE_CZ || type == GATE_TYPE_SWAP) && num_qubits != 2) return NULL;
    quantum_gate_t *g = (quantum_gate_t *)calloc(1, sizeof *g);
    if (!g) return NULL;
    g->type = type;
    g->num_qubits = num_qubits;
    g->target_qubits = (size_t *)malloc(num_qubits * sizeof(size_t));
    /* the unified array of QGTL's circuit interface (quantum_types.h:181) holds the same qubits */
    g->qubits = (size_t *)malloc(num_qubits * sizeof(size_t));
    const size_t dim = (size_t)1 << num_qubits;
    g->matrix = (ComplexFloat *)malloc(dim * dim * sizeof(ComplexFloat));
    if (!g->target_qubits || !g->qubits || !g->matrix) { destroy_quantum_gate(g); return NULL; }
    memcpy(g->target_qubits, qubits, num_qubits * sizeof(size_t));
    memcpy(g->qubits, qubits, num_qubits * sizeof(size_t));
    g->is_controlled = (type == GATE_TYPE_CNOT || type == GATE_TYPE_CZ);
    if (rot) {
        g->is_parameterized = true;
        g->parameters = (double *)malloc(sizeof(double));
        if (!g->parameters) { destroy_quantum_gate(g); return NULL; }
        g->parameters[0] = parameters[0];
        g->num_parameters = 1;
    }
    fill_gate_matrix(g);
    return g;
}

quantum_gate_t *copy_quantum_gate(const quantum_gate_t *gate)
{
    if (!gate) return NULL;
    quantum_gate_t *g = (quantum_gate_t *)calloc(1, sizeof *g);
    if (!g) return NULL;
    *g = *gate;
    g->target_qubits = NULL; g->control_qubits = NULL; g->qubits = NULL;
    g->parameters = NULL; g->matrix = NULL; g->custom_data = NULL;
    const size_t dim = (size_t)1 << gate->num_qubits;
    g->target_qubits = (size_t *)malloc((gate->num_qubits ? gate->num_qubits : 1) * sizeof(size_t));
    g->qubits = (size_t *)malloc((gate->num_qubits ? gate->num_qubits : 1) * sizeof(size_t));
    g->matrix = (ComplexFloat *)malloc(dim * dim * sizeof(ComplexFloat));
    if (!g->target_qubits || !g->qubits || !g->matrix) { destroy_quantum_gate(g); return NULL; }
    if (gate->target_qubits) {
        memcpy(g->target_qubits, gate->target_qubits, gate->num_qubits * sizeof(size_t));
        memcpy(g->qubits, gate->target_qubits, gate->num_qubits * sizeof(size_t));
    }
    if (gate->matrix) memcpy(g->matrix, gate->matrix, dim * dim * sizeof(ComplexFloat));
    else fill_gate_matrix(g);
    if (gate->num_controls && gate->control_qubits) {
        g->control_qubits = (size_t *)malloc(gate->num_controls * sizeof(size_t));
        if (!g->control_qubits) { destroy_quantum_gate(g); return NULL; }
        memcpy(g->control_qubits, gate->control_qubits, gate->num_controls * sizeof(size_t));
    }
    if (gate->num_parameters && gate->parameters) {
        g->parameters = (double *)malloc(gate->num_parameters * sizeof(double));
        if (!g->parameters) { destroy_quantum_gate(g); return NULL; }
        memcpy(g->parameters, gate->parameters, gate->num_parameters * sizeof(double));
    }
    return g;
}

bool update_gate_parameters(quantum_gate_t *gate, const double *parameters, size_t num_parameters)
{
    if (!gate || !gate->is_parameterized || !parameters || num_parameters != gate->num_parameters ||
        (gate->type != GATE_TYPE_RX && gate->type != GATE_TYPE_RY && gate->type != GATE_TYPE_RZ))
        return false;
    memcpy(gate->parameters, parameters, num_parameters * sizeof(double));
    if (gate->matrix) fill_gate_matrix(gate);
    return true;
}

void destroy_quantum_gate(quantum_gate_t *gate)
{
    if (!gate) return;
    free(gate->matrix);
    free(gate->target_qubits);
    free(gate->qubits);
    free(gate->control_qubits);
    free(gate->parameters);
    free(gate);
}

/* ---------------------------------------------------------------- network */
#define QGTN_CACHE_MAGIC 0x51475443u   /* "QGTC" */

typedef struct {
    uint32_t magic;
    uint64_t key;            /* hash of the circuit content (structure + angle bits) */
    int valid;
    int P;
    double *Q;               /* P x P complex128 */
    size_t *layer_cap;       /* gate capacity per layer */
} qgtn_cache_t;

static qgtn_cache_t *cache_of(const quantum_geometric_tensor_network_t *q)
{
    qgtn_cache_t *c = (qgtn_cache_t *)q->backend_state;
    return (c && c->magic == QGTN_CACHE_MAGIC) ? c : NULL;
}

quantum_geometric_tensor_network_t *create_quantum_geometric_tensor_network(
    size_t num_qubits, size_t num_layers, bool is_distributed, bool use_hardware_acceleration)
{
    if (num_qubits == 0 || num_qubits > QGTL_MAX_QUBITS) {
        set_qgtn_error("qubit count %zu outside 1..%d", num_qubits, QGTL_MAX_QUBITS);
        return NULL;
    }
    quantum_geometric_tensor_network_t *q = (quantum_geometric_tensor_network_t *)calloc(1, sizeof *q);
    quantum_circuit_t *c = (quantum_circuit_t *)calloc(1, sizeof *c);
    qgtn_cache_t *cache = (qgtn_cache_t *)calloc(1, sizeof *cache);
    const size_t nl = num_layers ? num_layers : 1;
    circuit_layer_t **layers = (circuit_layer_t **)calloc(nl, sizeof(circuit_layer_t *));
    size_t *cap = (size_t *)calloc(nl, sizeof(size_t));
    if (!q || !c || !cache || !layers || !cap) {
        free(q); free(c); free(cache); free(layers); free(cap);
        set_qgtn_error("Failed to allocate quantum geometric tensor network");
        return NULL;
    }
    cache->magic = QGTN_CACHE_MAGIC;
    cache->layer_cap = cap;
    c->layers = layers;
    c->num_layers = num_layers;
    c->layers_capacity = nl;
    c->num_qubits = num_qubits;
    q->circuit = c;
    q->num_qubits = num_qubits;
    q->num_layers = num_layers;
    q->is_distributed = is_distributed;
    q->use_hardware_acceleration = use_hardware_acceleration;
    q->hardware_config.type = QGTN_BACKEND_SIMULATOR;
    q->hardware_config.supports_gradients = true;
    q->hardware_config.supports_hybrid = true;
    q->backend_state = cache;
    return q;
}

void destroy_quantum_geometric_tensor_network(quantum_geometric_tensor_network_t *q)
{
    if (!q) return;
    if (q->circuit) {
        for (size_t l = 0; l < q->circuit->num_layers; l++) {
            circuit_layer_t *layer = q->circuit->layers ? q->circuit->layers[l] : NULL;
            if (!layer) continue;
            for (size_t g = 0; g < layer->num_gates; g++) destroy_quantum_gate(layer->gates[g]);
            free(layer->gates);
            free(layer);
        }
        free(q->circuit->layers);
        if (q->circuit->state) {
            free(q->circuit->state->coordinates);
            free(q->circuit->state);
        }
        free(q->circuit);
    }
    qgtn_cache_t *c = cache_of(q);
    if (c) {
        free(c->Q);
        free(c->layer_cap);
        free(c);
    }
    free(q);
}

bool apply_quantum_gate(quantum_geometric_tensor_network_t *q, const quantum_gate_t *gate,
                        const size_t *qubits, size_t num_qubits)
{
    if (!q || !gate || !qubits || num_qubits == 0) {
        set_qgtn_error("Invalid arguments to apply_quantum_gate");
        return false;
    }
    if (!q->circuit || !q->circuit->layers) {
        set_qgtn_error("Circuit not initialized");
        return false;
    }
    /* QGTL appends to the last layer (quantum_geometric_tensor_network.c:652) */
    const size_t li = q->num_layers > 0 ? q->num_layers - 1 : 0;
    if (li >= q->circuit->layers_capacity) {
        set_qgtn_error("Layer index out of bounds");
        return false;
    }
    qgtn_cache_t *cache = cache_of(q);
    circuit_layer_t *layer = q->circuit->layers[li];
    if (!layer) {
        layer = (circuit_layer_t *)calloc(1, sizeof *layer);
        if (!layer) { set_qgtn_error("Failed to create circuit layer"); return false; }
        q->circuit->layers[li] = layer;
    }
    size_t cap = cache ? cache->layer_cap[li] : layer->num_gates;
    if (layer->num_gates >= cap) {
        const size_t nc = cap ? 2 * cap : 16;
        quantum_gate_t **ng = (quantum_gate_t **)realloc(layer->gates, nc * sizeof(quantum_gate_t *));
        if (!ng) { set_qgtn_error("Failed to resize gates array"); return false; }
        layer->gates = ng;
        if (cache) cache->layer_cap[li] = nc;
    }
    quantum_gate_t *copy = copy_quantum_gate(gate);
    if (!copy) { set_qgtn_error("Failed to copy quantum gate"); return false; }
    layer->gates[layer->num_gates++] = copy;
    if (copy->is_parameterized) {
        layer->is_parameterized = true;
        q->circuit->is_parameterized = true;
    }
    if (cache) cache->valid = 0;
    return true;
}

/* Flatten the network's circuit into engine gates.  Parameters are numbered
 * in gate order.  Returns op count or a negative error. */
static int flatten(const quantum_geometric_tensor_network_t *q, qgtl_op_t **out, int *P_out)
{
    const quantum_circuit_t *c = q->circuit;
    size_t cap = 64, n = 0;
    int P = 0;
    qgtl_op_t *ops = (qgtl_op_t *)malloc(cap * sizeof(qgtl_op_t));
    if (!ops) return QGT_ERROR_NO_MEMORY;
    for (size_t l = 0; l < c->num_layers; l++) {
        const circuit_layer_t *layer = c->layers[l];
        if (!layer) continue;
        for (size_t gi = 0; gi < layer->num_gates; gi++) {
            const quantum_gate_t *g = layer->gates[gi];
            if (!g) continue;
            if (n + g->num_qubits + 1 >= cap) {
                cap *= 2;
                qgtl_op_t *no = (qgtl_op_t *)realloc(ops, cap * sizeof(qgtl_op_t));
                if (!no) { free(ops); return QGT_ERROR_NO_MEMORY; }
                ops = no;
            }
            const int two = (g->type == GATE_TYPE_CNOT || g->type == GATE_TYPE_CZ || g->type == GATE_TYPE_CY ||
                             g->type == GATE_TYPE_SWAP);
            if (two) {
                int a, b;
                if (g->num_controls > 0 && g->control_qubits && g->num_qubits >= 1) {
                    a = (int)g->control_qubits[0];          /* quantum_classical_algorithms.c:289-296 */
                    b = (int)g->target_qubits[0];
                } else if (g->num_qubits == 2) {
                    a = (int)g->target_qubits[0];           /* {control, target}: create_cnot_gate */
                    b = (int)g->target_qubits[1];
                } else {
                    free(ops);
                    return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "two-qubit gate without two qubits");
                }
                ops[n++] = (qgtl_op_t){g->type, a, b, -1, 0.0};
            } else {
                if (g->is_parameterized && g->num_qubits != 1) {
                    free(ops);
                    return qgtl_set_error(QGT_ERROR_NOT_IMPLEMENTED, "parameterised gate on %zu qubits", g->num_qubits);
                }
                const double angle = (g->parameters && g->num_parameters > 0) ? g->parameters[0] : 0.0;
                for (size_t t = 0; t < g->num_qubits; t++)
                    ops[n++] = (qgtl_op_t){g->type, (int)g->target_qubits[t], -1, g->is_parameterized ? P : -1, angle};
                if (g->is_parameterized) P++;
            }
        }
    }
    *out = ops;
    *P_out = P;
    return (int)n;
}

static uint64_t content_key(int n, int nops, const qgtl_op_t *ops)
{
    uint64_t h = qgtl_structure_hash(n, nops, ops, 0);
    for (int i = 0; i < nops; i++) {
        uint64_t bits;
        memcpy(&bits, &ops[i].angle, sizeof bits);
        h ^= bits + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    }
    return h;
}

static int qgtn_mode(void)
{
    const char *e = getenv("QGTL_QGT_MODE");
    return (e && (strcmp(e, "fp64") == 0 || strcmp(e, "plain") == 0)) ? QGTL_MODE_FP64_PLAIN : QGTL_MODE_MIXED;
}

/* Full QGT of the network (cached).  Returns P or a negative error. */
static int network_qgt(const quantum_geometric_tensor_network_t *q, const double **Qout)
{
    if (!q || !q->circuit) return qgtl_set_error(QGT_ERROR_INVALID_ARGUMENT, "null network");
    qgtn_cache_t *cache = cache_of(q);
    qgtl_op_t *ops = NULL;
    int P = 0;
    const int nops = flatten(q, &ops, &P);
    if (nops < 0) return nops;
    const uint64_t key = content_key((int)q->num_qubits, nops, ops);
    if (cache && cache->valid && cache->key == key) {
        free(ops);
        *Qout = cache->Q;
        return cache->P;
    }
    double *Q = (double *)malloc(sizeof(double) * 2 * (size_t)(P > 0 ? P : 1) * (size_t)(P > 0 ? P : 1));
    if (!Q) { free(ops); return QGT_ERROR_NO_MEMORY; }
    int rc = P > 0 ? qgtl_qgt((int)q->num_qubits, nops, ops, qgtn_mode(), Q, NULL) : 0;
    free(ops);
    if (rc) { free(Q); return rc; }
    if (!cache) { free(Q); return qgtl_set_error(QGT_ERROR_INVALID_STATE, "network was not created by this library"); }
    free(cache->Q);
    cache->Q = Q;
    cache->P = P;
    cache->key = key;
    cache->valid = 1;
    *Qout = Q;
    return P;
}

/* One gate of a flat circuit built by QGTL's circuit API
 * (src/quantum_geometric/core/quantum_circuit_operations.c:786-1135).  Those
 * builders set type, num_qubits, the unified qubits[] array (control first for
 * CNOT and CZ), parameters / num_parameters and custom_data only; the gate
 * struct comes from malloc, so target_qubits, control_qubits, matrix and
 * is_parameterized are indeterminate and are not read.  A rotation (RX, RY,
 * RZ) with a parameter is a parameterised gate, as quantum_circuit_execute
 * applies it (:1166-1177); quantum_circuit_phase's GATE_S carries an angle
 * that quantum_circuit_execute ignores (it applies RZ(pi/2), :1162-1165), so it
 * is S here; X, Y, Z, S differ from what quantum_circuit_execute applies
 * (RX(pi), RY(pi), RZ(pi), RZ(pi/2)) only by a global phase. */
static bool apply_flat_gate(quantum_geometric_tensor_network_t *q, const quantum_gate_t *gate)
{
    const size_t nq = gate->num_qubits;
    if (nq < 1 || nq > 2 || !gate->qubits) {
        set_qgtn_error("Invalid gate structure");
        return false;
    }
    for (size_t t = 0; t < nq; t++)
        if (gate->qubits[t] >= q->num_qubits) {
            set_qgtn_error("qubit index %zu out of range", gate->qubits[t]);
            return false;
        }
    const int rot = gate->type == GATE_TYPE_RX || gate->type == GATE_TYPE_RY || gate->type == GATE_TYPE_RZ;
    const int two = gate->type == GATE_TYPE_CNOT || gate->type == GATE_TYPE_CZ || gate->type == GATE_TYPE_CY ||
                    gate->type == GATE_TYPE_SWAP;
    const int one = gate->type == GATE_TYPE_H || gate->type == GATE_TYPE_X || gate->type == GATE_TYPE_Y ||
                    gate->type == GATE_TYPE_Z || gate->type == GATE_TYPE_S || gate->type == GATE_TYPE_T ||
                    gate->type == GATE_TYPE_I || rot;
    if ((two && nq != 2) || (one && nq != 1) || (!two && !one)) {
        set_qgtn_error("gate type %d on %zu qubits is not supported in a flat circuit", (int)gate->type, nq);
        return false;
    }
    if (rot && (gate->num_parameters < 1 || !gate->parameters)) {
        set_qgtn_error("rotation gate without its angle");
        return false;
    }
    quantum_gate_t *g = create_quantum_gate(gate->type, gate->qubits, nq, rot ? gate->parameters : NULL, rot ? 1 : 0);
    if (!g) {
        set_qgtn_error("Failed to create quantum gate");
        return false;
    }
    const bool ok = apply_quantum_gate(q, g, g->target_qubits, nq);
    destroy_quantum_gate(g);
    return ok;
}

bool apply_quantum_circuit(quantum_geometric_tensor_network_t *q, const quantum_circuit_t *circuit)
{
    QGTL_LOCKED();
    if (!q || !circuit) {
        set_qgtn_error("Invalid arguments to apply_quantum_circuit");
        return false;
    }
    for (size_t l = 0; l < circuit->num_layers; l++) {
        const circuit_layer_t *layer = circuit->layers ? circuit->layers[l] : NULL;
        if (!layer) continue;
        for (size_t g = 0; g < layer->num_gates; g++) {
            const quantum_gate_t *gate = layer->gates[g];
            if (!gate) continue;
            if (!gate->matrix || !gate->target_qubits) {
                set_qgtn_error("Invalid gate structure");
                return false;
            }
            if (!apply_quantum_gate(q, gate, gate->target_qubits, gate->num_qubits)) return false;
        }
    }
    /* a flat gate list (no layers), as QGTL's circuit builders make it */
    if (circuit->num_layers == 0)
        for (size_t g = 0; g < circuit->num_gates; g++) {
            const quantum_gate_t *gate = circuit->gates ? circuit->gates[g] : NULL;
            if (!gate) continue;
            if (!apply_flat_gate(q, gate)) return false;
        }
    /* refresh the network's state (QGTL keeps it in circuit->state) */
    ComplexFloat *sv = NULL;
    size_t dim = 0;
    if (!get_quantum_state(q, &sv, &dim)) return false;
    if (!q->circuit->state) {
        q->circuit->state = (quantum_geometric_state_t *)calloc(1, sizeof(quantum_geometric_state_t));
        if (!q->circuit->state) { free(sv); set_qgtn_error("Failed to allocate circuit state"); return false; }
    }
    free(q->circuit->state->coordinates);
    q->circuit->state->coordinates = sv;
    q->circuit->state->dimension = dim;
    q->circuit->state->manifold_dim = dim;
    q->circuit->state->num_qubits = q->num_qubits;
    q->circuit->state->is_normalized = true;
    q->circuit->state->hardware = HARDWARE_TYPE_GPU;
    return true;
}

bool get_quantum_state(const quantum_geometric_tensor_network_t *q, ComplexFloat **state_vector, size_t *dimension)
{
    QGTL_LOCKED();
    if (!q || !state_vector || !dimension || !q->circuit) {
        set_qgtn_error("Invalid arguments to get_quantum_state");
        return false;
    }
    qgtl_op_t *ops = NULL;
    int P = 0;
    const int nops = flatten(q, &ops, &P);
    if (nops < 0) { set_qgtn_error("%s", qgtl_last_error()); return false; }
    const size_t N = (size_t)1 << q->num_qubits;
    double *psi = (double *)malloc(N * 2 * sizeof(double));
    ComplexFloat *out = (ComplexFloat *)malloc(N * sizeof(ComplexFloat));
    if (!psi || !out) { free(ops); free(psi); free(out); set_qgtn_error("out of memory"); return false; }
    const int rc = qgtl_statevector((int)q->num_qubits, nops, ops, qgtn_mode(), psi);
    free(ops);
    if (rc) {
        free(psi); free(out);
        set_qgtn_error("Failed to compute state vector: %s", qgtl_last_error());
        return false;
    }
    for (size_t i = 0; i < N; i++) out[i] = (ComplexFloat){(float)psi[2 * i], (float)psi[2 * i + 1]};
    free(psi);
    *state_vector = out;
    *dimension = N;
    return true;
}

bool compute_quantum_geometric_tensor(const quantum_geometric_tensor_network_t *q, size_t i, size_t j,
                                      ComplexFloat *result)
{
    QGTL_LOCKED();
    if (!q || !result) {
        set_qgtn_error("Invalid arguments to compute_quantum_geometric_tensor");
        return false;
    }
    const double *Q = NULL;
    const int P = network_qgt(q, &Q);
    if (P < 0) { set_qgtn_error("%s", qgtl_last_error()); return false; }
    if (i >= (size_t)P || j >= (size_t)P) {
        set_qgtn_error("parameter index out of range (P = %d)", P);
        return false;
    }
    result->real = (float)Q[2 * (i * (size_t)P + j)];
    result->imag = (float)Q[2 * (i * (size_t)P + j) + 1];
    return true;
}

bool compute_quantum_metric(const quantum_geometric_tensor_network_t *q, size_t i, size_t j, double *result)
{
    if (!q || !result) { set_qgtn_error("Invalid arguments to compute_quantum_metric"); return false; }
    ComplexFloat t;
    if (!compute_quantum_geometric_tensor(q, i, j, &t)) return false;
    *result = t.real;
    return true;
}

bool compute_berry_curvature(const quantum_geometric_tensor_network_t *q, size_t i, size_t j, double *result)
{
    if (!q || !result) { set_qgtn_error("Invalid arguments to compute_berry_curvature"); return false; }
    ComplexFloat t;
    if (!compute_quantum_geometric_tensor(q, i, j, &t)) return false;
    *result = t.imag;
    return true;
}

/* ------------------------------------------------------ metric / curvature */
#define CHECK_NULL(p) do { if ((p) == NULL) return QGT_ERROR_INVALID_PARAMETER; } while (0)
#define CHECK_ARG(c) do { if (!(c)) return QGT_ERROR_INVALID_PARAMETER; } while (0)

qgt_error_t geometric_create_metric(quantum_geometric_metric_t **metric, geometric_metric_type_t type, size_t dimension)
{
    CHECK_NULL(metric);
    CHECK_ARG(dimension > 0 && dimension <= QGTL_MAX_PARAMS);
    if (type != GEOMETRIC_METRIC_EUCLIDEAN && type != GEOMETRIC_METRIC_MINKOWSKI && type != GEOMETRIC_METRIC_FUBINI_STUDY)
        return QGT_ERROR_INVALID_ARGUMENT;
    void *mem = NULL, *comp = NULL;
    if (posix_memalign(&mem, 32, sizeof(quantum_geometric_metric_t)) != 0) return QGT_ERROR_ALLOCATION_FAILED;
    if (posix_memalign(&comp, 32, dimension * dimension * sizeof(ComplexFloat)) != 0) { free(mem); return QGT_ERROR_ALLOCATION_FAILED; }
    quantum_geometric_metric_t *m = (quantum_geometric_metric_t *)mem;
    memset(m, 0, sizeof *m);
    m->components = (ComplexFloat *)comp;
    m->type = type;
    m->dimension = dimension;
    m->is_symmetric = true;
    m->hardware = HARDWARE_TYPE_CPU;
    memset(m->components, 0, dimension * dimension * sizeof(ComplexFloat));
    for (size_t i = 0; i < dimension; i++)
        m->components[i * dimension + i].real = (type == GEOMETRIC_METRIC_MINKOWSKI && i == 0) ? -1.f : 1.f;
    *metric = m;
    return QGT_SUCCESS;
}

void geometric_destroy_metric(quantum_geometric_metric_t *metric)
{
    if (!metric) return;
    free(metric->components);
    free(metric);
}

qgt_error_t geometric_create_curvature(quantum_geometric_curvature_t **curvature, geometric_curvature_type_t type,
                                       size_t dimension)
{
    CHECK_NULL(curvature);
    /* rank-4 tensors keep QGTL's QGT_MAX_DIMENSIONS = 16 */
    CHECK_ARG(dimension > 0 && dimension <= (type == GEOMETRIC_CURVATURE_BERRY ? (size_t)QGTL_MAX_PARAMS : 16u));
    quantum_geometric_curvature_t *c = (quantum_geometric_curvature_t *)calloc(1, sizeof *c);
    if (!c) return QGT_ERROR_ALLOCATION_FAILED;
    const size_t count = type == GEOMETRIC_CURVATURE_BERRY ? dimension * dimension
                                                           : dimension * dimension * dimension * dimension;
    c->components = (ComplexFloat *)calloc(count, sizeof(ComplexFloat));
    if (!c->components) { free(c); return QGT_ERROR_ALLOCATION_FAILED; }
    c->type = type;
    c->dimension = dimension;
    c->hardware = HARDWARE_TYPE_CPU;
    *curvature = c;
    return QGT_SUCCESS;
}

void geometric_destroy_curvature(quantum_geometric_curvature_t *curvature)
{
    if (!curvature) return;
    free(curvature->components);
    free(curvature);
}

/* P of the network, or an error when num_params exceeds it (QGTL fails the
 * element call for a parameter index it cannot find). */
static int network_qgt_for(const quantum_geometric_tensor_network_t *q, size_t num_params, const double **Q)
{
    const int P = network_qgt(q, Q);
    if (P < 0) return QGT_ERROR_COMPUTATION_FAILED;
    if (num_params > (size_t)P) return QGT_ERROR_COMPUTATION_FAILED;
    return P;
}

qgt_error_t geometric_compute_fubini_study_metric(quantum_geometric_metric_t *metric,
                                                  const quantum_geometric_tensor_network_t *qgtn, size_t num_params)
{
    QGTL_LOCKED();
    CHECK_NULL(metric);
    CHECK_NULL(qgtn);
    CHECK_ARG(num_params > 0);
    if (metric->type != GEOMETRIC_METRIC_FUBINI_STUDY) return QGT_ERROR_INVALID_ARGUMENT;
    if (metric->dimension != num_params) return QGT_ERROR_DIMENSION_MISMATCH;
    const double *Q = NULL;
    const int P = network_qgt_for(qgtn, num_params, &Q);
    if (P < 0) return (qgt_error_t)P;
    for (size_t mu = 0; mu < num_params; mu++)
        for (size_t nu = mu; nu < num_params; nu++) {
            const float g = (float)Q[2 * (mu * (size_t)P + nu)];
            metric->components[mu * num_params + nu] = (ComplexFloat){g, 0.f};
            metric->components[nu * num_params + mu] = (ComplexFloat){g, 0.f};
        }
    metric->is_symmetric = true;
    return QGT_SUCCESS;
}

qgt_error_t geometric_compute_fubini_study_element(const quantum_geometric_tensor_network_t *qgtn, size_t mu,
                                                   size_t nu, float *result)
{
    CHECK_NULL(qgtn);
    CHECK_NULL(result);
    double v;
    if (!compute_quantum_metric(qgtn, mu, nu, &v)) return QGT_ERROR_COMPUTATION_FAILED;
    *result = (float)v;
    return QGT_SUCCESS;
}

qgt_error_t geometric_compute_berry_curvature(quantum_geometric_curvature_t *curvature,
                                              const quantum_geometric_tensor_network_t *qgtn, size_t num_params)
{
    QGTL_LOCKED();
    CHECK_NULL(curvature);
    CHECK_NULL(qgtn);
    CHECK_ARG(num_params > 0);
    if (curvature->type != GEOMETRIC_CURVATURE_BERRY) return QGT_ERROR_INVALID_ARGUMENT;
    if (curvature->dimension != num_params) return QGT_ERROR_DIMENSION_MISMATCH;
    const double *Q = NULL;
    const int P = network_qgt_for(qgtn, num_params, &Q);
    if (P < 0) return (qgt_error_t)P;
    for (size_t mu = 0; mu < num_params; mu++) {
        curvature->components[mu * num_params + mu] = (ComplexFloat){0.f, 0.f};
        for (size_t nu = mu + 1; nu < num_params; nu++) {
            const float w = (float)Q[2 * (mu * (size_t)P + nu) + 1];
            curvature->components[mu * num_params + nu] = (ComplexFloat){w, 0.f};
            curvature->components[nu * num_params + mu] = (ComplexFloat){-w, 0.f};
        }
    }
    curvature->is_flat = false;
    return QGT_SUCCESS;
}

qgt_error_t geometric_compute_berry_curvature_element(const quantum_geometric_tensor_network_t *qgtn, size_t mu,
                                                      size_t nu, float *result)
{
    CHECK_NULL(qgtn);
    CHECK_NULL(result);
    double v;
    if (!compute_berry_curvature(qgtn, mu, nu, &v)) return QGT_ERROR_COMPUTATION_FAILED;
    *result = (float)v;
    return QGT_SUCCESS;
}

qgt_error_t geometric_compose_qgt(ComplexFloat *qgt, const quantum_geometric_metric_t *metric,
                                  const quantum_geometric_curvature_t *curvature, size_t dimension)
{
    CHECK_NULL(qgt);
    CHECK_NULL(metric);
    CHECK_NULL(curvature);
    if (metric->dimension != dimension || curvature->dimension != dimension) return QGT_ERROR_DIMENSION_MISMATCH;
    if (metric->type != GEOMETRIC_METRIC_FUBINI_STUDY) return QGT_ERROR_INVALID_ARGUMENT;
    if (curvature->type != GEOMETRIC_CURVATURE_BERRY) return QGT_ERROR_INVALID_ARGUMENT;
    for (size_t i = 0; i < dimension * dimension; i++)
        qgt[i] = (ComplexFloat){metric->components[i].real, curvature->components[i].real};
    return QGT_SUCCESS;
}

qgt_error_t geometric_compute_full_qgt(ComplexFloat *qgt, const quantum_geometric_tensor_network_t *qgtn,
                                       size_t num_params)
{
    QGTL_LOCKED();
    CHECK_NULL(qgt);
    CHECK_NULL(qgtn);
    CHECK_ARG(num_params > 0);
    const double *Q = NULL;
    const int P = network_qgt_for(qgtn, num_params, &Q);
    if (P < 0) return (qgt_error_t)P;
    for (size_t mu = 0; mu < num_params; mu++)
        for (size_t nu = mu; nu < num_params; nu++) {
            const ComplexFloat v = {(float)Q[2 * (mu * (size_t)P + nu)], (float)Q[2 * (mu * (size_t)P + nu) + 1]};
            qgt[mu * num_params + nu] = v;
            if (nu != mu) qgt[nu * num_params + mu] = (ComplexFloat){v.real, -v.imag};
        }
    return QGT_SUCCESS;
}

/* FP64 view of the cached QGT (additive): copies P x P complex128. */
int qgtl_network_qgt_f64(const quantum_geometric_tensor_network_t *qgtn, double *Q_out, int *P_out)
{
    QGTL_LOCKED();
    const double *Q = NULL;
    const int P = network_qgt(qgtn, &Q);
    if (P < 0) return P;
    if (P_out) *P_out = P;
    if (Q_out) memcpy(Q_out, Q, sizeof(double) * 2 * (size_t)P * (size_t)P);
    return 0;
}
