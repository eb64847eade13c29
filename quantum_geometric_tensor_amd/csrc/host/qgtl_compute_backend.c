/*
 * qgtl_compute_backend.c -- QGTL's compute-backend plugin ABI
 * (include/quantum_geometric/supercomputer/compute_backend.h:17-263) on the
 * MI355X: memory, streams and events map onto the library's HIP runtime, and
 * the quantum slots run this library's whole-vector kernels
 * (csrc/hip/qgtl_statevec.hip) with FP64 reductions.
 *
 * Behaviour follows QGTL's CPU backend
 * (src/quantum_geometric/supercomputer/backends/compute_cpu.c:341-460 for the
 * quantum slots, :480-660 for the single-node collectives, :664-790 for
 * execute / create_plan); pointers may be host memory (staged through device
 * buffers for the call, as QGTL's CUDA backend does in
 * backends/compute_cuda.cu:688-734) or device memory from this backend's
 * alloc (used in place).  All work is ordered on the library's one stream, so
 * a stream handle is a token and every call that hands results back in host
 * memory returns after they are there.  One process drives one GPU: a
 * configuration with more than one node is refused at init.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qgtl_compute_backend.h"
#include "qgtl_internal.h"

#define BACKEND_MAGIC 0x51474842u   /* "QGHB" */

struct ComputeBackend {
    uint32_t magic;
    double *partials, *sums;        /* device scratch of the reductions */
    ComputeMetrics metrics;
    size_t allocated, peak;
};

struct ComputeStream {
    int token;
};

struct ComputeEvent {
    qgtl_hip_event_t ev;
};

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static int valid(const ComputeBackend *b) { return b && b->magic == BACKEND_MAGIC; }

static ComputeResult map_rc(int rc)
{
    if (!rc) return COMPUTE_SUCCESS;
    return rc == QGT_ERROR_GPU_OUT_OF_MEMORY ? COMPUTE_ERROR_OUT_OF_MEMORY : COMPUTE_ERROR_KERNEL_FAILED;
}

/* ------------------------------------------------------------ lifecycle */
static ComputeBackend *hip_init(const ComputeDistributedConfig *config)
{
    QGTL_LOCKED();
    if (config && config->num_nodes > 1) {
        qgtl_set_error(QGT_ERROR_NOT_IMPLEMENTED, "the HIP backend drives one GPU of one node");
        return NULL;
    }
    if (qgtl_init_if_needed()) return NULL;
    ComputeBackend *b = (ComputeBackend *)calloc(1, sizeof *b);
    if (!b) return NULL;
    if (qgtl_hip_malloc((void **)&b->partials, sizeof(double) * (size_t)qgtl_hip_vec_partials_len()) ||
        qgtl_hip_malloc((void **)&b->sums, sizeof(double) * 4)) {
        qgtl_hip_free(b->partials);
        free(b);
        return NULL;
    }
    b->magic = BACKEND_MAGIC;
    return b;
}

static void hip_cleanup(ComputeBackend *b)
{
    QGTL_LOCKED();
    if (!valid(b)) return;
    qgtl_hip_sync();
    qgtl_hip_free(b->partials);
    qgtl_hip_free(b->sums);
    b->magic = 0;
    free(b);
}

static bool hip_probe(void) { return qgtl_hip_device_count() > 0; }

static ComputeResult hip_get_capabilities(ComputeBackend *b, int *num_devices, size_t *total_memory)
{
    QGTL_LOCKED();
    if (!valid(b)) return COMPUTE_ERROR_INVALID_ARGUMENT;
    if (num_devices) *num_devices = 1;
    if (total_memory && qgtl_hip_device_total_mem(total_memory)) return COMPUTE_ERROR_DEVICE_NOT_FOUND;
    return COMPUTE_SUCCESS;
}

/* --------------------------------------------------------------- memory */
static void *hip_alloc(ComputeBackend *b, size_t size, ComputeMemType type)
{
    QGTL_LOCKED();
    if (!valid(b) || size == 0) return NULL;
    void *p = NULL;
    switch (type) {
    case COMPUTE_MEM_HOST: p = malloc(size); break;
    case COMPUTE_MEM_DEVICE: if (qgtl_hip_malloc(&p, size)) p = NULL; break;
    /* pinned host memory is addressable from the device too, which is what
     * unified memory is asked for */
    case COMPUTE_MEM_PINNED:
    case COMPUTE_MEM_UNIFIED: if (qgtl_hip_host_alloc(&p, size)) p = NULL; break;
    default: return NULL;
    }
    if (p) {
        b->allocated += size;
        if (b->allocated > b->peak) b->peak = b->allocated;
    }
    return p;
}

static void hip_free(ComputeBackend *b, void *ptr, ComputeMemType type)
{
    QGTL_LOCKED();
    if (!ptr) return;
    switch (type) {
    case COMPUTE_MEM_HOST: free(ptr); break;
    case COMPUTE_MEM_DEVICE: qgtl_hip_sync(); qgtl_hip_free(ptr); break;
    default: qgtl_hip_sync(); qgtl_hip_host_free(ptr); break;
    }
    (void)b;
}

static ComputeResult hip_memcpy(ComputeBackend *b, void *dst, ComputeMemType dst_type, const void *src,
                                ComputeMemType src_type, size_t size, ComputeStream *stream)
{
    QGTL_LOCKED();
    if (!valid(b) || !dst || !src) return COMPUTE_ERROR_INVALID_ARGUMENT;
    if (size == 0) return COMPUTE_SUCCESS;
    const double t0 = now_ms();
    int rc;
    if (dst_type == COMPUTE_MEM_HOST && src_type == COMPUTE_MEM_HOST) {
        memcpy(dst, src, size);
        rc = 0;
    } else {
        rc = qgtl_hip_copy(dst, src, size);
        /* a host destination is readable when the call returns, as for a NULL stream */
        if (!rc && (stream == NULL || dst_type != COMPUTE_MEM_DEVICE)) rc = qgtl_hip_sync();
    }
    b->metrics.communication_time_ms += now_ms() - t0;
    return map_rc(rc);
}

static ComputeResult hip_memset(ComputeBackend *b, void *ptr, int value, size_t size, ComputeStream *stream)
{
    QGTL_LOCKED();
    if (!valid(b) || !ptr) return COMPUTE_ERROR_INVALID_ARGUMENT;
    if (!qgtl_hip_is_device_ptr(ptr)) {
        memset(ptr, value, size);
        return COMPUTE_SUCCESS;
    }
    int rc = qgtl_hip_memset(ptr, value, size);
    if (!rc && !stream) rc = qgtl_hip_sync();
    return map_rc(rc);
}

/* -------------------------------------------------------- streams, events */
static ComputeStream *hip_create_stream(ComputeBackend *b)
{
    if (!valid(b)) return NULL;
    return (ComputeStream *)calloc(1, sizeof(ComputeStream));
}

static void hip_destroy_stream(ComputeBackend *b, ComputeStream *s)
{
    (void)b;
    free(s);
}

static ComputeResult hip_synchronize_stream(ComputeBackend *b, ComputeStream *s)
{
    QGTL_LOCKED();
    (void)s;
    if (!valid(b)) return COMPUTE_ERROR_INVALID_ARGUMENT;
    return qgtl_hip_sync() ? COMPUTE_ERROR_SYNCHRONIZATION_FAILED : COMPUTE_SUCCESS;
}

static ComputeEvent *hip_create_event(ComputeBackend *b)
{
    QGTL_LOCKED();
    if (!valid(b)) return NULL;
    ComputeEvent *e = (ComputeEvent *)calloc(1, sizeof *e);
    if (e && qgtl_hip_event_create(&e->ev)) { free(e); e = NULL; }
    return e;
}

static void hip_destroy_event(ComputeBackend *b, ComputeEvent *e)
{
    QGTL_LOCKED();
    (void)b;
    if (!e) return;
    qgtl_hip_event_destroy(e->ev);
    free(e);
}

static ComputeResult hip_record_event(ComputeBackend *b, ComputeEvent *e, ComputeStream *s)
{
    QGTL_LOCKED();
    (void)s;
    if (!valid(b) || !e) return COMPUTE_ERROR_INVALID_ARGUMENT;
    return map_rc(qgtl_hip_event_record(e->ev));
}

/* every stream token is the library stream, which is ordered already */
static ComputeResult hip_wait_event(ComputeBackend *b, ComputeStream *s, ComputeEvent *e)
{
    (void)s;
    if (!valid(b) || !e) return COMPUTE_ERROR_INVALID_ARGUMENT;
    return COMPUTE_SUCCESS;
}

/* ------------------------------------------------------ operand staging */
typedef struct {
    void *dev;       /* device address of the operand */
    void *tmp;       /* staging buffer this call allocated, or NULL */
} operand_t;

static int stage(const void *p, size_t bytes, int upload, operand_t *o)
{
    o->tmp = NULL;
    if (qgtl_hip_is_device_ptr(p)) {
        o->dev = (void *)p;
        return 0;
    }
    int rc = qgtl_hip_malloc(&o->tmp, bytes);
    if (!rc && upload) rc = qgtl_hip_h2d(o->tmp, p, bytes);
    o->dev = o->tmp;
    return rc;
}

/* copy a staged in/out operand back to its host location, then free the stage */
static int unstage(void *host, size_t bytes, operand_t *o, int download)
{
    int rc = 0;
    if (o->tmp && download) rc = qgtl_hip_d2h(host, o->tmp, bytes);
    if (!rc) rc = qgtl_hip_sync();
    qgtl_hip_free(o->tmp);
    o->tmp = NULL;
    return rc;
}

/* four FP64 sums of <a|b> (see qgtl_hip_vec_dotc_c64) into host out4 */
static int dotc(ComputeBackend *b, const float *a, const float *bb, size_t n, double out4[4])
{
    operand_t oa, ob;
    int rc = stage(a, n * 8, 1, &oa);
    if (rc) return rc;
    rc = stage(bb, n * 8, 1, &ob);
    if (!rc) rc = qgtl_hip_vec_dotc_c64(oa.dev, ob.dev, n, b->partials, b->sums);
    if (!rc) rc = qgtl_hip_d2h(out4, b->sums, 4 * sizeof(double));
    const int r1 = unstage(NULL, 0, &oa, 0), r2 = unstage(NULL, 0, &ob, 0);
    return rc ? rc : (r1 ? r1 : r2);
}

/* ------------------------------------------------------ quantum slots */
static ComputeResult hip_quantum_unitary(ComputeBackend *b, float *state, size_t state_size, const float *unitary,
                                         size_t unitary_size, ComputeStream *stream)
{
    QGTL_LOCKED();
    (void)stream;
    if (!valid(b) || !state || !unitary || state_size == 0 || unitary_size != state_size)
        return COMPUTE_ERROR_INVALID_ARGUMENT;
    const double t0 = now_ms();
    operand_t os, ou, oy;
    int rc = stage(state, state_size * 8, 1, &os);
    if (rc) return map_rc(rc);
    rc = stage(unitary, state_size * state_size * 8, 1, &ou);
    oy.tmp = NULL;
    if (!rc) rc = qgtl_hip_malloc(&oy.tmp, state_size * 8);
    if (!rc) rc = qgtl_hip_vec_cgemv_c64(ou.dev, os.dev, oy.tmp, state_size, state_size);
    if (!rc) rc = qgtl_hip_d2d(os.dev, oy.tmp, state_size * 8);
    const int r1 = unstage(state, state_size * 8, &os, !rc);
    unstage(NULL, 0, &ou, 0);
    unstage(NULL, 0, &oy, 0);
    if (!rc) rc = r1;
    b->metrics.compute_time_ms += now_ms() - t0;
    b->metrics.flops += 8.0 * (double)state_size * (double)state_size;
    return map_rc(rc);
}

static ComputeResult hip_quantum_normalize(ComputeBackend *b, float *state, size_t size, ComputeStream *stream)
{
    QGTL_LOCKED();
    (void)stream;
    if (!valid(b) || !state || size == 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    const double t0 = now_ms();
    operand_t os;
    int rc = stage(state, size * 8, 1, &os);
    if (rc) return map_rc(rc);
    double s[4] = {0, 0, 0, 0};
    rc = qgtl_hip_vec_dotc_c64(os.dev, os.dev, size, b->partials, b->sums);
    if (!rc) rc = qgtl_hip_d2h(s, b->sums, sizeof s);
    if (!rc) rc = qgtl_hip_sync();
    const float norm = (float)sqrt(s[2]);
    int changed = 0;
    if (!rc && norm > 1e-10f) {
        rc = qgtl_hip_vec_scale_c64(os.dev, size, 1.0f / norm);
        changed = 1;
    }
    const int r1 = unstage(state, size * 8, &os, !rc && changed);
    b->metrics.compute_time_ms += now_ms() - t0;
    return map_rc(rc ? rc : r1);
}

static ComputeResult hip_quantum_tensor_contract(ComputeBackend *b, float *result, const float *a, const float *x,
                                                 size_t m, size_t n, size_t k, ComputeStream *stream)
{
    QGTL_LOCKED();
    (void)stream;
    if (!valid(b) || !result || !a || !x || m == 0 || n == 0 || k == 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    const double t0 = now_ms();
    operand_t oa, ox, oc;
    int rc = stage(a, m * n * 8, 1, &oa);
    if (rc) return map_rc(rc);
    rc = stage(x, n * k * 8, 1, &ox);
    if (!rc) rc = stage(result, m * k * 8, 0, &oc);
    else oc.tmp = NULL;
    if (!rc) rc = qgtl_hip_vec_cgemm_c64(oa.dev, ox.dev, oc.dev, m, n, k);
    const int r1 = unstage(result, m * k * 8, &oc, !rc);
    unstage(NULL, 0, &oa, 0);
    unstage(NULL, 0, &ox, 0);
    b->metrics.compute_time_ms += now_ms() - t0;
    b->metrics.flops += 8.0 * (double)m * (double)n * (double)k;
    return map_rc(rc ? rc : r1);
}

static ComputeResult hip_quantum_gradient(ComputeBackend *b, float *gradients, const float *forward_state,
                                          const float *backward_state, size_t size, ComputeStream *stream)
{
    QGTL_LOCKED();
    (void)stream;
    if (!valid(b) || !gradients || !forward_state || !backward_state || size == 0)
        return COMPUTE_ERROR_INVALID_ARGUMENT;
    double s[4];
    const int rc = dotc(b, backward_state, forward_state, size, s);
    if (rc) return map_rc(rc);
    gradients[0] = (float)s[0];
    gradients[1] = (float)s[1];
    return COMPUTE_SUCCESS;
}

static ComputeResult hip_quantum_inner_product(ComputeBackend *b, float *result, const float *state_a,
                                               const float *state_b, size_t size, ComputeStream *stream)
{
    QGTL_LOCKED();
    (void)stream;
    if (!valid(b) || !result || !state_a || !state_b || size == 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    double s[4];
    const int rc = dotc(b, state_a, state_b, size, s);
    if (rc) return map_rc(rc);
    result[0] = (float)s[0];
    result[1] = (float)s[1];
    return COMPUTE_SUCCESS;
}

static ComputeResult hip_quantum_expectation(ComputeBackend *b, float *result, const float *state,
                                             const float *observable, size_t size, ComputeStream *stream)
{
    QGTL_LOCKED();
    (void)stream;
    if (!valid(b) || !result || !state || !observable || size == 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    operand_t os, oo;
    int rc = stage(state, size * 8, 1, &os);
    if (rc) return map_rc(rc);
    rc = stage(observable, size * 4, 1, &oo);
    double s[4] = {0, 0, 0, 0};
    if (!rc) rc = qgtl_hip_vec_expect_diag_c64(os.dev, (const float *)oo.dev, size, b->partials, b->sums);
    if (!rc) rc = qgtl_hip_d2h(s, b->sums, sizeof s);
    const int r1 = unstage(NULL, 0, &os, 0), r2 = unstage(NULL, 0, &oo, 0);
    if (!rc) rc = r1 ? r1 : r2;
    if (!rc) *result = (float)s[0];
    return map_rc(rc);
}

/* ------------------------------------------------ single-node collectives */
static size_t dtype_size(ComputeDataType t)
{
    switch (t) {
    case COMPUTE_DTYPE_FLOAT32: case COMPUTE_DTYPE_INT32: return 4;
    case COMPUTE_DTYPE_FLOAT64: case COMPUTE_DTYPE_COMPLEX64: case COMPUTE_DTYPE_INT64: return 8;
    case COMPUTE_DTYPE_COMPLEX128: return 16;
    case COMPUTE_DTYPE_UINT8: return 1;
    default: return 0;
    }
}

static ComputeResult local_copy(ComputeBackend *b, void *dst, const void *src, size_t count, ComputeDataType t)
{
    QGTL_LOCKED();
    if (!valid(b) || !dst || !src || dtype_size(t) == 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    if (dst == src || count == 0) return COMPUTE_SUCCESS;
    int rc = qgtl_hip_copy(dst, src, count * dtype_size(t));
    if (!rc) rc = qgtl_hip_sync();
    return map_rc(rc);
}

static ComputeResult hip_barrier(ComputeBackend *b)
{
    QGTL_LOCKED();
    if (!valid(b)) return COMPUTE_ERROR_INVALID_ARGUMENT;
    return qgtl_hip_sync() ? COMPUTE_ERROR_SYNCHRONIZATION_FAILED : COMPUTE_SUCCESS;
}

static ComputeResult hip_broadcast(ComputeBackend *b, void *data, size_t size, ComputeDataType t, int root)
{
    (void)size; (void)t;
    if (!valid(b) || !data || root != 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    return COMPUTE_SUCCESS;
}

static ComputeResult hip_allreduce(ComputeBackend *b, const void *send, void *recv, size_t count, ComputeDataType t,
                                   ComputeReduceOp op)
{
    (void)op;
    return local_copy(b, recv, send, count, t);
}

static ComputeResult hip_scatter(ComputeBackend *b, const void *send, void *recv, size_t count, ComputeDataType t,
                                 int root)
{
    if (root != 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    return local_copy(b, recv, send, count, t);
}

static ComputeResult hip_gather(ComputeBackend *b, const void *send, void *recv, size_t count, ComputeDataType t,
                                int root)
{
    if (root != 0) return COMPUTE_ERROR_INVALID_ARGUMENT;
    return local_copy(b, recv, send, count, t);
}

static ComputeResult hip_allgather(ComputeBackend *b, const void *send, void *recv, size_t count, ComputeDataType t)
{
    return local_copy(b, recv, send, count, t);
}

static ComputeResult hip_reduce_scatter(ComputeBackend *b, const void *send, void *recv, size_t count,
                                        ComputeDataType t, ComputeReduceOp op)
{
    (void)op;
    return local_copy(b, recv, send, count, t);
}

/* ------------------------------------------------------------ execution */
static ComputeResult hip_execute(ComputeBackend *b, const ComputeQuantumOp *op, const ComputeExecutionPlan *plan,
                                 ComputeStream *stream)
{
    (void)plan;
    if (!valid(b) || !op) return COMPUTE_ERROR_INVALID_ARGUMENT;
    ComputeResult r;
    switch (op->type) {
    case QUANTUM_OP_UNITARY:
        r = hip_quantum_unitary(b, (float *)op->output_data, op->output_size, (const float *)op->parameters,
                                op->param_size, stream);
        break;
    case QUANTUM_OP_NORMALIZE:
        r = hip_quantum_normalize(b, (float *)op->output_data, op->output_size, stream);
        break;
    case QUANTUM_OP_TENSOR_CONTRACT:
        r = (op->num_dims >= 3 && op->dims)
                ? hip_quantum_tensor_contract(b, (float *)op->output_data, (const float *)op->input_data,
                                              (const float *)op->parameters, op->dims[0], op->dims[1], op->dims[2],
                                              stream)
                : COMPUTE_ERROR_INVALID_ARGUMENT;
        break;
    case QUANTUM_OP_GRADIENT:
        r = hip_quantum_gradient(b, (float *)op->output_data, (const float *)op->input_data,
                                 (const float *)op->parameters, op->input_size, stream);
        break;
    case QUANTUM_OP_INNER_PRODUCT:
        r = hip_quantum_inner_product(b, (float *)op->output_data, (const float *)op->input_data,
                                      (const float *)op->parameters, op->input_size, stream);
        break;
    case QUANTUM_OP_EXPECTATION:
        r = hip_quantum_expectation(b, (float *)op->output_data, (const float *)op->input_data,
                                    (const float *)op->parameters, op->input_size, stream);
        break;
    default:
        r = COMPUTE_ERROR_NOT_IMPLEMENTED;
        break;
    }
    b->metrics.operations_per_second += 1.0;   /* QGTL's CPU backend counts operations here */
    return r;
}

static ComputeExecutionPlan *hip_create_plan(ComputeBackend *b, const ComputeQuantumOp *op)
{
    if (!valid(b) || !op) return NULL;
    ComputeExecutionPlan *p = (ComputeExecutionPlan *)calloc(1, sizeof *p);
    if (!p) return NULL;
    p->num_partitions = 1;
    p->partition_size = op->input_size;
    p->node_assignments = (int *)calloc(1, sizeof(int));
    p->offsets = (size_t *)calloc(1, sizeof(size_t));
    p->sizes = (size_t *)calloc(1, sizeof(size_t));
    if (!p->node_assignments || !p->offsets || !p->sizes) {
        free(p->node_assignments); free(p->offsets); free(p->sizes); free(p);
        return NULL;
    }
    p->sizes[0] = op->input_size;
    return p;
}

static void hip_destroy_plan(ComputeBackend *b, ComputeExecutionPlan *p)
{
    (void)b;
    if (!p) return;
    free(p->node_assignments);
    free(p->offsets);
    free(p->sizes);
    free(p->send_targets);
    free(p->recv_sources);
    free(p->workspace);
    free(p);
}

static ComputeResult hip_get_metrics(ComputeBackend *b, ComputeMetrics *m)
{
    if (!valid(b) || !m) return COMPUTE_ERROR_INVALID_ARGUMENT;
    *m = b->metrics;
    m->peak_memory_bytes = b->peak;
    m->current_memory_bytes = b->allocated;
    m->memory_used = b->allocated;
    return COMPUTE_SUCCESS;
}

static ComputeResult hip_reset_metrics(ComputeBackend *b)
{
    if (!valid(b)) return COMPUTE_ERROR_INVALID_ARGUMENT;
    memset(&b->metrics, 0, sizeof b->metrics);
    return COMPUTE_SUCCESS;
}

static const ComputeBackendOps g_hip_ops = {
    .init = hip_init,
    .cleanup = hip_cleanup,
    .probe = hip_probe,
    .get_capabilities = hip_get_capabilities,
    .alloc = hip_alloc,
    .free = hip_free,
    .memcpy = hip_memcpy,
    .memset = hip_memset,
    .create_stream = hip_create_stream,
    .destroy_stream = hip_destroy_stream,
    .synchronize_stream = hip_synchronize_stream,
    .create_event = hip_create_event,
    .destroy_event = hip_destroy_event,
    .record_event = hip_record_event,
    .wait_event = hip_wait_event,
    .quantum_unitary = hip_quantum_unitary,
    .quantum_normalize = hip_quantum_normalize,
    .quantum_tensor_contract = hip_quantum_tensor_contract,
    .quantum_gradient = hip_quantum_gradient,
    .quantum_inner_product = hip_quantum_inner_product,
    .quantum_expectation = hip_quantum_expectation,
    .barrier = hip_barrier,
    .broadcast = hip_broadcast,
    .allreduce = hip_allreduce,
    .scatter = hip_scatter,
    .gather = hip_gather,
    .allgather = hip_allgather,
    .reduce_scatter = hip_reduce_scatter,
    .execute = hip_execute,
    .create_plan = hip_create_plan,
    .destroy_plan = hip_destroy_plan,
    .get_metrics = hip_get_metrics,
    .reset_metrics = hip_reset_metrics,
};

static const ComputeBackendInfo g_hip_info = {
    .type = COMPUTE_BACKEND_CUDA,
    .name = "HIP (AMD MI355X)",
    .version = "1.0.0",
    .priority = 100,
    .ops = &g_hip_ops,
};

const ComputeBackendOps *qgtl_hip_compute_backend_ops(void) { return &g_hip_ops; }

const ComputeBackendInfo *qgtl_hip_compute_backend_info(void) { return &g_hip_info; }

/* QGTL's registry (compute_backend.c:41-64), when the program links it; a weak
 * reference, so this library also loads on its own.  The same registration
 * that COMPUTE_REGISTER_BACKEND (compute_backend.h:523-534) performs at load
 * time. */
extern ComputeResult compute_register_backend(const ComputeBackendInfo *info) __attribute__((weak));

__attribute__((constructor)) static void register_hip_backend(void)
{
    if (compute_register_backend) compute_register_backend(&g_hip_info);
}
