/*
 * qgtl_compute_backend.h -- a HIP implementation of QGTL's compute-backend
 * plugin ABI (include/quantum_geometric/supercomputer/compute_backend.h:17-263
 * and compute_types.h:72-348), so that QGTL's distributed/supercomputer
 * engine runs its state-vector operations on the MI355X through
 * libqgtl_mi355x.
 *
 * The types below keep QGTL's layouts and enum values (the vtable is called
 * through QGTL's own pointers, so member order is ABI).  QGTL has no HIP entry
 * in ComputeBackendType; this backend registers itself in the GPU slot,
 * COMPUTE_BACKEND_CUDA (priority 100, the priority QGTL's CUDA backend uses),
 * through QGTL's compute_register_backend when the program links QGTL's
 * registry (compute_backend.c:41-64; a weak reference, so this library also
 * loads without it), and qgtl_hip_compute_backend_info() returns the same
 * registration record for programs that register backends themselves.
 *
 * Semantics of the quantum slots follow QGTL's CPU backend
 * (src/quantum_geometric/supercomputer/backends/compute_cpu.c:341-460), with
 * FP64 accumulation on the device instead of FP32 on the host:
 *   quantum_unitary       state <- U state, U a state_size x state_size
 *                         row-major complex64 matrix (unitary_size must equal
 *                         state_size)
 *   quantum_normalize     state <- state / |state| when |state| > 1e-10
 *   quantum_inner_product result[0..1] = <a|b> = sum conj(a_i) b_i
 *   quantum_gradient      gradients[0..1] = <backward|forward>
 *   quantum_expectation   *result = sum |psi_i|^2 observable[i] (diagonal
 *                         observable of `size` real values)
 * Pointers may be host memory (staged through device buffers for the call,
 * as QGTL's CUDA backend stages them) or device memory from this backend's
 * alloc (used in place).  On one GPU the collectives are local copies.
 */
#ifndef QGTL_COMPUTE_BACKEND_H
#define QGTL_COMPUTE_BACKEND_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* compute_types.h:72-79 */
typedef enum {
    COMPUTE_BACKEND_AUTO = -1, COMPUTE_BACKEND_CPU = 0, COMPUTE_BACKEND_CUDA, COMPUTE_BACKEND_METAL,
    COMPUTE_BACKEND_OPENCL, COMPUTE_BACKEND_COUNT
} ComputeBackendType;

/* compute_types.h:85-90 */
typedef enum { COMPUTE_MEM_HOST = 0, COMPUTE_MEM_DEVICE, COMPUTE_MEM_UNIFIED, COMPUTE_MEM_PINNED } ComputeMemType;

/* compute_types.h:96-104 */
typedef enum {
    COMPUTE_DTYPE_FLOAT32 = 0, COMPUTE_DTYPE_FLOAT64, COMPUTE_DTYPE_COMPLEX64, COMPUTE_DTYPE_COMPLEX128,
    COMPUTE_DTYPE_INT32, COMPUTE_DTYPE_INT64, COMPUTE_DTYPE_UINT8
} ComputeDataType;

/* compute_types.h:110-116 */
typedef enum {
    COMPUTE_REDUCE_SUM = 0, COMPUTE_REDUCE_PROD, COMPUTE_REDUCE_MIN, COMPUTE_REDUCE_MAX, COMPUTE_REDUCE_AVG
} ComputeReduceOp;

/* compute_types.h:122-133 */
typedef enum {
    COMPUTE_SUCCESS = 0, COMPUTE_ERROR_INVALID_ARGUMENT, COMPUTE_ERROR_OUT_OF_MEMORY, COMPUTE_ERROR_DEVICE_NOT_FOUND,
    COMPUTE_ERROR_BACKEND_NOT_AVAILABLE, COMPUTE_ERROR_COMMUNICATION_FAILED, COMPUTE_ERROR_SYNCHRONIZATION_FAILED,
    COMPUTE_ERROR_KERNEL_FAILED, COMPUTE_ERROR_NOT_IMPLEMENTED, COMPUTE_ERROR_INTERNAL
} ComputeResult;

/* compute_types.h:139-148 */
typedef enum {
    QUANTUM_OP_UNITARY = 0, QUANTUM_OP_MEASUREMENT, QUANTUM_OP_TENSOR_CONTRACT, QUANTUM_OP_GRADIENT,
    QUANTUM_OP_NORMALIZE, QUANTUM_OP_INNER_PRODUCT, QUANTUM_OP_EXPECTATION, QUANTUM_OP_DENSITY_MATRIX
} QuantumOpType;

/* compute_types.h:154-162 */
typedef enum {
    COMPUTE_TOPO_SINGLE = 0, COMPUTE_TOPO_RING, COMPUTE_TOPO_TREE, COMPUTE_TOPO_MESH, COMPUTE_TOPO_TORUS,
    COMPUTE_TOPO_FULLY_CONNECTED, COMPUTE_TOPO_HYBRID
} ComputeTopologyType;

/* compute_types.h:169-176 */
typedef struct {
    bool enable_timing;
    bool enable_memory_tracking;
    bool enable_bandwidth_tracking;
    bool enable_power_monitoring;
    size_t sample_interval_ms;
    size_t history_size;
} ComputeMonitorConfig;

typedef struct ComputeBackendConfig ComputeBackendConfig;

/* compute_types.h:193-229 */
typedef struct {
    int num_nodes;
    int devices_per_node;
    ComputeTopologyType topology;
    int node_rank;
    int local_rank;
    int local_size;
    size_t device_buffer_size;
    size_t host_buffer_size;
    size_t comm_buffer_size;
    bool use_nccl;
    bool use_rdma;
    bool use_compression;
    ComputeBackendType preferred_backend;
    bool allow_fallback;
    int num_streams;
    int num_threads_per_node;
    bool enable_async;
    ComputeMonitorConfig *monitor_config;
    ComputeBackendConfig *backend_config;
} ComputeDistributedConfig;

/* compute_types.h:235-261 */
typedef struct {
    QuantumOpType type;
    void *input_data;
    size_t input_size;
    ComputeDataType input_dtype;
    void *output_data;
    size_t output_size;
    ComputeDataType output_dtype;
    void *parameters;
    size_t param_size;
    size_t *dims;
    size_t num_dims;
    size_t num_qubits;
    size_t *target_qubits;
    size_t num_targets;
    float *parameter_gradients;
    size_t num_parameters;
} ComputeQuantumOp;

/* compute_types.h:267-290 */
typedef struct {
    int *node_assignments;
    size_t num_partitions;
    size_t partition_size;
    size_t *offsets;
    size_t *sizes;
    int *send_targets;
    int *recv_sources;
    size_t num_comm_ops;
    void *workspace;
    size_t workspace_size;
    int priority;
    bool requires_sync;
} ComputeExecutionPlan;

typedef struct ComputeBackend ComputeBackend;
typedef struct ComputeStream ComputeStream;
typedef struct ComputeEvent ComputeEvent;

/* compute_types.h:320-348 */
typedef struct {
    double total_time_ms;
    double compute_time_ms;
    double communication_time_ms;
    double synchronization_time_ms;
    double execution_time;
    double operations_per_second;
    double flops;
    double bandwidth_gbps;
    size_t peak_memory_bytes;
    size_t current_memory_bytes;
    size_t memory_used;
    double memory_efficiency;
    size_t bytes_sent;
    size_t bytes_received;
    size_t num_messages;
    double avg_latency_us;
    double gate_fidelity;
    double state_fidelity;
    double error_rate;
} ComputeMetrics;

/* compute_backend.h:17-424, member order kept */
typedef struct ComputeBackendOps {
    ComputeBackend *(*init)(const ComputeDistributedConfig *config);
    void (*cleanup)(ComputeBackend *backend);
    bool (*probe)(void);
    ComputeResult (*get_capabilities)(ComputeBackend *backend, int *num_devices, size_t *total_memory);
    void *(*alloc)(ComputeBackend *backend, size_t size, ComputeMemType mem_type);
    void (*free)(ComputeBackend *backend, void *ptr, ComputeMemType mem_type);
    ComputeResult (*memcpy)(ComputeBackend *backend, void *dst, ComputeMemType dst_type, const void *src,
                            ComputeMemType src_type, size_t size, ComputeStream *stream);
    ComputeResult (*memset)(ComputeBackend *backend, void *ptr, int value, size_t size, ComputeStream *stream);
    ComputeStream *(*create_stream)(ComputeBackend *backend);
    void (*destroy_stream)(ComputeBackend *backend, ComputeStream *stream);
    ComputeResult (*synchronize_stream)(ComputeBackend *backend, ComputeStream *stream);
    ComputeEvent *(*create_event)(ComputeBackend *backend);
    void (*destroy_event)(ComputeBackend *backend, ComputeEvent *event);
    ComputeResult (*record_event)(ComputeBackend *backend, ComputeEvent *event, ComputeStream *stream);
    ComputeResult (*wait_event)(ComputeBackend *backend, ComputeStream *stream, ComputeEvent *event);
    ComputeResult (*quantum_unitary)(ComputeBackend *backend, float *state, size_t state_size, const float *unitary,
                                     size_t unitary_size, ComputeStream *stream);
    ComputeResult (*quantum_normalize)(ComputeBackend *backend, float *state, size_t size, ComputeStream *stream);
    ComputeResult (*quantum_tensor_contract)(ComputeBackend *backend, float *result, const float *a, const float *b,
                                             size_t m, size_t n, size_t k, ComputeStream *stream);
    ComputeResult (*quantum_gradient)(ComputeBackend *backend, float *gradients, const float *forward_state,
                                      const float *backward_state, size_t size, ComputeStream *stream);
    ComputeResult (*quantum_inner_product)(ComputeBackend *backend, float *result, const float *state_a,
                                           const float *state_b, size_t size, ComputeStream *stream);
    ComputeResult (*quantum_expectation)(ComputeBackend *backend, float *result, const float *state,
                                         const float *observable, size_t size, ComputeStream *stream);
    ComputeResult (*barrier)(ComputeBackend *backend);
    ComputeResult (*broadcast)(ComputeBackend *backend, void *data, size_t size, ComputeDataType dtype, int root);
    ComputeResult (*allreduce)(ComputeBackend *backend, const void *send_data, void *recv_data, size_t count,
                               ComputeDataType dtype, ComputeReduceOp op);
    ComputeResult (*scatter)(ComputeBackend *backend, const void *send_data, void *recv_data, size_t count,
                             ComputeDataType dtype, int root);
    ComputeResult (*gather)(ComputeBackend *backend, const void *send_data, void *recv_data, size_t count,
                            ComputeDataType dtype, int root);
    ComputeResult (*allgather)(ComputeBackend *backend, const void *send_data, void *recv_data, size_t count,
                               ComputeDataType dtype);
    ComputeResult (*reduce_scatter)(ComputeBackend *backend, const void *send_data, void *recv_data, size_t count,
                                    ComputeDataType dtype, ComputeReduceOp op);
    ComputeResult (*execute)(ComputeBackend *backend, const ComputeQuantumOp *op, const ComputeExecutionPlan *plan,
                             ComputeStream *stream);
    ComputeExecutionPlan *(*create_plan)(ComputeBackend *backend, const ComputeQuantumOp *op);
    void (*destroy_plan)(ComputeBackend *backend, ComputeExecutionPlan *plan);
    ComputeResult (*get_metrics)(ComputeBackend *backend, ComputeMetrics *metrics);
    ComputeResult (*reset_metrics)(ComputeBackend *backend);
} ComputeBackendOps;

/* compute_backend.h:430-436 */
typedef struct {
    ComputeBackendType type;
    const char *name;
    const char *version;
    int priority;
    const ComputeBackendOps *ops;
} ComputeBackendInfo;

/* Additive (no counterpart in QGTL): the HIP backend's vtable and its
 * registration record (type COMPUTE_BACKEND_CUDA, name "HIP (AMD MI355X)",
 * priority 100). */
const ComputeBackendOps *qgtl_hip_compute_backend_ops(void);
const ComputeBackendInfo *qgtl_hip_compute_backend_info(void);

#ifdef __cplusplus
}
#endif
#endif
