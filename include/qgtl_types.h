/*
 * qgtl_types.h -- the QGTL types that the QGT path's C API passes across the
 * library boundary, laid out field-for-field as QGTL declares them so that a
 * program compiled against QGTL's headers can link against libqgtl_mi355x.so.
 *
 * Each definition cites the QGTL declaration whose memory layout it keeps
 * (paths relative to the QGTL source root).  Pointers to subsystems this
 * library does not implement (tensor networks, computational graphs,
 * stabilizer/anyon data) are kept as opaque pointers of the same size.
 */
#ifndef QGTL_TYPES_H
#define QGTL_TYPES_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/quantum_geometric/core/quantum_complex.h:13-22 */
#ifndef COMPLEX_FLOAT_DEFINED
#define COMPLEX_FLOAT_DEFINED
typedef struct { float real; float imag; } ComplexFloat;
typedef struct { double real; double imag; } ComplexDouble;
#endif

/* include/quantum_geometric/core/quantum_base_types.h:32-72 (values kept) */
typedef enum {
    GATE_TYPE_I = 0, GATE_TYPE_X = 1, GATE_TYPE_Y = 2, GATE_TYPE_Z = 3, GATE_TYPE_H = 4,
    GATE_TYPE_S = 5, GATE_TYPE_T = 6, GATE_TYPE_RX = 7, GATE_TYPE_RY = 8, GATE_TYPE_RZ = 9,
    GATE_TYPE_CNOT = 10, GATE_TYPE_CY = 11, GATE_TYPE_CZ = 12, GATE_TYPE_SWAP = 13,
    GATE_TYPE_CUSTOM = 14, GATE_TYPE_U1 = 15, GATE_TYPE_U2 = 16, GATE_TYPE_U3 = 17,
    GATE_TYPE_CCX = 18, GATE_TYPE_PHASE = 19, GATE_TYPE_CSWAP = 20, GATE_TYPE_ISWAP = 21,
    GATE_TYPE_CRX = 22, GATE_TYPE_CRY = 23, GATE_TYPE_CRZ = 24, GATE_TYPE_CH = 25,
    GATE_TYPE_SDG = 26, GATE_TYPE_TDG = 27, GATE_TYPE_ECR = 28, GATE_TYPE_SX = 29,
    GATE_TYPE_MEASURE = 30, GATE_TYPE_RESET = 31, GATE_TYPE_BARRIER = 32,
    GATE_TYPE_XX = 33, GATE_TYPE_YY = 34, GATE_TYPE_ZZ = 35
} gate_type_t;

/* include/quantum_geometric/core/quantum_base_types.h:127-141 */
typedef enum {
    HARDWARE_TYPE_CPU, HARDWARE_TYPE_GPU, HARDWARE_TYPE_QPU, HARDWARE_TYPE_SIMULATOR,
    HARDWARE_TYPE_METAL, HARDWARE_TYPE_CUDA, HARDWARE_TYPE_FPGA, HARDWARE_TYPE_IBM,
    HARDWARE_TYPE_RIGETTI, HARDWARE_TYPE_DWAVE, HARDWARE_TYPE_CUSTOM
} HardwareType;

/* include/quantum_geometric/core/quantum_base_types.h:167-175 */
typedef enum {
    QUANTUM_STATE_PURE, QUANTUM_STATE_MIXED, QUANTUM_STATE_THERMAL, QUANTUM_STATE_COHERENT,
    QUANTUM_STATE_SQUEEZED, QUANTUM_STATE_ENTANGLED, QUANTUM_STATE_CUSTOM
} quantum_state_type_t;

/* include/quantum_geometric/core/error_codes.h:17-202 (the codes this path returns) */
typedef enum {
    QGT_SUCCESS = 0,
    QGT_ERROR_INVALID_PARAMETER = -1,
    QGT_ERROR_INVALID_ARGUMENT = -1,
    QGT_ERROR_MEMORY_ALLOCATION = -2,
    QGT_ERROR_OUT_OF_MEMORY = -2,
    QGT_ERROR_ALLOCATION_FAILED = -2,
    QGT_ERROR_DIMENSION_MISMATCH = -3,
    QGT_ERROR_INVALID_STATE = -4,
    QGT_ERROR_VALIDATION_FAILED = -5,
    QGT_ERROR_HARDWARE_FAILURE = -6,
    QGT_ERROR_NOT_IMPLEMENTED = -7,
    QGT_ERROR_IO_ERROR = -16,
    QGT_ERROR_IO = -16,
    QGT_ERROR_INCOMPATIBLE = -19,
    QGT_ERROR_INVALID_DIMENSION = -22,
    QGT_ERROR_NO_MEMORY = -23,
    QGT_ERROR_NUMERICAL_INSTABILITY = -24,
    QGT_ERROR_MATRIX_SINGULAR = -55,
    QGT_ERROR_GPU_INVALID_VALUE = -118,
    QGT_ERROR_GPU_OUT_OF_MEMORY = -119,
    QGT_ERROR_GPU_INTERNAL = -120,
    QGT_ERROR_GPU_LAUNCH_FAILED = -121,
    QGT_ERROR_GPU_SYNC_FAILED = -122,
    QGT_ERROR_GPU_TRANSFER_FAILED = -123,
    QGT_ERROR_COMPUTATION_FAILED = -900,
    QGT_ERROR_CONVERGENCE_FAILED = -901,
    QGT_ERROR_NUMERICAL_ERROR = -902
} qgt_error_t;

/* include/quantum_geometric/core/quantum_types.h:170-192 */
typedef struct quantum_gate_t {
    gate_type_t type;
    size_t num_qubits;
    bool is_controlled;
    bool is_parameterized;
    size_t *target_qubits;
    size_t *control_qubits;
    size_t num_controls;
    size_t *qubits;
    double *parameters;
    size_t num_parameters;
    ComplexFloat *matrix;
    void *custom_data;
} quantum_gate_t;

/* include/quantum_geometric/core/quantum_types.h:234-238 */
typedef struct circuit_layer_t {
    quantum_gate_t **gates;
    size_t num_gates;
    bool is_parameterized;
} circuit_layer_t;

/* include/quantum_geometric/core/quantum_geometric_types.h:265-277 */
typedef struct quantum_geometric_state_t {
    int type;                      /* geometric_state_type_t */
    size_t dimension;
    size_t manifold_dim;
    size_t num_qubits;
    ComplexFloat *coordinates;
    ComplexFloat *metric;
    ComplexFloat *connection;
    double *error_rates;
    void *auxiliary_data;
    bool is_normalized;
    HardwareType hardware;
} quantum_geometric_state_t;

/* include/quantum_geometric/core/quantum_types.h:240-268 */
typedef struct quantum_circuit_t {
    size_t num_qubits;
    bool is_parameterized;
    circuit_layer_t **layers;
    size_t num_layers;
    size_t layers_capacity;
    quantum_gate_t **gates;
    size_t num_gates;
    size_t max_gates;
    int optimization_level;
    bool is_compiled;
    void *graph;                   /* computational_graph_t* */
    quantum_geometric_state_t *state;
    void **nodes;                  /* quantum_compute_node_t** */
    size_t num_nodes;
    size_t capacity;
} quantum_circuit_t;

/* include/quantum_geometric/core/quantum_geometric_tensor_network.h:15-43 */
typedef enum {
    QGTN_BACKEND_SIMULATOR, QGTN_BACKEND_IBM, QGTN_BACKEND_RIGETTI, QGTN_BACKEND_DWAVE,
    QGTN_BACKEND_IONQ, QGTN_BACKEND_XANADU, QGTN_BACKEND_CUSTOM
} qgtn_backend_type_t;

typedef struct {
    qgtn_backend_type_t type;
    void *backend_specific;
    bool supports_gradients;
    bool supports_hybrid;
} qgtn_hardware_config_t;

typedef struct quantum_geometric_tensor_network {
    void *network;                 /* tensor_network_t*: not used by this library (NULL) */
    quantum_circuit_t *circuit;
    size_t num_qubits;
    size_t num_layers;
    bool is_distributed;
    bool use_hardware_acceleration;
    qgtn_hardware_config_t hardware_config;
    void *backend_state;           /* this library's per-network cache */
} quantum_geometric_tensor_network_t;

/* include/quantum_geometric/core/quantum_types.h:195-231 */
typedef struct quantum_state_t {
    quantum_state_type_t type;
    size_t num_qubits;
    size_t dimension;
    size_t manifold_dim;
    ComplexFloat *coordinates;
    ComplexFloat *metric;
    ComplexFloat *connection;
    void *auxiliary_data;
    bool is_normalized;
    HardwareType hardware;
    size_t lattice_width;
    size_t lattice_height;
    size_t num_stabilizers;
    size_t num_plaquettes;
    size_t num_vertices;
    void *stabilizers;
    size_t num_anyons;
    size_t max_anyons;
    void *anyons;
    double *syndrome_values;
    size_t syndrome_size;
    double *measurement_confidence;
    size_t confidence_size;
} quantum_state_t;

/* include/quantum_geometric/hardware/hardware_capabilities.h:28-54 */
typedef struct {
    size_t size;
    size_t line_size;
    size_t associativity;
    bool inclusive;
} CacheConfig;

typedef struct {
    uint32_t feature_flags;
    size_t num_physical_cores;
    size_t num_logical_cores;
    size_t vector_register_width;
    size_t max_vector_elements;
    CacheConfig cache_config[3];
    bool has_fma;
    bool has_avx;
    bool has_avx2;
    bool has_avx512;
    bool has_neon;
    bool has_sve;
    bool has_amx;
    size_t page_size;
    size_t huge_page_size;
    bool supports_huge_pages;
} SystemCapabilities;

/* include/quantum_geometric/hardware/quantum_hardware_types.h:120-151 */
typedef struct QuantumHardwareCapabilities {
    bool supports_gpu;
    bool supports_metal;
    bool supports_cuda;
    bool supports_openmp;
    bool supports_mpi;
    bool supports_distributed;
    bool supports_feedback;
    bool supports_reset;
    bool supports_gates;
    bool supports_measurement;
    bool supports_annealing;
    bool supports_conditional;
    bool supports_parallel;
    bool supports_error_correction;
    uint32_t max_qubits;
    uint32_t max_gates;
    uint32_t max_depth;
    uint32_t max_shots;
    uint32_t max_parallel_jobs;
    size_t memory_size;
    double coherence_time;
    double gate_time;
    double readout_time;
    double *connectivity;
    size_t connectivity_size;
    size_t num_gates;
    char **available_gates;
    void *extensions;
    void *device_specific;
    void *backend_specific;
} QuantumHardwareCapabilities;

/* include/quantum_geometric/hardware/quantum_hardware_types.h:264-277 (passed by value to
 * quantum_state_to_device / _from_device / _is_on_device; this library reads `type`) */
typedef struct {
    HardwareType type;
    uint32_t num_qubits;
    uint32_t num_classical_bits;
    SystemCapabilities sys_caps;
    QuantumHardwareCapabilities caps;
    char device_name[256];
    void *device_data;
    uint32_t device_id;
    void *device_handle;
    void *context;
    void *command_queue;
} quantum_hardware_t;

/* include/quantum_geometric/core/quantum_geometric_types.h:232-250 */
typedef enum {
    GEOMETRIC_TENSOR_SCALAR, GEOMETRIC_TENSOR_VECTOR, GEOMETRIC_TENSOR_COVECTOR, GEOMETRIC_TENSOR_BIVECTOR,
    GEOMETRIC_TENSOR_TRIVECTOR, GEOMETRIC_TENSOR_UNITARY, GEOMETRIC_TENSOR_HERMITIAN, GEOMETRIC_TENSOR_SYMMETRIC,
    GEOMETRIC_TENSOR_CUSTOM
} geometric_tensor_type_t;

/* include/quantum_geometric/core/quantum_geometric_types.h:290-298 */
typedef enum { QGT_MEM_STANDARD = 0, QGT_MEM_HUGE_PAGES, QGT_MEM_PINNED, QGT_MEM_UNIFIED } qgt_memory_type_t;

/* include/quantum_geometric/core/quantum_geometric_types.h:331-353 */
typedef struct quantum_geometric_tensor_t {
    geometric_tensor_type_t type;
    size_t *dimensions;
    size_t rank;
    size_t dimension;
    ComplexFloat *components;
    size_t num_spins;
    void *auxiliary_data;
    bool is_symmetric;
    bool is_unitary;
    bool is_hermitian;
    HardwareType hardware;
    qgt_memory_type_t mem_type;
    size_t total_elements;
    size_t aligned_elements;
} quantum_geometric_tensor_t;

/* include/quantum_geometric/core/quantum_geometric_types.h:371-395 */
typedef enum {
    GEOMETRIC_METRIC_EUCLIDEAN, GEOMETRIC_METRIC_MINKOWSKI, GEOMETRIC_METRIC_FUBINI_STUDY,
    GEOMETRIC_METRIC_KAHLER, GEOMETRIC_METRIC_CUSTOM
} geometric_metric_type_t;

typedef enum {
    GEOMETRIC_CURVATURE_RIEMANN, GEOMETRIC_CURVATURE_RICCI, GEOMETRIC_CURVATURE_SCALAR,
    GEOMETRIC_CURVATURE_WEYL, GEOMETRIC_CURVATURE_BERRY, GEOMETRIC_CURVATURE_CUSTOM
} geometric_curvature_type_t;

/* include/quantum_geometric/core/quantum_geometric_types.h:433-459 */
typedef struct quantum_geometric_metric_t {
    geometric_metric_type_t type;
    size_t dimension;
    ComplexFloat *components;
    bool is_symmetric;
    void *auxiliary_data;
    HardwareType hardware;
} quantum_geometric_metric_t;

typedef struct quantum_geometric_curvature_t {
    geometric_curvature_type_t type;
    size_t dimension;
    ComplexFloat *components;
    void *auxiliary_data;
    bool is_flat;
    HardwareType hardware;
} quantum_geometric_curvature_t;

/* include/quantum_geometric/core/quantum_geometric_gradient.h:188-201 */
typedef struct {
    float regularization_param;
    float condition_threshold;
    bool use_adaptive_regularization;
    bool use_pseudoinverse_fallback;
    float singular_value_cutoff;
} natural_gradient_config_t;

#ifdef __cplusplus
}
#endif
#endif
