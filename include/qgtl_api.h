/*
 * qgtl_api.h -- QGTL's C API for the quantum-geometric-tensor path, exported
 * by libqgtl_mi355x.so with QGTL's names, signatures, argument meaning and
 * error behaviour.  Each declaration cites the QGTL declaration it replaces
 * (paths relative to the QGTL source root).
 *
 * Behaviour notes that differ from QGTL on purpose:
 *  - Derivative states are exact (generator insertion on the GPU), not the
 *    shifted differences of compute_higher_order_gradient.
 *  - Every element call reads one full P x P QGT computed once per circuit
 *    content and cached on the network (backend_state); the element-wise
 *    loops of the metric/curvature/full-QGT wrappers therefore cost one
 *    GPU QGT, not P(P+1)/2 of them.
 *  - geometric_create_metric/curvature accept dimensions up to
 *    QGTL_MAX_PARAMS (QGTL caps them at QGT_MAX_DIMENSIONS = 16, which rules
 *    out any real circuit).
 */
#ifndef QGTL_API_H
#define QGTL_API_H

#include "qgtl_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define QGTL_MAX_PARAMS 4096

/* ---- gates: include/quantum_geometric/core/quantum_gate_operations.h ---- */
/* :19-24 */
quantum_gate_t *create_quantum_gate(gate_type_t type, const size_t *qubits, size_t num_qubits,
                                    const double *parameters, size_t num_parameters);
/* :32 */
quantum_gate_t *copy_quantum_gate(const quantum_gate_t *gate);
/* :41-44 */
bool update_gate_parameters(quantum_gate_t *gate, const double *parameters, size_t num_parameters);
/* quantum_gate_operations.c:430-438 */
void destroy_quantum_gate(quantum_gate_t *gate);

/* ---- network: include/quantum_geometric/core/quantum_geometric_tensor_network.h ---- */
/* :46-51 */
quantum_geometric_tensor_network_t *create_quantum_geometric_tensor_network(
    size_t num_qubits, size_t num_layers, bool is_distributed, bool use_hardware_acceleration);
/* :53-55 */
void destroy_quantum_geometric_tensor_network(quantum_geometric_tensor_network_t *qgtn);
/* :82-87: appends a copy of the gate to the last layer */
bool apply_quantum_gate(quantum_geometric_tensor_network_t *qgtn, const quantum_gate_t *gate,
                        const size_t *qubits, size_t num_qubits);
/* :89-92 */
bool apply_quantum_circuit(quantum_geometric_tensor_network_t *qgtn, const quantum_circuit_t *circuit);
/* :101-105: *state_vector is malloc'd; caller frees with free() */
bool get_quantum_state(const quantum_geometric_tensor_network_t *qgtn, ComplexFloat **state_vector,
                       size_t *dimension);
/* :108-113 */
bool compute_quantum_geometric_tensor(const quantum_geometric_tensor_network_t *qgtn, size_t param_i,
                                      size_t param_j, ComplexFloat *result);
/* :115-120 */
bool compute_quantum_metric(const quantum_geometric_tensor_network_t *qgtn, size_t param_i,
                            size_t param_j, double *result);
/* :122-127 */
bool compute_berry_curvature(const quantum_geometric_tensor_network_t *qgtn, size_t param_i,
                             size_t param_j, double *result);
/* :151 */
const char *get_quantum_geometric_tensor_network_error(void);

/* ---- metric: include/quantum_geometric/core/quantum_geometric_metric.h ---- */
/* :13-15 */
qgt_error_t geometric_create_metric(quantum_geometric_metric_t **metric, geometric_metric_type_t type,
                                    size_t dimension);
/* :18 */
void geometric_destroy_metric(quantum_geometric_metric_t *metric);
/* :48-51 */
qgt_error_t geometric_compute_fubini_study_metric(quantum_geometric_metric_t *metric,
                                                  const quantum_geometric_tensor_network_t *qgtn,
                                                  size_t num_params);
/* :64-68 */
qgt_error_t geometric_compute_fubini_study_element(const quantum_geometric_tensor_network_t *qgtn,
                                                   size_t param_mu, size_t param_nu, float *result);

/* ---- curvature: include/quantum_geometric/core/quantum_geometric_curvature.h ---- */
/* :15-17 */
qgt_error_t geometric_create_curvature(quantum_geometric_curvature_t **curvature,
                                       geometric_curvature_type_t type, size_t dimension);
/* :20 */
void geometric_destroy_curvature(quantum_geometric_curvature_t *curvature);
/* :49-52 */
qgt_error_t geometric_compute_berry_curvature(quantum_geometric_curvature_t *curvature,
                                              const quantum_geometric_tensor_network_t *qgtn,
                                              size_t num_params);
/* :65-69 */
qgt_error_t geometric_compute_berry_curvature_element(const quantum_geometric_tensor_network_t *qgtn,
                                                      size_t param_mu, size_t param_nu, float *result);
/* :83-87 */
qgt_error_t geometric_compose_qgt(ComplexFloat *qgt, const quantum_geometric_metric_t *metric,
                                  const quantum_geometric_curvature_t *curvature, size_t dimension);
/* :100-103 */
qgt_error_t geometric_compute_full_qgt(ComplexFloat *qgt, const quantum_geometric_tensor_network_t *qgtn,
                                       size_t num_params);

/* ---- natural gradient: include/quantum_geometric/core/quantum_geometric_gradient.h ---- */
/* :208 */
natural_gradient_config_t get_default_natural_gradient_config(void);
/* :220-223 */
bool compute_complex_matrix_condition_number(const ComplexFloat *matrix, size_t dimension,
                                             float *condition_number);
/* :239-244 */
bool compute_regularized_natural_gradient(const ComplexFloat *gradient, const ComplexFloat *metric,
                                          ComplexFloat *natural_gradient, size_t dimension,
                                          const natural_gradient_config_t *config);
/* :261-266 */
bool compute_pseudoinverse_natural_gradient(const ComplexFloat *gradient, const ComplexFloat *metric,
                                            ComplexFloat *natural_gradient, size_t dimension,
                                            float singular_cutoff);

/* ---- state: include/quantum_geometric/core/quantum_state.h ---- */
/* :15-17 */
qgt_error_t quantum_state_create(quantum_state_t **state, quantum_state_type_t type, size_t dimension);
/* :18 */
void quantum_state_destroy(quantum_state_t *state);
/* :19-20 */
qgt_error_t quantum_state_clone(quantum_state_t **dest, const quantum_state_t *src);
/* :23-24 */
qgt_error_t quantum_state_initialize(quantum_state_t *state, const ComplexFloat *amplitudes);
/* :25 */
qgt_error_t quantum_state_initialize_zero(quantum_state_t *state);
/* :26-27 */
qgt_error_t quantum_state_initialize_basis(quantum_state_t *state, size_t basis_index);
/* :58-60 */
qgt_error_t quantum_state_inner_product(ComplexFloat *result, const quantum_state_t *a,
                                        const quantum_state_t *b);
/* :61-62 */
qgt_error_t quantum_state_norm(float *norm, const quantum_state_t *state);
/* :63 */
qgt_error_t quantum_state_normalize(quantum_state_t *state);
/* :68-70 */
qgt_error_t quantum_state_fidelity(float *fidelity, const quantum_state_t *a, const quantum_state_t *b);
/* :86 */
bool quantum_state_is_normalized(const quantum_state_t *state);
/* :93-94 (declared, never defined in QGTL): a complex64 copy in device memory for
 * hardware.type == HARDWARE_TYPE_GPU (HARDWARE_TYPE_CPU moves it back; other types
 * QGT_ERROR_NOT_IMPLEMENTED); while it is there the device copy is authoritative */
qgt_error_t quantum_state_to_device(quantum_state_t *state, quantum_hardware_t hardware);
/* :95-96: copies the device values into coordinates and frees the device copy */
qgt_error_t quantum_state_from_device(quantum_state_t *state, quantum_hardware_t hardware);
/* :97-98 */
bool quantum_state_is_on_device(const quantum_state_t *state, quantum_hardware_t hardware);

/* ---- files: include/quantum_geometric/core/quantum_geometric_tensor.h ---- */
/* :144-145 (declared, never defined in QGTL): type, rank, dimensions, flags and the
 * total_elements components in the library's checksummed container (magic QGTLTEN1) */
qgt_error_t geometric_tensor_save(const quantum_geometric_tensor_t *tensor, const char *filename);
/* :146-147: a new tensor allocated the way QGTL's geometric_tensor_create allocates one
 * (aligned_alloc(64) struct and components, malloc'd dimensions), so QGTL's
 * geometric_tensor_destroy frees it */
qgt_error_t geometric_tensor_load(quantum_geometric_tensor_t **tensor, const char *filename);

#ifdef __cplusplus
}
#endif
#endif
