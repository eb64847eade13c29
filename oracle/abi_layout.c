/*
 * ABI layout probe (test infrastructure).  Prints sizeof/offsetof of every
 * struct that crosses the QGT path's C API, and the enum values it uses.
 * Built twice: against QGTL's headers (oracle/Makefile target ref-abi, output
 * oracle/_ref/abi_layout_ref) and against this library's include/ (tests);
 * tests/test_abi.py requires identical output.
 */
#include <stddef.h>
#include <stdio.h>

#ifdef QGTL_REFERENCE_HEADERS
#include "quantum_geometric/core/quantum_geometric_tensor_network.h"
#include "quantum_geometric/core/quantum_geometric_metric.h"
#include "quantum_geometric/core/quantum_geometric_curvature.h"
#include "quantum_geometric/core/quantum_geometric_gradient.h"
#include "quantum_geometric/core/quantum_state.h"
#include "quantum_geometric/core/quantum_geometric_types.h"
#include "quantum_geometric/hardware/quantum_hardware_types.h"
#include "quantum_geometric/supercomputer/compute_backend.h"
#else
#include "qgtl_api.h"
#include "qgtl_compute_backend.h"
#endif

#define SZ(T) printf("sizeof %s %zu\n", #T, sizeof(T))
#define OFF(T, f) printf("offsetof %s.%s %zu\n", #T, #f, offsetof(T, f))
#define VAL(e) printf("value %s %d\n", #e, (int)(e))

int main(void)
{
    SZ(ComplexFloat); SZ(ComplexDouble);
    SZ(quantum_gate_t);
    OFF(quantum_gate_t, type); OFF(quantum_gate_t, num_qubits); OFF(quantum_gate_t, is_controlled);
    OFF(quantum_gate_t, is_parameterized); OFF(quantum_gate_t, target_qubits); OFF(quantum_gate_t, control_qubits);
    OFF(quantum_gate_t, num_controls); OFF(quantum_gate_t, qubits); OFF(quantum_gate_t, parameters);
    OFF(quantum_gate_t, num_parameters); OFF(quantum_gate_t, matrix); OFF(quantum_gate_t, custom_data);
    SZ(circuit_layer_t);
    OFF(circuit_layer_t, gates); OFF(circuit_layer_t, num_gates); OFF(circuit_layer_t, is_parameterized);
    SZ(quantum_circuit_t);
    OFF(quantum_circuit_t, num_qubits); OFF(quantum_circuit_t, is_parameterized); OFF(quantum_circuit_t, layers);
    OFF(quantum_circuit_t, num_layers); OFF(quantum_circuit_t, layers_capacity); OFF(quantum_circuit_t, gates);
    OFF(quantum_circuit_t, num_gates); OFF(quantum_circuit_t, max_gates); OFF(quantum_circuit_t, optimization_level);
    OFF(quantum_circuit_t, is_compiled); OFF(quantum_circuit_t, graph); OFF(quantum_circuit_t, state);
    OFF(quantum_circuit_t, nodes); OFF(quantum_circuit_t, num_nodes); OFF(quantum_circuit_t, capacity);
    SZ(quantum_geometric_state_t);
    OFF(quantum_geometric_state_t, dimension); OFF(quantum_geometric_state_t, coordinates);
    OFF(quantum_geometric_state_t, is_normalized); OFF(quantum_geometric_state_t, hardware);
    SZ(quantum_geometric_tensor_network_t);
    OFF(quantum_geometric_tensor_network_t, network); OFF(quantum_geometric_tensor_network_t, circuit);
    OFF(quantum_geometric_tensor_network_t, num_qubits); OFF(quantum_geometric_tensor_network_t, num_layers);
    OFF(quantum_geometric_tensor_network_t, is_distributed);
    OFF(quantum_geometric_tensor_network_t, use_hardware_acceleration);
    OFF(quantum_geometric_tensor_network_t, hardware_config); OFF(quantum_geometric_tensor_network_t, backend_state);
    SZ(quantum_state_t);
    OFF(quantum_state_t, type); OFF(quantum_state_t, num_qubits); OFF(quantum_state_t, dimension);
    OFF(quantum_state_t, coordinates); OFF(quantum_state_t, auxiliary_data); OFF(quantum_state_t, is_normalized);
    OFF(quantum_state_t, hardware); OFF(quantum_state_t, lattice_width); OFF(quantum_state_t, stabilizers);
    OFF(quantum_state_t, anyons); OFF(quantum_state_t, measurement_confidence); OFF(quantum_state_t, confidence_size);
    SZ(quantum_geometric_metric_t);
    OFF(quantum_geometric_metric_t, type); OFF(quantum_geometric_metric_t, dimension);
    OFF(quantum_geometric_metric_t, components); OFF(quantum_geometric_metric_t, is_symmetric);
    OFF(quantum_geometric_metric_t, auxiliary_data); OFF(quantum_geometric_metric_t, hardware);
    SZ(quantum_geometric_curvature_t);
    OFF(quantum_geometric_curvature_t, type); OFF(quantum_geometric_curvature_t, dimension);
    OFF(quantum_geometric_curvature_t, components); OFF(quantum_geometric_curvature_t, auxiliary_data);
    OFF(quantum_geometric_curvature_t, is_flat); OFF(quantum_geometric_curvature_t, hardware);
    SZ(natural_gradient_config_t);
    OFF(natural_gradient_config_t, regularization_param); OFF(natural_gradient_config_t, condition_threshold);
    OFF(natural_gradient_config_t, use_adaptive_regularization);
    OFF(natural_gradient_config_t, use_pseudoinverse_fallback); OFF(natural_gradient_config_t, singular_value_cutoff);
    SZ(CacheConfig); SZ(SystemCapabilities); SZ(QuantumHardwareCapabilities);
    OFF(SystemCapabilities, cache_config); OFF(SystemCapabilities, has_fma); OFF(SystemCapabilities, page_size);
    OFF(QuantumHardwareCapabilities, max_qubits); OFF(QuantumHardwareCapabilities, memory_size);
    OFF(QuantumHardwareCapabilities, backend_specific);
    SZ(quantum_hardware_t);
    OFF(quantum_hardware_t, type); OFF(quantum_hardware_t, sys_caps); OFF(quantum_hardware_t, caps);
    OFF(quantum_hardware_t, device_name); OFF(quantum_hardware_t, device_data); OFF(quantum_hardware_t, device_id);
    OFF(quantum_hardware_t, command_queue);
    SZ(quantum_geometric_tensor_t);
    OFF(quantum_geometric_tensor_t, type); OFF(quantum_geometric_tensor_t, dimensions);
    OFF(quantum_geometric_tensor_t, rank); OFF(quantum_geometric_tensor_t, dimension);
    OFF(quantum_geometric_tensor_t, components); OFF(quantum_geometric_tensor_t, num_spins);
    OFF(quantum_geometric_tensor_t, auxiliary_data); OFF(quantum_geometric_tensor_t, is_symmetric);
    OFF(quantum_geometric_tensor_t, is_hermitian); OFF(quantum_geometric_tensor_t, hardware);
    OFF(quantum_geometric_tensor_t, mem_type); OFF(quantum_geometric_tensor_t, total_elements);
    OFF(quantum_geometric_tensor_t, aligned_elements);
    SZ(ComputeBackendOps);
    OFF(ComputeBackendOps, init); OFF(ComputeBackendOps, memcpy); OFF(ComputeBackendOps, wait_event);
    OFF(ComputeBackendOps, quantum_unitary); OFF(ComputeBackendOps, quantum_normalize);
    OFF(ComputeBackendOps, quantum_tensor_contract); OFF(ComputeBackendOps, quantum_gradient);
    OFF(ComputeBackendOps, quantum_inner_product); OFF(ComputeBackendOps, quantum_expectation);
    OFF(ComputeBackendOps, barrier); OFF(ComputeBackendOps, reduce_scatter); OFF(ComputeBackendOps, execute);
    OFF(ComputeBackendOps, reset_metrics);
    SZ(ComputeBackendInfo); OFF(ComputeBackendInfo, name); OFF(ComputeBackendInfo, priority); OFF(ComputeBackendInfo, ops);
    SZ(ComputeQuantumOp); OFF(ComputeQuantumOp, output_data); OFF(ComputeQuantumOp, parameters);
    OFF(ComputeQuantumOp, param_size); OFF(ComputeQuantumOp, dims); OFF(ComputeQuantumOp, num_dims);
    OFF(ComputeQuantumOp, num_parameters);
    SZ(ComputeExecutionPlan); OFF(ComputeExecutionPlan, workspace); OFF(ComputeExecutionPlan, requires_sync);
    SZ(ComputeMetrics); OFF(ComputeMetrics, flops); OFF(ComputeMetrics, current_memory_bytes);
    OFF(ComputeMetrics, error_rate);
    SZ(ComputeDistributedConfig); OFF(ComputeDistributedConfig, preferred_backend);
    OFF(ComputeDistributedConfig, backend_config);
    VAL(COMPUTE_BACKEND_CUDA); VAL(COMPUTE_MEM_DEVICE); VAL(COMPUTE_MEM_PINNED); VAL(COMPUTE_DTYPE_COMPLEX128);
    VAL(COMPUTE_ERROR_NOT_IMPLEMENTED); VAL(COMPUTE_ERROR_INTERNAL); VAL(QUANTUM_OP_EXPECTATION);
    VAL(QUANTUM_OP_DENSITY_MATRIX); VAL(COMPUTE_REDUCE_AVG); VAL(GEOMETRIC_TENSOR_HERMITIAN); VAL(QGT_MEM_UNIFIED);
    VAL(HARDWARE_TYPE_QPU);
    VAL(GATE_TYPE_X); VAL(GATE_TYPE_Y); VAL(GATE_TYPE_Z); VAL(GATE_TYPE_H); VAL(GATE_TYPE_S); VAL(GATE_TYPE_T);
    VAL(GATE_TYPE_RX); VAL(GATE_TYPE_RY); VAL(GATE_TYPE_RZ); VAL(GATE_TYPE_CNOT); VAL(GATE_TYPE_CY);
    VAL(GATE_TYPE_CZ); VAL(GATE_TYPE_SWAP); VAL(GATE_TYPE_U1); VAL(GATE_TYPE_PHASE); VAL(GATE_TYPE_SDG);
    VAL(GATE_TYPE_TDG); VAL(GATE_TYPE_SX);
    VAL(QGT_SUCCESS); VAL(QGT_ERROR_INVALID_PARAMETER); VAL(QGT_ERROR_INVALID_ARGUMENT); VAL(QGT_ERROR_DIMENSION_MISMATCH);
    VAL(QGT_ERROR_ALLOCATION_FAILED); VAL(QGT_ERROR_INVALID_STATE); VAL(QGT_ERROR_INCOMPATIBLE);
    VAL(QGT_ERROR_COMPUTATION_FAILED); VAL(QGT_ERROR_NOT_IMPLEMENTED); VAL(QGT_ERROR_NO_MEMORY);
    VAL(QGT_ERROR_MATRIX_SINGULAR); VAL(QGT_ERROR_CONVERGENCE_FAILED); VAL(QGT_ERROR_GPU_OUT_OF_MEMORY);
    VAL(GEOMETRIC_METRIC_EUCLIDEAN); VAL(GEOMETRIC_METRIC_MINKOWSKI); VAL(GEOMETRIC_METRIC_FUBINI_STUDY);
    VAL(GEOMETRIC_CURVATURE_BERRY); VAL(HARDWARE_TYPE_CPU); VAL(HARDWARE_TYPE_GPU); VAL(QUANTUM_STATE_PURE);
    VAL(QGTN_BACKEND_SIMULATOR);
    return 0;
}
