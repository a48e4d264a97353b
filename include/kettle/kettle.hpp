#pragma once

// Umbrella header: kettle's simulate() hot-path API on MI355X.
#include <kettle/calculations/measurements.hpp>
#include <kettle/calculations/probabilities.hpp>
#include <kettle/circuit/circuit.hpp>
#include <kettle/circuit/classical_register.hpp>
#include <kettle/circuit/control_flow_predicate.hpp>
#include <kettle/circuit_loggers/circuit_logger.hpp>
#include <kettle/common/matrix2x2.hpp>
#include <kettle/gates/common_u_gates.hpp>
#include <kettle/gates/primitive_gate.hpp>
#include <kettle/gates/random_u_gates.hpp>
#include <kettle/gradient/energy_gradient.hpp>
#include <kettle/io/read_pauli_operator.hpp>
#include <kettle/io/read_tangelo_file.hpp>
#include <kettle/io/write_tangelo_file.hpp>
#include <kettle/operator/pauli/pauli_operator.hpp>
#include <kettle/operator/pauli/sparse_pauli_string.hpp>
#include <kettle/optimize/n_local.hpp>
#include <kettle/parameter/parameter.hpp>
#include <kettle/simulation/simulate.hpp>
#include <kettle/simulation/simulate_density_matrix.hpp>
#include <kettle/simulation/simulate_pauli.hpp>
#include <kettle/state/density_matrix.hpp>
#include <kettle/state/endian.hpp>
#include <kettle/state/qubit_state_conversion.hpp>
#include <kettle/state/statevector.hpp>
