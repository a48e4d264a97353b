"""kettle_amd: kettle's simulate() hot path on MI355X (gfx950) through libkq's C-ABI.

Importing the package does not touch the GPU; the first call that needs libkq loads
``kettle_amd/lib/libkq.so`` and fails loudly if it has not been built.
"""
from ._lib import KQ_C128, KQ_C64, KqError
from .circuit import (ControlFlowBooleanKind, ControlFlowPredicate, Gate, Matrix2X2, ParameterID, QuantumCircuit,
                      extend_circuit)
from .api import (DensityMatrix, EnergyGradient, FusionPlan, PackedPauliOperator, PauliOperator, PauliPhase,
                  PauliTerm, QuantumNoise, debug_jit_compile, debug_jit_source, debug_plan, jit_info, set_jit,
                  set_jit_tuning, set_jit_priority, set_jit_nontemporal,
                  SimulationResult, SparsePauliString, Statevector, apply_gates, calculate_probabilities_raw,
                  energy_and_gradient, expectation_value, partial_trace, perform_measurements_as_counts,
                  perform_measurements_as_counts_marginal, perform_measurements_as_counts_raw,
                  perform_measurements_as_memory, perform_measurements_as_memory_circuit, plan_stats, set_fusion, set_fusion_autotune, fusion_autotune_info, set_fusion_low_qubits, set_fusion_relabel, simulate, simulate_pauli_string,
                  state_index_to_bitstring, statevector_to_density_matrix)
from .io import (Endian, load_statevector, read_numpy_statevector, read_pauli_operator, read_tangelo_circuit,
                 save_statevector, write_tangelo_circuit)

__all__ = [name for name in dir() if not name.startswith("_")]
