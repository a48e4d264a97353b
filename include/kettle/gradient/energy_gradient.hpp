#pragma once

#include <cstddef>
#include <vector>

#include "kettle/circuit/circuit.hpp"
#include "kettle/operator/pauli/pauli_operator.hpp"
#include "kettle/parameter/parameter.hpp"
#include "kettle/state/statevector.hpp"

/*
    Energy and its gradient for a parameterised circuit (new: kettle's VQE drivers are
    derivative-free).  E(theta) = Re <psi(theta)| H |psi(theta)>, psi(theta) = U(theta)|0...0>,
    and dE/dtheta_p for every id in `parameter_ids`, by one forward and one reverse (adjoint)
    sweep on the GPU.  Precision::C64 keeps the states in complex<float> and every reduction in
    FP64.  The circuit must be straight-line (no measurement, control flow or loggers).
*/
namespace ket
{

struct EnergyGradient
{
    double energy;
    std::vector<double> gradient;
    std::size_t checkpoint_restores;
};

struct AdjointOptions
{
    Precision precision = Precision::C64;
    std::size_t checkpoint_every = 0;   // 0: sqrt(number of gates)
    double drift_tolerance = 1.0e-5;    // max-abs drift before a checkpoint is restored
};

auto energy_and_gradient(const QuantumCircuit& circuit, const std::vector<ket::param::ParameterID>& parameter_ids,
                         const PauliOperator& hamiltonian, const AdjointOptions& options = {}) -> EnergyGradient;

}  // namespace ket
