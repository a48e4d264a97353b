#pragma once

#include <optional>
#include <vector>

#include "kettle/circuit/circuit.hpp"
#include "kettle/circuit/classical_register.hpp"
#include "kettle/state/density_matrix.hpp"

// ket::simulate for density matrices (simulate_density_matrix.hpp:18-46 in the reference).
namespace ket
{

class DensityMatrixSimulator
{
public:
    explicit DensityMatrixSimulator(std::size_t n_qubits);
    void run(const QuantumCircuit& circuit, DensityMatrix& state, std::optional<int> prng_seed = std::nullopt);
    [[nodiscard]] auto has_been_run() const -> bool;
    [[nodiscard]] auto classical_register() const -> const ClassicalRegister&;

private:
    std::size_t n_qubits_;
    std::optional<ClassicalRegister> cregister_;
    bool has_been_run_ {false};
};

void simulate(const QuantumCircuit& circuit, DensityMatrix& state, std::optional<int> prng_seed = std::nullopt);

}  // namespace ket
