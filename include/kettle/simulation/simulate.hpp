#pragma once

#include <optional>
#include <vector>

#include "kettle/circuit/circuit.hpp"
#include "kettle/circuit/classical_register.hpp"
#include "kettle/circuit_loggers/circuit_logger.hpp"
#include "kettle/state/statevector.hpp"

// ket::simulate for statevectors (include/kettle/simulation/simulate.hpp:16-41 in the reference),
// executed on the GPU through libkq.
namespace ket
{

class StatevectorSimulator
{
public:
    void run(const QuantumCircuit& circuit, Statevector& state, std::optional<int> prng_seed = std::nullopt);
    [[nodiscard]] auto has_been_run() const -> bool;
    [[nodiscard]] auto classical_register() const -> const ClassicalRegister&;
    auto classical_register() -> ClassicalRegister&;
    [[nodiscard]] auto circuit_loggers() const -> const std::vector<CircuitLogger>&;

private:
    std::optional<ClassicalRegister> cregister_;
    bool has_been_run_ {false};
    std::vector<CircuitLogger> circuit_loggers_;
};

void simulate(const QuantumCircuit& circuit, Statevector& state, std::optional<int> prng_seed = std::nullopt);

}  // namespace ket
