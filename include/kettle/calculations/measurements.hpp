#pragma once

#include <cstddef>
#include <map>
#include <optional>
#include <string>
#include <vector>

#include "kettle/calculations/probabilities.hpp"
#include "kettle/circuit/circuit.hpp"
#include "kettle/state/statevector.hpp"

// Sampling (include/kettle/calculations/measurements.hpp:42-97 in the reference).  The CDF and
// the per-shot lower_bound run on the GPU; the uniform draws are kettle's.
namespace ket
{

auto memory_to_counts(const std::vector<std::size_t>& measurements) -> std::map<std::size_t, std::size_t>;
auto memory_to_fractions(const std::vector<std::size_t>& measurements) -> std::map<std::size_t, double>;

auto perform_measurements_as_memory(const std::vector<double>& probabilities_raw, std::size_t n_shots,
                                    std::optional<int> seed = std::nullopt) -> std::vector<std::size_t>;
auto perform_measurements_as_memory(const Statevector& state, std::size_t n_shots, const QuantumNoise* noise = nullptr,
                                    std::optional<int> seed = std::nullopt) -> std::vector<std::size_t>;
auto perform_measurements_as_counts_raw(const std::vector<double>& probabilities_raw, std::size_t n_shots,
                                        std::optional<int> seed = std::nullopt) -> std::map<std::size_t, std::size_t>;
auto perform_measurements_as_counts_raw(const Statevector& state, std::size_t n_shots, const QuantumNoise* noise = nullptr,
                                        std::optional<int> seed = std::nullopt) -> std::map<std::size_t, std::size_t>;
auto perform_measurements_as_counts_marginal(const std::vector<double>& probabilities_raw, std::size_t n_shots,
                                             const std::vector<std::size_t>& marginal_qubits = {},
                                             std::optional<int> seed = std::nullopt) -> std::map<std::string, std::size_t>;
auto perform_measurements_as_counts_marginal(const Statevector& state, std::size_t n_shots,
                                             const std::vector<std::size_t>& marginal_qubits = {},
                                             const QuantumNoise* noise = nullptr,
                                             std::optional<int> seed = std::nullopt) -> std::map<std::string, std::size_t>;
// re-simulates the circuit from `original_state` for every shot (measurements.cpp:178-208)
auto perform_measurements_as_counts_marginal(const QuantumCircuit& circuit, const Statevector& original_state,
                                             std::size_t n_shots, const std::vector<std::size_t>& marginal_qubits = {},
                                             const QuantumNoise* noise = nullptr,
                                             std::optional<int> seed = std::nullopt) -> std::map<std::string, std::size_t>;
auto perform_measurements_as_counts(const Statevector& state, std::size_t n_shots, const QuantumNoise* noise = nullptr,
                                    std::optional<int> seed = std::nullopt) -> std::map<std::string, std::size_t>;

}  // namespace ket
