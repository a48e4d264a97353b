#pragma once

#include <cstddef>
#include <map>
#include <string>
#include <vector>

#include "kettle/state/density_matrix.hpp"
#include "kettle/state/statevector.hpp"

// Probabilities with optional bit-flip noise (include/kettle/calculations/probabilities.hpp).
namespace ket
{

class QuantumNoise
{
public:
    explicit QuantumNoise(std::size_t n_qubits);
    void set(std::size_t index, double noise);
    [[nodiscard]] auto get(std::size_t index) const -> const double&;
    [[nodiscard]] auto n_qubits() const -> std::size_t { return n_qubits_; }
    [[nodiscard]] auto values() const -> const std::vector<double>& { return noise_; }

private:
    std::size_t n_qubits_;
    std::vector<double> noise_;
    void check_index_(std::size_t index) const;
};

auto calculate_probabilities_raw(const Statevector& state, const QuantumNoise* noise = nullptr) -> std::vector<double>;
// diag(rho); kettle has no density-matrix overload, this one feeds the same sampler
auto calculate_probabilities_raw(const DensityMatrix& state, const QuantumNoise* noise = nullptr) -> std::vector<double>;
auto calculate_probabilities(const Statevector& state, const QuantumNoise* noise = nullptr) -> std::map<std::string, double>;

}  // namespace ket
