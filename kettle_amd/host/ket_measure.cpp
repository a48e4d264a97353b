// Probabilities and sampling (calculations/probabilities.cpp, calculations/measurements.cpp) over
// kq_probabilities / kq_sample / kq_sample_probabilities.
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "kettle/calculations/measurements.hpp"
#include "kettle/calculations/probabilities.hpp"
#include "kettle/simulation/simulate.hpp"

#include "ket_internal.hpp"

namespace ket
{

namespace ki = ket::internal;

QuantumNoise::QuantumNoise(std::size_t n_qubits) : n_qubits_ {n_qubits}, noise_(n_qubits, 0.0) {}

void QuantumNoise::check_index_(std::size_t index) const
{
    if (index >= n_qubits_) {
        throw std::runtime_error {"ERROR: Out-of-bounds access for QuantumNoise probability."};
    }
}

void QuantumNoise::set(std::size_t index, double noise)
{
    check_index_(index);
    if (!(0.0 <= noise && noise <= 1.0)) {
        throw std::runtime_error {"ERROR: Cannot set probability in QuantumNoise object outside of [0, 1]."};
    }
    noise_[index] = noise;
}

auto QuantumNoise::get(std::size_t index) const -> const double&
{
    check_index_(index);
    return noise_[index];
}

namespace
{

auto noise_ptr(const QuantumNoise* noise, std::size_t n_qubits) -> const double*
{
    if (noise == nullptr) {
        return nullptr;
    }
    if (noise->n_qubits() != n_qubits) {
        throw std::runtime_error {"ERROR: Out-of-bounds access for QuantumNoise probability."};
    }
    return noise->values().data();
}

auto seed_args(std::optional<int> seed) -> std::pair<int, int32_t> { return {seed ? 1 : 0, seed.value_or(0)}; }

auto is_power_of_2(std::size_t v) -> bool { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace

auto calculate_probabilities_raw(const Statevector& state, const QuantumNoise* noise) -> std::vector<double>
{
    std::vector<double> out(state.n_states());
    ki::kq_check(kq_probabilities(state.device_state(), noise_ptr(noise, state.n_qubits()), out.data()));
    return out;
}

auto calculate_probabilities_raw(const DensityMatrix& state, const QuantumNoise* noise) -> std::vector<double>
{
    std::vector<double> out(state.n_states());
    ki::kq_check(kq_probabilities(state.device_state(), noise_ptr(noise, state.n_qubits()), out.data()));
    return out;
}

auto calculate_probabilities(const Statevector& state, const QuantumNoise* noise) -> std::map<std::string, double>
{
    const auto raw = calculate_probabilities_raw(state, noise);
    auto out = std::map<std::string, double> {};
    for (std::size_t i = 0; i < raw.size(); ++i) {
        out[state_index_to_bitstring(i, state.n_qubits(), Endian::LITTLE)] = raw[i];
    }
    return out;
}

auto memory_to_counts(const std::vector<std::size_t>& measurements) -> std::map<std::size_t, std::size_t>
{
    auto map = std::map<std::size_t, std::size_t> {};
    for (auto i : measurements) {
        ++map[i];
    }
    return map;
}

auto memory_to_fractions(const std::vector<std::size_t>& measurements) -> std::map<std::size_t, double>
{
    auto map = std::map<std::size_t, double> {};
    for (auto i : measurements) {
        map[i] += 1.0;
    }
    const auto n = static_cast<double>(measurements.size());
    for (auto& [k, v] : map) {
        v /= n;
    }
    return map;
}

auto perform_measurements_as_memory(const std::vector<double>& probabilities_raw, std::size_t n_shots,
                                    std::optional<int> seed) -> std::vector<std::size_t>
{
    std::vector<uint64_t> idx(n_shots);
    const auto [hs, sv] = seed_args(seed);
    ki::kq_check(kq_sample_probabilities(probabilities_raw.data(), probabilities_raw.size(), n_shots, hs, sv, idx.data()));
    return {idx.begin(), idx.end()};
}

auto perform_measurements_as_memory(const Statevector& state, std::size_t n_shots, const QuantumNoise* noise,
                                    std::optional<int> seed) -> std::vector<std::size_t>
{
    std::vector<uint64_t> idx(n_shots);
    const auto [hs, sv] = seed_args(seed);
    ki::kq_check(kq_sample(state.device_state(), noise_ptr(noise, state.n_qubits()), n_shots, hs, sv, idx.data()));
    return {idx.begin(), idx.end()};
}

auto perform_measurements_as_counts_raw(const std::vector<double>& probabilities_raw, std::size_t n_shots,
                                        std::optional<int> seed) -> std::map<std::size_t, std::size_t>
{
    return memory_to_counts(perform_measurements_as_memory(probabilities_raw, n_shots, seed));
}

auto perform_measurements_as_counts_raw(const Statevector& state, std::size_t n_shots, const QuantumNoise* noise,
                                        std::optional<int> seed) -> std::map<std::size_t, std::size_t>
{
    return memory_to_counts(perform_measurements_as_memory(state, n_shots, noise, seed));
}

auto perform_measurements_as_counts_marginal(const std::vector<double>& probabilities_raw, std::size_t n_shots,
                                             const std::vector<std::size_t>& marginal_qubits, std::optional<int> seed)
    -> std::map<std::string, std::size_t>
{
    if (!is_power_of_2(probabilities_raw.size())) {
        throw std::runtime_error {"The number of probabilities must be a power of 2.\n"};
    }
    std::size_t n_qubits = 0;
    while ((std::size_t {1} << n_qubits) < probabilities_raw.size()) {
        ++n_qubits;
    }
    for (auto q : marginal_qubits) {
        if (q >= n_qubits) {
            throw std::runtime_error {"ERROR: marginal qubit index out of range."};
        }
    }
    auto out = std::map<std::string, std::size_t> {};
    for (auto i : perform_measurements_as_memory(probabilities_raw, n_shots, seed)) {
        ++out[state_index_to_bitstring_marginal(i, marginal_qubits, n_qubits, Endian::LITTLE)];
    }
    return out;
}

auto perform_measurements_as_counts_marginal(const Statevector& state, std::size_t n_shots,
                                             const std::vector<std::size_t>& marginal_qubits, const QuantumNoise* noise,
                                             std::optional<int> seed) -> std::map<std::string, std::size_t>
{
    for (auto q : marginal_qubits) {
        if (q >= state.n_qubits()) {
            throw std::runtime_error {"ERROR: marginal qubit index out of range."};
        }
    }
    auto out = std::map<std::string, std::size_t> {};
    for (auto i : perform_measurements_as_memory(state, n_shots, noise, seed)) {
        ++out[state_index_to_bitstring_marginal(i, marginal_qubits, state.n_qubits(), Endian::LITTLE)];
    }
    return out;
}

// measurements.cpp:178-208: one simulation and one draw per shot.  As in kettle, the circuit is
// simulated without a seed and every shot's sampler is built from the same `seed`.
auto perform_measurements_as_counts_marginal(const QuantumCircuit& circuit, const Statevector& original_state,
                                             std::size_t n_shots, const std::vector<std::size_t>& marginal_qubits,
                                             const QuantumNoise* noise, std::optional<int> seed)
    -> std::map<std::string, std::size_t>
{
    const auto n_qubits = circuit.n_qubits();
    for (auto q : marginal_qubits) {
        if (q >= n_qubits) {
            throw std::runtime_error {"ERROR: marginal qubit index out of range."};
        }
    }
    if (circuit.n_qubits() != original_state.n_qubits()) {
        throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
    }
    // every shot re-simulates the circuit; shots with the same mid-circuit outcome history share one
    // simulated branch on the device (kq_sample_circuit_shots)
    const auto flat = circuit.flatten();
    std::vector<uint64_t> idx(n_shots);
    const auto [hs, sv] = seed_args(seed);
    constexpr uint64_t cache_bytes = uint64_t {8} << 30;
    ki::kq_check(kq_sample_circuit_shots(original_state.device_state(), static_cast<uint32_t>(n_qubits), flat.ops.data(),
                                         flat.ops.size(), flat.predicates.data(), flat.predicates.size(), circuit.n_bits(),
                                         noise_ptr(noise, n_qubits), n_shots, hs, sv, cache_bytes, idx.data(), nullptr));
    auto out = std::map<std::string, std::size_t> {};
    for (const uint64_t i : idx) {
        ++out[state_index_to_bitstring_marginal(i, marginal_qubits, n_qubits, Endian::LITTLE)];
    }
    return out;
}

auto perform_measurements_as_counts(const Statevector& state, std::size_t n_shots, const QuantumNoise* noise,
                                    std::optional<int> seed) -> std::map<std::string, std::size_t>
{
    return perform_measurements_as_counts_marginal(state, n_shots, {}, noise, seed);
}

}  // namespace ket
