#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <optional>
#include <unordered_map>
#include <utility>
#include <vector>

// kettle's SparsePauliString (include/kettle/operator/pauli/sparse_pauli_string.hpp:42-144).
namespace ket
{

enum class PauliTerm : std::uint8_t
{
    I,
    X,
    Y,
    Z,
};

enum class PauliPhase : std::uint8_t
{
    PLUS_ONE,
    PLUS_EYE,
    MINUS_ONE,
    MINUS_EYE,
};

const auto PAULI_PHASE_MAP = std::unordered_map<PauliPhase, std::complex<double>> {
    {PauliPhase::PLUS_ONE, {1.0, 0.0}},
    {PauliPhase::PLUS_EYE, {0.0, 1.0}},
    {PauliPhase::MINUS_ONE, {-1.0, 0.0}},
    {PauliPhase::MINUS_EYE, {0.0, -1.0}},
};

class SparsePauliString
{
public:
    explicit SparsePauliString(std::size_t n_qubits, PauliPhase phase = PauliPhase::PLUS_ONE);
    SparsePauliString(std::vector<std::pair<std::size_t, PauliTerm>> pauli_indexed_terms, std::size_t n_qubits,
                      PauliPhase phase = PauliPhase::PLUS_ONE);
    SparsePauliString(const std::vector<PauliTerm>& pauli_terms, PauliPhase phase = PauliPhase::PLUS_ONE);          // NOLINT
    SparsePauliString(const std::initializer_list<PauliTerm>& pauli_terms, PauliPhase phase = PauliPhase::PLUS_ONE); // NOLINT

    void set_phase(PauliPhase phase) noexcept { phase_ = phase; }
    [[nodiscard]] auto phase() const noexcept -> PauliPhase { return phase_; }
    [[nodiscard]] auto n_qubits() const noexcept -> std::size_t { return n_qubits_; }
    [[nodiscard]] auto terms() const noexcept -> const std::vector<std::pair<std::size_t, PauliTerm>>& { return terms_; }
    [[nodiscard]] auto size() const noexcept -> std::size_t { return terms_.size(); }
    [[nodiscard]] auto at(std::size_t qubit_index) const -> PauliTerm;
    void add(std::size_t qubit_index, PauliTerm term);
    void overwrite(std::size_t qubit_index, PauliTerm term);
    void remove(std::size_t qubit_index);
    [[nodiscard]] auto contains_index(std::size_t qubit_index) const noexcept -> bool;
    [[nodiscard]] auto equal_up_to_phase(const SparsePauliString& other) const -> bool;
    friend auto operator==(const SparsePauliString& left, const SparsePauliString& right) -> bool;

private:
    std::size_t n_qubits_;
    PauliPhase phase_ {};
    std::vector<std::pair<std::size_t, PauliTerm>> terms_;

    [[nodiscard]] auto vector_index_(std::size_t qubit_index) const -> std::optional<std::size_t>;
    void check_index_in_qubit_range_(std::size_t index) const;
};

}  // namespace ket
