#pragma once

#include <complex>
#include <cstddef>
#include <initializer_list>
#include <vector>

#include "kettle/common/tolerance.hpp"
#include "kettle/operator/pauli/sparse_pauli_string.hpp"
#include "kettle/state/statevector.hpp"

// kettle's PauliOperator and expectation values (include/kettle/operator/pauli/pauli_operator.hpp).
namespace ket
{

struct WeightedPauliString
{
    std::complex<double> coefficient;
    SparsePauliString pauli_string;
};

class PauliOperator
{
public:
    explicit PauliOperator(std::size_t n_qubits);
    PauliOperator(std::vector<WeightedPauliString> weighted_pauli_strings);                    // NOLINT
    PauliOperator(const std::initializer_list<WeightedPauliString>& weighted_pauli_strings);

    [[nodiscard]] auto n_qubits() const noexcept -> std::size_t { return n_qubits_; }
    [[nodiscard]] auto size() const noexcept -> std::size_t { return strings_.size(); }
    [[nodiscard]] auto weighted_pauli_strings() const noexcept -> const std::vector<WeightedPauliString>& { return strings_; }
    [[nodiscard]] auto at(std::size_t index) const -> const WeightedPauliString& { return strings_[index]; }
    auto at(std::size_t index) -> WeightedPauliString& { return strings_[index]; }
    void add(WeightedPauliString value);
    void add(std::complex<double> coefficient, SparsePauliString pauli_string);
    void remove(std::size_t index);

private:
    std::size_t n_qubits_;
    std::vector<WeightedPauliString> strings_;
};

auto almost_eq(const PauliOperator& left_op, const PauliOperator& right_op,
               double coeff_tolerance = COMPLEX_ALMOST_EQ_TOLERANCE_SQ) -> bool;
auto expectation_value(const PauliOperator& pauli_op, const Statevector& state) -> std::complex<double>;
auto expectation_value(const SparsePauliString& sparse_pauli_string, const Statevector& state) -> std::complex<double>;

}  // namespace ket
