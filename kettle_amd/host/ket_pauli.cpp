// SparsePauliString, PauliOperator and expectation values (operator/pauli/*.cpp in the reference).
// The expectation runs on the device through kq_pauli_expectation: one pass per distinct X/Y flip
// mask with FP64 accumulation, instead of kettle's copy + apply + inner product per string.
#include <algorithm>
#include <stdexcept>
#include <vector>

#include "kettle/operator/pauli/pauli_operator.hpp"
#include "kettle/operator/pauli/sparse_pauli_string.hpp"

#include "ket_internal.hpp"

namespace ket
{

namespace ki = ket::internal;

namespace
{

using Terms = std::vector<std::pair<std::size_t, PauliTerm>>;

auto sorted_non_identity(const Terms& terms) -> Terms
{
    auto out = Terms {};
    for (const auto& t : terms) {
        if (t.second != PauliTerm::I) {
            out.push_back(t);
        }
    }
    std::ranges::sort(out, [](const auto& l, const auto& r) { return l.first < r.first; });
    return out;
}

void check_n_qubits(std::size_t n)
{
    if (n == 0) {
        throw std::runtime_error {"ERROR: SparsePauliString cannot be constructed with 0 qubits.\n"};
    }
}

struct Packed
{
    std::vector<uint32_t> offsets {0};
    std::vector<uint32_t> qubits;
    std::vector<uint8_t> kinds;
    std::vector<uint8_t> phases;
    std::vector<double> coeffs;

    void add(std::complex<double> c, const SparsePauliString& s)
    {
        for (const auto& [q, t] : s.terms()) {
            qubits.push_back(static_cast<uint32_t>(q));
            kinds.push_back(static_cast<uint8_t>(t));
        }
        offsets.push_back(static_cast<uint32_t>(qubits.size()));
        phases.push_back(static_cast<uint8_t>(s.phase()));
        coeffs.push_back(c.real());
        coeffs.push_back(c.imag());
    }
};

auto device_expectation(const Packed& p, const Statevector& state) -> std::complex<double>
{
    double out[2] = {0.0, 0.0};
    ki::kq_check(kq_pauli_expectation(state.device_state(), p.phases.size(), p.offsets.data(), p.qubits.data(),
                                      p.kinds.data(), p.phases.data(), p.coeffs.data(), out));
    return {out[0], out[1]};
}

}  // namespace

SparsePauliString::SparsePauliString(std::size_t n_qubits, PauliPhase phase) : n_qubits_ {n_qubits}, phase_ {phase}
{
    check_n_qubits(n_qubits_);
}

SparsePauliString::SparsePauliString(std::vector<std::pair<std::size_t, PauliTerm>> terms, std::size_t n_qubits,
                                     PauliPhase phase)
    : n_qubits_ {n_qubits}, phase_ {phase}, terms_ {std::move(terms)}
{
    check_n_qubits(n_qubits_);
}

SparsePauliString::SparsePauliString(const std::vector<PauliTerm>& pauli_terms, PauliPhase phase)
    : n_qubits_ {pauli_terms.size()}, phase_ {phase}
{
    check_n_qubits(n_qubits_);
    for (std::size_t i = 0; i < pauli_terms.size(); ++i) {
        terms_.emplace_back(i, pauli_terms[i]);
    }
}

SparsePauliString::SparsePauliString(const std::initializer_list<PauliTerm>& pauli_terms, PauliPhase phase)
    : SparsePauliString {std::vector<PauliTerm> {pauli_terms}, phase}
{}

auto SparsePauliString::vector_index_(std::size_t qubit_index) const -> std::optional<std::size_t>
{
    if (qubit_index >= n_qubits_) {
        return std::nullopt;
    }
    for (std::size_t i = 0; i < terms_.size(); ++i) {
        if (terms_[i].first == qubit_index) {
            return i;
        }
    }
    return std::nullopt;
}

void SparsePauliString::check_index_in_qubit_range_(std::size_t index) const
{
    if (index >= n_qubits_) {
        throw std::runtime_error {"ERROR: cannot perform operation on SparsePauliString with index beyond qubit range.\n"};
    }
}

auto SparsePauliString::at(std::size_t qubit_index) const -> PauliTerm
{
    if (const auto i = vector_index_(qubit_index)) {
        return terms_[*i].second;
    }
    throw std::runtime_error {"ERROR: no Pauli term found for provided qubit index.\n"};
}

void SparsePauliString::add(std::size_t qubit_index, PauliTerm term)
{
    check_index_in_qubit_range_(qubit_index);
    if (vector_index_(qubit_index)) {
        throw std::runtime_error {"ERROR: Pauli term is already present in the string\n"};
    }
    terms_.emplace_back(qubit_index, term);
}

void SparsePauliString::overwrite(std::size_t qubit_index, PauliTerm term)
{
    check_index_in_qubit_range_(qubit_index);
    if (const auto i = vector_index_(qubit_index)) {
        terms_[*i].second = term;
    }
    else {
        terms_.emplace_back(qubit_index, term);
    }
}

void SparsePauliString::remove(std::size_t qubit_index)
{
    if (const auto i = vector_index_(qubit_index)) {
        terms_.erase(terms_.begin() + static_cast<std::ptrdiff_t>(*i));
    }
}

auto SparsePauliString::contains_index(std::size_t qubit_index) const noexcept -> bool
{
    return std::ranges::any_of(terms_, [qubit_index](const auto& p) { return p.first == qubit_index; });
}

auto SparsePauliString::equal_up_to_phase(const SparsePauliString& other) const -> bool
{
    return n_qubits_ == other.n_qubits_ && sorted_non_identity(terms_) == sorted_non_identity(other.terms_);
}

auto operator==(const SparsePauliString& left, const SparsePauliString& right) -> bool
{
    return left.phase_ == right.phase_ && left.equal_up_to_phase(right);
}

// ---- PauliOperator ----------------------------------------------------------------------------------------

PauliOperator::PauliOperator(std::size_t n_qubits) : n_qubits_ {n_qubits} {}

PauliOperator::PauliOperator(std::vector<WeightedPauliString> weighted) : n_qubits_ {0}, strings_ {std::move(weighted)}
{
    if (strings_.empty()) {
        throw std::runtime_error {
            "ERROR: construction of `PauliOperator` with only WeightedPauliString instances requires\n"
            "a non-empty vector.\n"};
    }
    const auto differ = [](const auto& l, const auto& r) { return l.pauli_string.n_qubits() != r.pauli_string.n_qubits(); };
    if (std::ranges::adjacent_find(strings_, differ) != strings_.end()) {
        throw std::runtime_error {
            "ERROR: construction of `PauliOperator` with WeightedPauliString instances requires\n"
            "all pauli strings to have the same number of qubits.\n"};
    }
    n_qubits_ = strings_[0].pauli_string.n_qubits();
}

PauliOperator::PauliOperator(const std::initializer_list<WeightedPauliString>& weighted)
    : PauliOperator {std::vector<WeightedPauliString> {weighted}}
{}

void PauliOperator::add(WeightedPauliString value) { strings_.emplace_back(std::move(value)); }

void PauliOperator::add(std::complex<double> coefficient, SparsePauliString pauli_string)
{
    strings_.push_back(WeightedPauliString {coefficient, std::move(pauli_string)});
}

void PauliOperator::remove(std::size_t index)
{
    if (index >= strings_.size()) {
        throw std::runtime_error {"ERROR: attempted to remove out of bounds element from PauliOperator.\n"};
    }
    strings_.erase(strings_.begin() + static_cast<std::ptrdiff_t>(index));
}

auto almost_eq(const PauliOperator& left_op, const PauliOperator& right_op, double coeff_tolerance) -> bool
{
    if (left_op.size() != right_op.size()) {
        return false;
    }
    for (std::size_t i = 0; i < left_op.size(); ++i) {
        const auto& l = left_op.at(i);
        const auto& r = right_op.at(i);
        if (!almost_eq(l.coefficient, r.coefficient, coeff_tolerance) || !(l.pauli_string == r.pauli_string)) {
            return false;
        }
    }
    return true;
}

auto expectation_value(const PauliOperator& pauli_op, const Statevector& state) -> std::complex<double>
{
    auto p = Packed {};
    for (const auto& [coeff, s] : pauli_op.weighted_pauli_strings()) {
        if (s.n_qubits() != state.n_qubits()) {
            throw std::runtime_error {"Invalid simulation; SparsePauliString and state have different number of qubits."};
        }
        p.add(coeff, s);
    }
    if (p.phases.empty()) {
        return {0.0, 0.0};
    }
    return device_expectation(p, state);
}

auto expectation_value(const SparsePauliString& sparse_pauli_string, const Statevector& state) -> std::complex<double>
{
    if (sparse_pauli_string.n_qubits() != state.n_qubits()) {
        throw std::runtime_error {"Invalid simulation; SparsePauliString and state have different number of qubits."};
    }
    auto p = Packed {};
    p.add({1.0, 0.0}, sparse_pauli_string);
    return device_expectation(p, state);
}

}  // namespace ket
