// Host-side pieces of the ket:: API that need no device: 2x2 matrices, gate matrices, random
// unitaries, parameter ids, control-flow predicates and bitstring maps.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <random>
#include <sstream>
#include <stdexcept>

#include "kettle/circuit/control_flow_predicate.hpp"
#include "kettle/common/matrix2x2.hpp"
#include "kettle/gates/common_u_gates.hpp"
#include "kettle/gates/random_u_gates.hpp"
#include "kettle/parameter/parameter.hpp"
#include "kettle/state/qubit_state_conversion.hpp"

#include "ket_internal.hpp"

namespace ket
{

auto operator*(const Matrix2X2& l, const Matrix2X2& r) noexcept -> Matrix2X2
{
    return {l.elem00 * r.elem00 + l.elem01 * r.elem10, l.elem00 * r.elem01 + l.elem01 * r.elem11,
            l.elem10 * r.elem00 + l.elem11 * r.elem10, l.elem10 * r.elem01 + l.elem11 * r.elem11};
}

auto operator+(const Matrix2X2& l, const Matrix2X2& r) noexcept -> Matrix2X2
{
    return {l.elem00 + r.elem00, l.elem01 + r.elem01, l.elem10 + r.elem10, l.elem11 + r.elem11};
}

auto conjugate_transpose(const Matrix2X2& m) -> Matrix2X2
{
    return {std::conj(m.elem00), std::conj(m.elem10), std::conj(m.elem01), std::conj(m.elem11)};
}

auto determinant(const Matrix2X2& m) -> std::complex<double> { return m.elem00 * m.elem11 - m.elem01 * m.elem10; }

auto trace(const Matrix2X2& m) -> std::complex<double> { return m.elem00 + m.elem11; }

auto almost_eq(const std::complex<double>& left, const std::complex<double>& right, double tolerance_sq) noexcept -> bool
{
    const auto dr = left.real() - right.real();
    const auto di = left.imag() - right.imag();
    return dr * dr + di * di < tolerance_sq;
}

auto almost_eq(const Matrix2X2& l, const Matrix2X2& r, double tolerance_sq) noexcept -> bool
{
    return almost_eq(l.elem00, r.elem00, tolerance_sq) && almost_eq(l.elem01, r.elem01, tolerance_sq)
        && almost_eq(l.elem10, r.elem10, tolerance_sq) && almost_eq(l.elem11, r.elem11, tolerance_sq);
}

// ---- gate matrices (the textbook matrices of gates/common_u_gates.cpp) --------------------------

auto i_gate() noexcept -> Matrix2X2 { return {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}}; }
auto h_gate() noexcept -> Matrix2X2 { return {{M_SQRT1_2, 0.0}, {M_SQRT1_2, 0.0}, {M_SQRT1_2, 0.0}, {-M_SQRT1_2, 0.0}}; }
auto x_gate() noexcept -> Matrix2X2 { return {{0.0, 0.0}, {1.0, 0.0}, {1.0, 0.0}, {0.0, 0.0}}; }
auto y_gate() noexcept -> Matrix2X2 { return {{0.0, 0.0}, {0.0, -1.0}, {0.0, 1.0}, {0.0, 0.0}}; }
auto z_gate() noexcept -> Matrix2X2 { return {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {-1.0, 0.0}}; }
auto s_gate() noexcept -> Matrix2X2 { return {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 1.0}}; }
auto sdag_gate() noexcept -> Matrix2X2 { return {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, -1.0}}; }
auto t_gate() noexcept -> Matrix2X2 { return {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {M_SQRT1_2, M_SQRT1_2}}; }
auto tdag_gate() noexcept -> Matrix2X2 { return {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {M_SQRT1_2, -M_SQRT1_2}}; }
auto sx_gate() noexcept -> Matrix2X2 { return {{0.5, 0.5}, {0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}}; }
auto sxdag_gate() noexcept -> Matrix2X2 { return {{0.5, -0.5}, {0.5, 0.5}, {0.5, 0.5}, {0.5, -0.5}}; }

auto rx_gate(double angle) noexcept -> Matrix2X2
{
    const auto c = std::cos(angle / 2.0), s = std::sin(angle / 2.0);
    return {{c, 0.0}, {0.0, -s}, {0.0, -s}, {c, 0.0}};
}

auto ry_gate(double angle) noexcept -> Matrix2X2
{
    const auto c = std::cos(angle / 2.0), s = std::sin(angle / 2.0);
    return {{c, 0.0}, {-s, 0.0}, {s, 0.0}, {c, 0.0}};
}

auto rz_gate(double angle) noexcept -> Matrix2X2
{
    const auto c = std::cos(angle / 2.0), s = std::sin(angle / 2.0);
    return {{c, -s}, {0.0, 0.0}, {0.0, 0.0}, {c, s}};
}

auto p_gate(double angle) noexcept -> Matrix2X2
{
    return {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {std::cos(angle), std::sin(angle)}};
}

auto non_angle_gate(Gate gate) -> Matrix2X2
{
    switch (gate) {
        case Gate::H: case Gate::CH: return h_gate();
        case Gate::X: case Gate::CX: return x_gate();
        case Gate::Y: case Gate::CY: return y_gate();
        case Gate::Z: case Gate::CZ: return z_gate();
        case Gate::S: case Gate::CS: return s_gate();
        case Gate::SDAG: case Gate::CSDAG: return sdag_gate();
        case Gate::T: case Gate::CT: return t_gate();
        case Gate::TDAG: case Gate::CTDAG: return tdag_gate();
        case Gate::SX: case Gate::CSX: return sx_gate();
        case Gate::SXDAG: case Gate::CSXDAG: return sxdag_gate();
        default: break;
    }
    throw std::runtime_error {"UNREACHABLE: dev error, invalid non-angle gate provided\n"};
}

auto angle_gate(Gate gate, double angle) -> Matrix2X2
{
    switch (gate) {
        case Gate::RX: case Gate::CRX: return rx_gate(angle);
        case Gate::RY: case Gate::CRY: return ry_gate(angle);
        case Gate::RZ: case Gate::CRZ: return rz_gate(angle);
        case Gate::P: case Gate::CP: return p_gate(angle);
        default: break;
    }
    throw std::runtime_error {"UNREACHABLE: dev error, invalid angle gate provided\n"};
}

// ---- random U(2) (generate_random_unitary2x2, random_u_gates.cpp:13-54) ----------------------------

auto generate_random_unitary2x2(std::mt19937& prng) -> Matrix2X2
{
    auto uniform = std::uniform_real_distribution<double> {0.0, 1.0};
    const auto alpha = 2.0 * M_PI * uniform(prng);
    const auto phi = 2.0 * M_PI * uniform(prng);
    const auto chi = 2.0 * M_PI * uniform(prng);
    const auto psi = std::asin(std::sqrt(uniform(prng)));
    const auto g = std::complex<double> {std::cos(alpha), std::sin(alpha)};
    const auto ep = std::complex<double> {std::cos(phi), std::sin(phi)};
    const auto ec = std::complex<double> {std::cos(chi), std::sin(chi)};
    return {g * std::cos(psi) * ep, g * std::sin(psi) * ec, -g * std::sin(psi) / ec, g * std::cos(psi) / ep};
}

auto generate_random_unitary2x2(int seed) -> Matrix2X2
{
    auto prng = internal::get_prng(seed);
    return generate_random_unitary2x2(prng);
}

auto generate_random_unitary2x2() -> Matrix2X2
{
    auto prng = internal::get_prng(std::nullopt);
    return generate_random_unitary2x2(prng);
}

// ---- control flow predicate (control_flow_predicate.cpp:12-69) -----------------------------------

ControlFlowPredicate::ControlFlowPredicate(std::vector<std::size_t> bit_indices_to_check, std::vector<int> expected_bits,
                                           ControlFlowBooleanKind control_kind)
    : bit_indices_to_check_ {std::move(bit_indices_to_check)}
    , expected_bits_ {std::move(expected_bits)}
    , control_kind_ {control_kind}
{
    if (bit_indices_to_check_.size() != expected_bits_.size()) {
        auto err = std::stringstream {};
        err << "There is a mismatch in the number of bits to check, and the number of expected bits:\n";
        err << "bit_indices_to_check.size() = " << bit_indices_to_check_.size() << '\n';
        err << "expected_bits.size() = " << expected_bits_.size() << '\n';
        throw std::runtime_error {err.str()};
    }
    if (bit_indices_to_check_.empty()) {
        throw std::runtime_error {"The control flow predicate must depend on at least one bit.\n"};
    }
    if (!std::ranges::all_of(expected_bits_, [](int x) { return x == 0 || x == 1; })) {
        throw std::runtime_error {"The expected bits must all be 0 or 1.\n"};
    }
}

auto ControlFlowPredicate::operator()(const ClassicalRegister& creg) const -> bool
{
    auto output = true;
    for (std::size_t i = 0; i < bit_indices_to_check_.size(); ++i) {
        const auto bit = bit_indices_to_check_[i];
        if (!creg.is_measured(bit)) {
            auto err = std::stringstream {};
            err << "There is no measured bit at classical register " << bit << '\n';
            throw std::runtime_error {err.str()};
        }
        if (creg.get(bit) != expected_bits_[i]) {
            output = false;
            break;
        }
    }
    return control_kind_ == ControlFlowBooleanKind::IF ? output : !output;
}

// ---- bitstrings (qubit_state_conversion.cpp:37-180, marginal.cpp:19-43) ---------------------------

auto endian_flip(std::size_t value, std::size_t n_relevant_bits) -> std::size_t
{
    auto backward = std::size_t {0};
    for (std::size_t i = 0; i < n_relevant_bits; ++i) {
        backward <<= 1U;
        backward |= (value & 1U);
        value >>= 1U;
    }
    return backward;
}

auto bitstring_to_state_index(const std::string& bitstring, Endian input_endian) -> std::size_t
{
    auto index = std::size_t {0};
    const auto n = bitstring.size();
    for (std::size_t i = 0; i < n; ++i) {
        const char ch = bitstring[i];
        if (ch != '0' && ch != '1') {
            throw std::runtime_error {"A bitstring must contain only '0' and '1'.\n"};
        }
        const auto bit = input_endian == Endian::LITTLE ? i : n - 1 - i;
        if (ch == '1') {
            index |= std::size_t {1} << bit;
        }
    }
    return index;
}

auto state_index_to_bitstring(std::size_t i_state, std::size_t n_qubits, Endian input_endian) -> std::string
{
    auto out = std::string(n_qubits, '0');
    for (std::size_t i = 0; i < n_qubits; ++i) {
        const auto pos = input_endian == Endian::LITTLE ? i : n_qubits - 1 - i;
        out[pos] = ((i_state >> i) & 1U) ? '1' : '0';
    }
    return out;
}

auto state_index_to_bitstring_marginal(std::size_t i_state, const std::vector<std::size_t>& marginal_qubits,
                                       std::size_t n_qubits, Endian input_endian) -> std::string
{
    auto out = state_index_to_bitstring(i_state, n_qubits, input_endian);
    for (auto q : marginal_qubits) {
        if (q >= n_qubits) {
            throw std::runtime_error {"ERROR: marginal qubit index out of range."};
        }
        out[input_endian == Endian::LITTLE ? q : n_qubits - 1 - q] = 'x';
    }
    return out;
}

}  // namespace ket

// ---- parameters (parameter/parameter.cpp) ------------------------------------------------------------
namespace ket::param
{

namespace
{
auto random_id(std::mt19937& prng) -> ParameterID
{
    auto dist = std::uniform_int_distribution<int> {0, 255};
    auto id = ParameterID {};
    for (auto& b : id) {
        b = static_cast<std::uint8_t>(dist(prng));
    }
    return id;
}
}  // namespace

Parameter::Parameter(std::string name, const ParameterID& id) : name_ {std::move(name)}, id_ {id} {}

Parameter::Parameter(std::string name, int seed) : name_ {std::move(name)}
{
    auto prng = std::mt19937 {static_cast<std::mt19937::result_type>(seed)};
    id_ = random_id(prng);
}

Parameter::Parameter(std::string name) : name_ {std::move(name)}
{
    auto device = std::random_device {};
    auto prng = std::mt19937 {device()};
    id_ = random_id(prng);
}

auto ParameterIdHash::operator()(const ParameterID& id) const noexcept -> std::size_t
{
    std::size_t h = 1469598103934665603ULL;
    for (auto b : id) {
        h = (h ^ b) * 1099511628211ULL;
    }
    return h;
}

}  // namespace ket::param
