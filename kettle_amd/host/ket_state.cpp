// Statevector and DensityMatrix: host objects with a lazily synchronised device mirror.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "kettle/state/density_matrix.hpp"
#include "kettle/state/statevector.hpp"

#include "ket_internal.hpp"

namespace ket
{

namespace ki = ket::internal;

namespace
{

auto create_device(std::size_t n_qubits, Precision p, bool dm) -> std::shared_ptr<kq_state>
{
    kq_state* h = nullptr;
    if (dm) {
        ki::kq_check(kq_dm_create(&h, static_cast<uint32_t>(n_qubits), ki::kq_precision(p)));
    }
    else {
        ki::kq_check(kq_sv_create(&h, static_cast<uint32_t>(n_qubits), ki::kq_precision(p)));
    }
    return std::shared_ptr<kq_state> {h, ki::KqStateDeleter {}};
}

auto is_power_of_2(std::size_t v) -> bool { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace

// ---- Statevector (state/state.cpp:21-123) ------------------------------------------------------------

Statevector::Statevector(std::size_t n_qubits, Precision precision)
    : n_qubits_ {n_qubits}, n_states_ {std::size_t {1} << n_qubits}, precision_ {precision}
{
    if (n_qubits == 0) {
        throw std::runtime_error {"There must be at least 1 qubit in the Statevector.\n"};
    }
    basis_ = 0;
}

Statevector::Statevector(std::vector<std::complex<double>> coefficients, Endian input_endian, double normalization_tolerance)
    : n_qubits_ {0}, n_states_ {coefficients.size()}, precision_ {Precision::C128}, host_ {std::move(coefficients)}
{
    if (host_.size() < 2) {
        throw std::runtime_error {"There must be at least 2 coefficients, representing the states for one qubit.\n"};
    }
    if (!is_power_of_2(host_.size())) {
        auto err = std::stringstream {};
        err << "The provided coefficients must have a size equal to a power of 2.\n";
        err << "Found size = " << host_.size();
        throw std::runtime_error {err.str()};
    }
    auto sum = 0.0;
    for (const auto& x : host_) {
        sum += std::norm(x);
    }
    if (!(std::fabs(sum - 1.0) < normalization_tolerance)) {
        auto err = std::stringstream {};
        err << "The provided coefficients are not properly normalized.\n";
        err << "Found sum of squared norms : " << std::fixed << std::setprecision(14) << sum;
        throw std::runtime_error {err.str()};
    }
    while ((std::size_t {1} << n_qubits_) < host_.size()) {
        ++n_qubits_;
    }
    if (input_endian == Endian::BIG) {
        for (std::size_t i = 0; i < n_states_; ++i) {
            const auto j = endian_flip(i, n_qubits_);
            if (i < j) {
                std::swap(host_[i], host_[j]);
            }
        }
    }
    host_valid_ = true;
}

Statevector::Statevector(const std::string& computational_state, Endian input_endian)
    : n_qubits_ {computational_state.size()}, n_states_ {std::size_t {1} << computational_state.size()}
    , precision_ {Precision::C128}
{
    if (n_qubits_ == 0) {
        throw std::runtime_error {"There must be at least 1 qubit in the Statevector.\n"};
    }
    basis_ = bitstring_to_state_index(computational_state, input_endian);
}

Statevector::Statevector(const Statevector& other)
    : n_qubits_ {other.n_qubits_}, n_states_ {other.n_states_}, precision_ {other.precision_}
{
    if (other.basis_) {
        basis_ = other.basis_;
        return;
    }
    if (other.device_valid_ && !other.host_valid_) {
        // device-to-device copy keeps large states off the host
        device_ = create_device(n_qubits_, precision_, false);
        ki::kq_check(kq_state_copy(device_.get(), other.device_.get()));
        device_valid_ = true;
        return;
    }
    other.ensure_host_();
    host_ = other.host_;
    host_valid_ = true;
}

Statevector::Statevector(Statevector&& other) noexcept = default;

auto Statevector::operator=(const Statevector& other) -> Statevector&
{
    if (this != &other) {
        auto tmp = Statevector {other};
        *this = std::move(tmp);
    }
    return *this;
}

auto Statevector::operator=(Statevector&& other) noexcept -> Statevector& = default;

Statevector::~Statevector() = default;

void Statevector::ensure_host_() const
{
    if (host_valid_) {
        return;
    }
    if (basis_) {
        host_.assign(n_states_, {0.0, 0.0});
        host_[*basis_] = {1.0, 0.0};
    }
    else if (device_valid_) {
        host_.resize(n_states_);
        ki::kq_check(kq_state_download(device_.get(), reinterpret_cast<double*>(host_.data())));
    }
    host_valid_ = true;
}

auto Statevector::coefficients() const -> const std::vector<std::complex<double>>&
{
    ensure_host_();
    return host_;
}

auto Statevector::device_state() const -> kq_state*
{
    if (!device_) {
        device_ = create_device(n_qubits_, precision_, false);
        device_valid_ = false;
    }
    if (!device_valid_) {
        if (basis_) {
            ki::kq_check(kq_state_set_basis(device_.get(), *basis_));
        }
        else {
            ensure_host_();
            ki::kq_check(kq_state_upload(device_.get(), reinterpret_cast<const double*>(host_.data())));
        }
        device_valid_ = true;
    }
    return device_.get();
}

void Statevector::mark_device_modified() const
{
    device_valid_ = true;
    host_valid_ = false;
    basis_.reset();
}

void Statevector::check_index_(std::size_t index) const
{
    if (index >= n_states_) {
        throw std::runtime_error {"Out-of-bounds access for the quantum state.\n"};
    }
}

auto Statevector::operator[](std::size_t index) const -> const std::complex<double>&
{
    ensure_host_();
    return host_[index];
}

auto Statevector::operator[](std::size_t index) -> std::complex<double>&
{
    ensure_host_();
    device_valid_ = false;
    basis_.reset();
    return host_[index];
}

auto Statevector::at(std::size_t index) const -> const std::complex<double>&
{
    check_index_(index);
    return (*this)[index];
}

auto Statevector::at(std::size_t index) -> std::complex<double>&
{
    check_index_(index);
    return (*this)[index];
}

auto Statevector::at(const std::string& bitstring, Endian endian) const -> const std::complex<double>&
{
    return at(bitstring_to_state_index(bitstring, endian));
}

auto almost_eq(const Statevector& left, const Statevector& right, double tolerance_sq) noexcept -> bool
{
    if (left.n_qubits() != right.n_qubits()) {
        return false;
    }
    try {
        const auto& a = left.coefficients();
        const auto& b = right.coefficients();
        for (std::size_t i = 0; i < a.size(); ++i) {
            if (!almost_eq(a[i], b[i], tolerance_sq)) {
                return false;
            }
        }
    }
    catch (...) {
        return false;
    }
    return true;
}

auto inner_product(const Statevector& bra_state, const Statevector& ket_state) -> std::complex<double>
{
    if (bra_state.n_states() != ket_state.n_states()) {
        throw std::runtime_error {"ERROR: cannot calculate inner product between two states of different sizes.\n"};
    }
    const auto& a = bra_state.coefficients();
    const auto& b = ket_state.coefficients();
    auto out = std::complex<double> {};
    for (std::size_t i = 0; i < a.size(); ++i) {
        out += std::conj(a[i]) * b[i];
    }
    return out;
}

auto inner_product_norm_squared(const Statevector& left, const Statevector& right) -> double
{
    return std::norm(inner_product(left, right));
}

auto tensor_product(const Statevector& left, const Statevector& right) -> Statevector
{
    std::vector<std::complex<double>> out;
    out.reserve(left.n_states() * right.n_states());
    for (std::size_t ir = 0; ir < right.n_states(); ++ir) {
        for (std::size_t il = 0; il < left.n_states(); ++il) {
            out.push_back(left[il] * right[ir]);
        }
    }
    return Statevector {std::move(out)};
}

// ---- DensityMatrix (state/density_matrix.cpp) ---------------------------------------------------------

auto ComplexMatrix::trace() const -> std::complex<double>
{
    auto t = std::complex<double> {};
    for (std::size_t i = 0; i < std::min(rows_, cols_); ++i) {
        t += (*this)(i, i);
    }
    return t;
}

DensityMatrix::DensityMatrix(std::size_t n_qubits, Precision precision)
    : n_qubits_ {n_qubits}, n_states_ {std::size_t {1} << n_qubits}, precision_ {precision}
{
    if (n_qubits == 0) {
        throw std::runtime_error {"ERROR: cannot perform a DensityMatrix simulation with 0 qubits.\n"};
    }
    basis_ = 0;
}

DensityMatrix::DensityMatrix(ComplexMatrix matrix, const density_matrix_nocheck&)
    : n_qubits_ {0}, n_states_ {matrix.cols()}, precision_ {Precision::C128}, host_ {std::move(matrix)}
{
    while ((std::size_t {1} << n_qubits_) < n_states_) {
        ++n_qubits_;
    }
    host_valid_ = true;
}

DensityMatrix::DensityMatrix(ComplexMatrix matrix, double trace_tolerance, double hermitian_tolerance)
    : DensityMatrix {std::move(matrix), density_matrix_nocheck {}}
{
    const auto& m = host_;
    if (m.rows() != m.cols()) {
        throw std::runtime_error {"ERROR: the density matrix must be a square matrix.\n"};
    }
    if (!is_power_of_2(m.cols()) || m.cols() < 2) {
        throw std::runtime_error {"ERROR: the side length of the density matrix must be a power of 2.\n"};
    }
    if (std::norm(m.trace() - std::complex<double> {1.0, 0.0}) >= trace_tolerance) {
        throw std::runtime_error {"ERROR: the density matrix must have a trace of 1.\n"};
    }
    for (std::size_t r = 0; r < m.rows(); ++r) {
        for (std::size_t c = 0; c < m.cols(); ++c) {
            if (std::abs(m(r, c) - std::conj(m(c, r))) > hermitian_tolerance) {
                throw std::runtime_error {"ERROR: the density matrix must be Hermitian.\n"};
            }
        }
    }
}

DensityMatrix::DensityMatrix(const std::string& bitstring, Endian input_endian)
    : n_qubits_ {bitstring.size()}, n_states_ {std::size_t {1} << bitstring.size()}, precision_ {Precision::C128}
{
    if (n_qubits_ == 0) {
        throw std::runtime_error {"ERROR: cannot perform a DensityMatrix simulation with 0 qubits.\n"};
    }
    basis_ = bitstring_to_state_index(bitstring, input_endian);
}

DensityMatrix::DensityMatrix(const DensityMatrix& other)
    : n_qubits_ {other.n_qubits_}, n_states_ {other.n_states_}, precision_ {other.precision_}
{
    if (other.basis_) {
        basis_ = other.basis_;
        return;
    }
    if (other.device_valid_ && !other.host_valid_) {
        device_ = create_device(n_qubits_, precision_, true);
        ki::kq_check(kq_state_copy(device_.get(), other.device_.get()));
        device_valid_ = true;
        return;
    }
    other.ensure_host_();
    host_ = other.host_;
    host_valid_ = true;
}

auto DensityMatrix::operator=(const DensityMatrix& other) -> DensityMatrix&
{
    if (this != &other) {
        auto tmp = DensityMatrix {other};
        *this = std::move(tmp);
    }
    return *this;
}

void DensityMatrix::ensure_host_() const
{
    if (host_valid_) {
        return;
    }
    host_ = ComplexMatrix {n_states_, n_states_};
    if (basis_) {
        host_(*basis_, *basis_) = {1.0, 0.0};
    }
    else if (device_valid_) {
        ki::kq_check(kq_state_download(device_.get(), reinterpret_cast<double*>(host_.data().data())));
    }
    host_valid_ = true;
}

auto DensityMatrix::matrix() const -> const ComplexMatrix&
{
    ensure_host_();
    return host_;
}

auto DensityMatrix::matrix() -> ComplexMatrix&
{
    ensure_host_();
    device_valid_ = false;
    basis_.reset();
    return host_;
}

auto DensityMatrix::trace() const -> std::complex<double>
{
    if (host_valid_) {
        return host_.trace();
    }
    double out[2];
    ki::kq_check(kq_dm_trace(device_state(), out));
    return {out[0], out[1]};
}

auto DensityMatrix::trace_of_square() const -> std::complex<double>
{
    double out[2];
    ki::kq_check(kq_dm_trace_of_square(device_state(), out));
    return {out[0], out[1]};
}

auto DensityMatrix::is_pure(double tolerance) const -> bool
{
    return std::norm(trace_of_square() - std::complex<double> {1.0, 0.0}) <= tolerance;
}

auto DensityMatrix::device_state() const -> kq_state*
{
    if (!device_) {
        device_ = create_device(n_qubits_, precision_, true);
        device_valid_ = false;
    }
    if (!device_valid_) {
        if (basis_) {
            ki::kq_check(kq_state_set_basis(device_.get(), *basis_));
        }
        else {
            ensure_host_();
            ki::kq_check(kq_state_upload(device_.get(), reinterpret_cast<const double*>(host_.data().data())));
        }
        device_valid_ = true;
    }
    return device_.get();
}

void DensityMatrix::mark_device_modified() const
{
    device_valid_ = true;
    host_valid_ = false;
    basis_.reset();
}

auto statevector_to_density_matrix(const Statevector& statevector) -> DensityMatrix
{
    auto dm = DensityMatrix {statevector.n_qubits(), statevector.precision()};
    ki::kq_check(kq_dm_from_statevector(dm.device_state(), statevector.device_state()));
    dm.mark_device_modified();
    return dm;
}

auto partial_trace(const DensityMatrix& density_matrix, std::vector<std::size_t> qubit_indices) -> DensityMatrix
{
    const auto n = density_matrix.n_qubits();
    if (std::ranges::any_of(qubit_indices, [n](auto q) { return q >= n; })) {
        throw std::runtime_error {"ERROR: Cannot take partial trace; found qubit index out of bounds\n"};
    }
    std::vector<uint32_t> qs(qubit_indices.begin(), qubit_indices.end());
    const auto side = std::size_t {1} << (n - qs.size());
    auto out = ComplexMatrix {side, side};
    ki::kq_check(kq_dm_partial_trace(density_matrix.device_state(), qs.data(), qs.size(),
                                     reinterpret_cast<double*>(out.data().data())));
    return DensityMatrix {std::move(out), density_matrix_nocheck {}};
}

auto almost_eq(const DensityMatrix& left, const DensityMatrix& right, double tolerance) -> bool
{
    if (left.n_qubits() != right.n_qubits()) {
        return false;
    }
    const auto& a = left.matrix().data();
    const auto& b = right.matrix().data();
    double diff = 0.0, norm = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) {
        diff += std::norm(a[i] - b[i]);
        norm += std::min(std::norm(a[i]), std::norm(b[i]));
    }
    return std::sqrt(diff) <= tolerance * std::sqrt(std::max(norm, 1e-300)) || std::sqrt(diff) < 1e-12;
}

}  // namespace ket
