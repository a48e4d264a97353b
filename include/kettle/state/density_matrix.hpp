#pragma once

#include <complex>
#include <cstddef>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "kettle/common/tolerance.hpp"
#include "kettle/state/endian.hpp"
#include "kettle/state/statevector.hpp"

struct kq_state;

/*
    kettle's DensityMatrix (include/kettle/state/density_matrix.hpp in the reference), resident
    in device memory as the column-major vector vec(rho).  kettle exposes an Eigen::MatrixXcd;
    Eigen is not a dependency here, so matrix() returns ComplexMatrix, a column-major matrix with
    the same element access (rows(), cols(), operator()(row, col)).
*/
namespace ket
{

class ComplexMatrix
{
public:
    ComplexMatrix() = default;
    ComplexMatrix(std::size_t rows, std::size_t cols) : rows_ {rows}, cols_ {cols}, data_(rows * cols) {}
    [[nodiscard]] auto rows() const -> std::size_t { return rows_; }
    [[nodiscard]] auto cols() const -> std::size_t { return cols_; }
    auto operator()(std::size_t r, std::size_t c) -> std::complex<double>& { return data_[r + c * rows_]; }
    auto operator()(std::size_t r, std::size_t c) const -> const std::complex<double>& { return data_[r + c * rows_]; }
    [[nodiscard]] auto data() const -> const std::vector<std::complex<double>>& { return data_; }
    auto data() -> std::vector<std::complex<double>>& { return data_; }
    [[nodiscard]] auto trace() const -> std::complex<double>;

private:
    std::size_t rows_ {0};
    std::size_t cols_ {0};
    std::vector<std::complex<double>> data_;
};

struct density_matrix_nocheck {};

class DensityMatrix
{
public:
    // |0...0><0...0| of n qubits
    explicit DensityMatrix(std::size_t n_qubits, Precision precision = Precision::C128);
    explicit DensityMatrix(ComplexMatrix matrix, double trace_tolerance = DENSITY_MATRIX_TRACE_TOLERANCE,
                           double hermitian_tolerance = MATRIX_HERMITIAN_TOLERANCE);
    DensityMatrix(ComplexMatrix matrix, const density_matrix_nocheck& key);
    explicit DensityMatrix(const std::string& bitstring, Endian input_endian = Endian::LITTLE);

    DensityMatrix(const DensityMatrix& other);
    DensityMatrix(DensityMatrix&&) noexcept = default;
    auto operator=(const DensityMatrix& other) -> DensityMatrix&;
    auto operator=(DensityMatrix&&) noexcept -> DensityMatrix& = default;
    ~DensityMatrix() = default;

    [[nodiscard]] auto matrix() const -> const ComplexMatrix&;
    auto matrix() -> ComplexMatrix&;
    [[nodiscard]] auto n_states() const noexcept -> std::size_t { return n_states_; }
    [[nodiscard]] auto n_qubits() const noexcept -> std::size_t { return n_qubits_; }
    [[nodiscard]] auto precision() const noexcept -> Precision { return precision_; }
    [[nodiscard]] auto trace() const -> std::complex<double>;
    // |Trace(matrix^2) - 1| within tolerance, as std::norm (density_matrix.hpp:86-91)
    [[nodiscard]] auto is_pure(double tolerance = DENSITY_MATRIX_TRACE_TOLERANCE) const -> bool;
    // the trace of matrix * matrix (density_matrix.hpp:100-104)
    [[nodiscard]] auto trace_of_square() const -> std::complex<double>;

    [[nodiscard]] auto device_state() const -> kq_state*;
    void mark_device_modified() const;

private:
    std::size_t n_qubits_;
    std::size_t n_states_;
    Precision precision_;
    mutable ComplexMatrix host_;
    mutable std::shared_ptr<kq_state> device_;
    mutable bool host_valid_ {false};
    mutable bool device_valid_ {false};
    mutable std::optional<std::size_t> basis_;

    void ensure_host_() const;
};

auto statevector_to_density_matrix(const Statevector& statevector) -> DensityMatrix;
// traces out `qubit_indices` (density_matrix.cpp:231-278)
auto partial_trace(const DensityMatrix& density_matrix, std::vector<std::size_t> qubit_indices) -> DensityMatrix;
auto almost_eq(const DensityMatrix& left, const DensityMatrix& right, double tolerance = 1.0e-6) -> bool;

}  // namespace ket
