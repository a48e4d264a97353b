#pragma once

#include <complex>

#include "kettle/common/mathtools.hpp"
#include "kettle/common/tolerance.hpp"

// kettle's 2x2 complex matrix (include/kettle/common/matrix2x2.hpp in the reference).
namespace ket
{

struct Matrix2X2
{
    std::complex<double> elem00;
    std::complex<double> elem01;
    std::complex<double> elem10;
    std::complex<double> elem11;
};

auto operator*(const Matrix2X2& lhs, const Matrix2X2& rhs) noexcept -> Matrix2X2;
auto operator+(const Matrix2X2& lhs, const Matrix2X2& rhs) noexcept -> Matrix2X2;
auto conjugate_transpose(const Matrix2X2& mat) -> Matrix2X2;
auto determinant(const Matrix2X2& mat) -> std::complex<double>;
auto trace(const Matrix2X2& mat) -> std::complex<double>;
auto almost_eq(const Matrix2X2& left, const Matrix2X2& right,
               double tolerance_sq = COMPLEX_ALMOST_EQ_TOLERANCE_SQ) noexcept -> bool;

}  // namespace ket
