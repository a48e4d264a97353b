#pragma once

#include <complex>

#include "kettle/common/tolerance.hpp"

// Complex comparison (include/kettle/common/mathtools.hpp in the reference): |left - right|^2 < tol.
namespace ket
{

auto almost_eq(const std::complex<double>& left, const std::complex<double>& right,
               double tolerance_sq = COMPLEX_ALMOST_EQ_TOLERANCE_SQ) noexcept -> bool;

}  // namespace ket
