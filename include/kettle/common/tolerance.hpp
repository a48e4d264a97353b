#pragma once

// Tolerances with kettle's values (include/kettle/common/tolerance.hpp:6-12 in the reference).
namespace ket
{

constexpr inline auto PROJECTION_NORMALIZATION_TOLERANCE = double {1.0e-8};
constexpr inline auto CONSTRUCTION_NORMALIZATION_TOLERANCE = double {1.0e-6};
constexpr inline auto COMPLEX_ALMOST_EQ_TOLERANCE_SQ = double {1.0e-6};
constexpr inline auto MATRIX_2X2_SQRT_TOLERANCE = double {1.0e-6};
constexpr inline auto MATCHING_PARAMETER_VALUE_TOLERANCE = double {1.0e-6};
constexpr inline auto DENSITY_MATRIX_TRACE_TOLERANCE = double {1.0e-8};
constexpr inline auto MATRIX_HERMITIAN_TOLERANCE = double {1.0e-8};

}  // namespace ket
