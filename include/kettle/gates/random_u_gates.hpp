#pragma once

#include <random>

#include "kettle/common/matrix2x2.hpp"

// Haar-random U(2) (include/kettle/gates/random_u_gates.hpp in the reference).
namespace ket
{

auto generate_random_unitary2x2(std::mt19937& prng) -> Matrix2X2;
auto generate_random_unitary2x2(int seed) -> Matrix2X2;
auto generate_random_unitary2x2() -> Matrix2X2;

}  // namespace ket
