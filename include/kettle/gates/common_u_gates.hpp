#pragma once

#include "kettle/common/matrix2x2.hpp"
#include "kettle/gates/primitive_gate.hpp"

// Standard 2x2 gate matrices (include/kettle/gates/common_u_gates.hpp in the reference).
namespace ket
{

auto i_gate() noexcept -> Matrix2X2;
auto h_gate() noexcept -> Matrix2X2;
auto x_gate() noexcept -> Matrix2X2;
auto y_gate() noexcept -> Matrix2X2;
auto z_gate() noexcept -> Matrix2X2;
auto s_gate() noexcept -> Matrix2X2;
auto sdag_gate() noexcept -> Matrix2X2;
auto t_gate() noexcept -> Matrix2X2;
auto tdag_gate() noexcept -> Matrix2X2;
auto sx_gate() noexcept -> Matrix2X2;
auto sxdag_gate() noexcept -> Matrix2X2;
auto rx_gate(double angle) noexcept -> Matrix2X2;
auto ry_gate(double angle) noexcept -> Matrix2X2;
auto rz_gate(double angle) noexcept -> Matrix2X2;
auto p_gate(double angle) noexcept -> Matrix2X2;
auto non_angle_gate(Gate gate) -> Matrix2X2;
auto angle_gate(Gate gate, double angle) -> Matrix2X2;

}  // namespace ket
