#pragma once

#include <cstddef>
#include <cstdint>
#include <optional>

#include "kettle/common/matrix2x2.hpp"
#include "kettle/parameter/parameter.hpp"

// kettle's primitive gates (include/kettle/gates/primitive_gate.hpp:19-79 in the reference).
// The enumerator values equal the gate codes of libkq's C-ABI (include/kq.h).
namespace ket
{

enum class Gate : std::uint8_t
{
    H, X, Y, Z, S, SDAG, T, TDAG, SX, SXDAG,
    RX, RY, RZ, P,
    CH, CX, CY, CZ, CS, CSDAG, CT, CTDAG, CSX, CSXDAG,
    CRX, CRY, CRZ, CP,
    U, CU,
    M
};

/*
    One primitive gate: arg0 = target (1-qubit gates, M qubit) or control (2-qubit gates),
    arg1 = target (2-qubit gates) or classical bit (M), arg2 = angle.  A parameterised gate
    carries its parameter id instead of a fixed angle; U / CU carry their matrix.
*/
struct GateInfo
{
    Gate gate;
    std::size_t arg0;
    std::size_t arg1;
    double arg2;
    std::optional<Matrix2X2> unitary;
    std::optional<ket::param::ParameterID> parameter;
};

}  // namespace ket
