#pragma once

#include <cstddef>
#include <cstdint>
#include <tuple>
#include <vector>

#include "kettle/circuit/circuit.hpp"
#include "kettle/gates/primitive_gate.hpp"
#include "kettle/parameter/parameter.hpp"

// kettle's n_local ansatz builder (include/kettle/optimize/n_local.hpp) for primitive gates.
namespace ket
{

inline constexpr auto DEFAULT_NLOCAL_GATE_PARAMETER = double {0.0};

enum class NLocalEntangelement : std::uint8_t
{
    FULL,
    LINEAR
};

enum class SkipLastRotationLayerFlag : std::uint8_t
{
    TRUE,
    FALSE
};

auto n_local(std::size_t n_qubits, const std::vector<Gate>& rotation_blocks, const std::vector<Gate>& entanglement_blocks,
             NLocalEntangelement entanglement_kind, std::size_t n_repetitions,
             SkipLastRotationLayerFlag flag = SkipLastRotationLayerFlag::FALSE)
    -> std::tuple<QuantumCircuit, std::vector<ket::param::ParameterID>>;

}  // namespace ket
