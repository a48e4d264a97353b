#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "kettle/state/endian.hpp"

// Bitstring <-> index maps (qubit_state_conversion.hpp in the reference).  Little endian:
// character i of the bitstring is qubit i; marginal qubits print as 'x'.
namespace ket
{

auto endian_flip(std::size_t value, std::size_t n_relevant_bits) -> std::size_t;
auto bitstring_to_state_index(const std::string& bitstring, Endian input_endian = Endian::LITTLE) -> std::size_t;
auto state_index_to_bitstring(std::size_t i_state, std::size_t n_qubits, Endian input_endian = Endian::LITTLE) -> std::string;
auto state_index_to_bitstring_marginal(std::size_t i_state, const std::vector<std::size_t>& marginal_qubits,
                                       std::size_t n_qubits, Endian input_endian = Endian::LITTLE) -> std::string;

}  // namespace ket
