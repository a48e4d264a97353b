#pragma once

#include <filesystem>
#include <istream>

#include "kettle/state/statevector.hpp"

/*
    Reader of statevectors written by numpy (include/kettle/io/numpy_statevector.hpp and
    source/kettle_internal/io/numpy_statevector.cpp:35-76 in the reference): the number of qubits
    n on the first line, then 2^n amplitudes as "(re+imj)", one per line.
*/
namespace ket
{

auto read_numpy_statevector(std::istream& instream, Endian input_endian = Endian::LITTLE) -> Statevector;

auto read_numpy_statevector(const std::filesystem::path& filepath, Endian input_endian = Endian::LITTLE)
    -> Statevector;

}  // namespace ket
