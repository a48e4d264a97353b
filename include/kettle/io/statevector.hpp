#pragma once

#include <filesystem>
#include <istream>
#include <ostream>

#include "kettle/state/statevector.hpp"

/*
    Statevector text snapshots (include/kettle/io/statevector.hpp and
    source/kettle_internal/io/statevector.cpp:74-169 in the reference):

        ENDIANNESS: LITTLE            (or BIG: amplitudes listed in bit-reversed index order)
        NUMBER OF STATES: 8
         0.35355339059327   -0.00000000000000
        ...

    one amplitude per line, real and imaginary parts fixed with 14 decimals, non-negative values
    padded by a space.  The state is downloaded from the device for a save; a load builds a
    Statevector (validated like the amplitude constructor) that is uploaded on first use.
*/
namespace ket
{

void save_statevector(std::ostream& outstream, const Statevector& state, Endian endian = Endian::LITTLE);

void save_statevector(const std::filesystem::path& filepath, const Statevector& state, Endian endian = Endian::LITTLE);

auto load_statevector(std::istream& instream) -> Statevector;

auto load_statevector(const std::filesystem::path& filepath) -> Statevector;

}  // namespace ket
