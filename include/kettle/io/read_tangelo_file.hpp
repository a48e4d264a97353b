#pragma once

#include <cstddef>
#include <filesystem>
#include <istream>
#include <optional>

#include "kettle/circuit/circuit.hpp"

/*
    Reader of tangelo-style circuit files (include/kettle/io/read_tangelo_file.hpp:29-45 and
    source/kettle_internal/io/read_tangelo_file.cpp:241-376 in the reference).  One gate per line:

        H         target : [4]
        RX        target : [5]   parameter : 1.5707963267948966
        CNOT      target : [4]   control : [2]
        SWAP      target : [12, 9]
        M         target : [1]   bit : [4]
        U         target : [1]                  followed by two lines of [re, im]   [re, im]
        IF BITS[0, 1] == [1, 0]                 followed by the body, each line indented 4 spaces
        ELSE                                    (after an IF block) followed by the else body

    Tangelo names CNOT, CPHASE and PHASE are read as CX, CP and P; SWAP becomes three CX gates.
    An unknown gate name throws std::runtime_error.
*/
namespace ket
{

// Reads from `stream` after skipping `n_skip_lines` lines.  With `line_starts_with_spaces`, reading
// stops (and the stream is rewound to that line) at the first line without that many leading spaces.
auto read_tangelo_circuit(std::size_t n_qubits, std::istream& stream, std::size_t n_skip_lines,
                          std::optional<std::size_t> line_starts_with_spaces = std::nullopt) -> QuantumCircuit;

auto read_tangelo_circuit(std::size_t n_qubits, const std::filesystem::path& filepath, std::size_t n_skip_lines)
    -> QuantumCircuit;

}  // namespace ket
