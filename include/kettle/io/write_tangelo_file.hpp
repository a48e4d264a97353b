#pragma once

#include <cstddef>
#include <filesystem>
#include <ostream>

#include "kettle/circuit/circuit.hpp"

/*
    Writer of tangelo-style circuit files (include/kettle/io/write_tangelo_file.hpp:31-39 and
    source/kettle_internal/io/write_tangelo_file.cpp:24-213 in the reference), the inverse of
    read_tangelo_circuit: gate names padded to 10 columns, angles with 16 decimals, U / CU matrix
    elements as the first 16 characters of "% .17f", control-flow bodies indented by 4 spaces.
    Circuit loggers are skipped.  A parameterised gate is written with the placeholder angle 0.0, as
    the reference does (its parameterised gates carry arg2 = 0.0, gate_create.cpp:27,84).
*/
namespace ket
{

void write_tangelo_circuit(const QuantumCircuit& circuit, std::ostream& stream, std::size_t n_leading_whitespace = 0);

void write_tangelo_circuit(const QuantumCircuit& circuit, const std::filesystem::path& filepath);

void print_tangelo_circuit(const QuantumCircuit& circuit);

}  // namespace ket
