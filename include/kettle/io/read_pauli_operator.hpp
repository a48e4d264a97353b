#pragma once

#include <cstddef>
#include <filesystem>
#include <istream>

#include "kettle/operator/pauli/pauli_operator.hpp"

/*
    kettle's Hamiltonian text reader (include/kettle/io/read_pauli_operator.hpp and
    source/kettle_internal/io/read_pauli_operator.cpp:29-69 in the reference).  One weighted
    Pauli string per line:

        <re> <im> : (q, P) (q, P) ...      P in {X, Y, Z}; no terms = the identity string
*/
namespace ket
{

auto read_pauli_operator(std::istream& instream, std::size_t n_qubits) -> ket::PauliOperator;

auto read_pauli_operator(const std::filesystem::path& filepath, std::size_t n_qubits) -> ket::PauliOperator;

}  // namespace ket
