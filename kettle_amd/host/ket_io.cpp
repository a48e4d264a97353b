// Hamiltonian text reader (io/read_pauli_operator.cpp:29-69 in the reference): the PauliOperator
// that feeds expectation_value() and energy_and_gradient().
#include <fstream>
#include <ios>
#include <sstream>
#include <stdexcept>
#include <string>

#include "kettle/io/read_pauli_operator.hpp"

namespace ket
{

namespace
{

auto pauli_term_of(char c) -> PauliTerm
{
    switch (c) {
        case 'X': return PauliTerm::X;
        case 'Y': return PauliTerm::Y;
        case 'Z': return PauliTerm::Z;
        default: break;
    }
    throw std::out_of_range {std::string {"ERROR: unknown Pauli term '"} + c + "'\n"};
}

}  // namespace

auto read_pauli_operator(std::istream& instream, std::size_t n_qubits) -> ket::PauliOperator
{
    auto op = PauliOperator {n_qubits};
    std::string line;
    while (std::getline(instream, line)) {
        std::istringstream in {line};
        double re = 0.0;
        double im = 0.0;
        in >> re >> im;
        char sep = 0;
        in >> sep;   // ':'
        auto pauli_string = SparsePauliString {n_qubits};
        char c = 0;
        while (in >> c) {
            if (c != '(') {
                continue;
            }
            std::size_t qubit = 0;
            char comma = 0;
            char term = 0;
            char close = 0;
            in >> qubit >> comma >> term >> close;
            pauli_string.add(qubit, pauli_term_of(term));
        }
        op.add(std::complex<double> {re, im}, std::move(pauli_string));
    }
    return op;
}

auto read_pauli_operator(const std::filesystem::path& filepath, std::size_t n_qubits) -> ket::PauliOperator
{
    auto in = std::ifstream {filepath};
    if (!in.is_open()) {
        throw std::ios::failure {"ERROR: unable to open: " + filepath.string()};
    }
    return read_pauli_operator(in, n_qubits);
}

}  // namespace ket
