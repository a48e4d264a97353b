#pragma once

#include "kettle/operator/pauli/sparse_pauli_string.hpp"
#include "kettle/state/statevector.hpp"

// ket::simulate(SparsePauliString, Statevector&) (simulate_pauli.hpp:22 in the reference).
namespace ket
{

class StatevectorPauliStringSimulator
{
public:
    void run(const SparsePauliString& pauli_string, Statevector& state);
    [[nodiscard]] auto has_been_run() const -> bool { return has_been_run_; }

private:
    bool has_been_run_ {false};
};

void simulate(const SparsePauliString& pauli_string, Statevector& state);

}  // namespace ket
