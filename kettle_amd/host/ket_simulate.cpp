// ket::simulate for statevectors and density matrices, over kq_simulate (include/kq.h).
#include <cstring>
#include <stdexcept>
#include <vector>

#include "kettle/simulation/simulate.hpp"
#include "kettle/simulation/simulate_density_matrix.hpp"
#include "kettle/simulation/simulate_pauli.hpp"

#include "ket_internal.hpp"

namespace ket
{

namespace ki = ket::internal;

namespace
{

void check_valid_number_of_qubits(std::size_t circuit_n, std::size_t state_n)
{
    if (circuit_n != state_n) {
        throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
    }
    if (circuit_n == 0) {
        throw std::runtime_error {"Cannot simulate a circuit or state with zero qubits."};
    }
}

auto to_register(const std::vector<int8_t>& cells, std::size_t n_bits) -> ClassicalRegister
{
    auto reg = ClassicalRegister {n_bits};
    for (std::size_t i = 0; i < n_bits; ++i) {
        if (cells[i] >= 0) {
            reg.set(i, cells[i]);
        }
    }
    return reg;
}

struct LoggerSink
{
    std::size_t n_qubits;
    std::vector<CircuitLogger> loggers;
};

// called by libkq at every logger element, in execution order
void on_logger(void* user, int32_t kind, const double* amps, const int8_t* creg, size_t n_bits)
{
    auto* sink = static_cast<LoggerSink*>(user);
    if (kind == KQ_OP_LOG_STATE) {
        const auto n_states = std::size_t {1} << sink->n_qubits;
        std::vector<std::complex<double>> coeffs(n_states);
        std::memcpy(static_cast<void*>(coeffs.data()), amps, n_states * sizeof(std::complex<double>));
        auto logger = StatevectorCircuitLogger {};
        logger.add_statevector(Statevector {std::move(coeffs), Endian::LITTLE, 1.0e300});
        sink->loggers.emplace_back(std::move(logger));
    }
    else {
        auto logger = ClassicalRegisterCircuitLogger {};
        logger.add_classical_register(to_register(std::vector<int8_t>(creg, creg + n_bits), n_bits));
        sink->loggers.emplace_back(std::move(logger));
    }
}

}  // namespace

// StatevectorSimulator::run (simulate.cpp:539-560): flatten, then one kq_simulate call; gates
// between measurements, control flow and loggers are fused on the device.
void StatevectorSimulator::run(const QuantumCircuit& circuit, Statevector& state, std::optional<int> prng_seed)
{
    check_valid_number_of_qubits(circuit.n_qubits(), state.n_qubits());
    const auto flat = circuit.flatten();
    std::vector<int8_t> creg(std::max<std::size_t>(circuit.n_bits(), 1), -1);
    auto sink = LoggerSink {state.n_qubits(), {}};
    kq_state* dev = state.device_state();
    const int status = kq_simulate(dev, static_cast<uint32_t>(circuit.n_qubits()), flat.ops.data(), flat.ops.size(),
                                   flat.predicates.data(), flat.predicates.size(), prng_seed ? 1 : 0,
                                   prng_seed.value_or(0), creg.data(), circuit.n_bits(), on_logger, &sink);
    state.mark_device_modified();
    ki::kq_check(status);
    cregister_ = to_register(creg, circuit.n_bits());
    circuit_loggers_ = std::move(sink.loggers);
    has_been_run_ = true;
}

auto StatevectorSimulator::has_been_run() const -> bool { return has_been_run_; }

auto StatevectorSimulator::classical_register() const -> const ClassicalRegister&
{
    if (!cregister_) {
        throw std::runtime_error {"ERROR: Cannot access classical register; no simulation has been run\n"};
    }
    return *cregister_;
}

auto StatevectorSimulator::classical_register() -> ClassicalRegister&
{
    if (!cregister_) {
        throw std::runtime_error {"ERROR: Cannot access classical register; no simulation has been run\n"};
    }
    return *cregister_;
}

auto StatevectorSimulator::circuit_loggers() const -> const std::vector<CircuitLogger>& { return circuit_loggers_; }

void simulate(const QuantumCircuit& circuit, Statevector& state, std::optional<int> prng_seed)
{
    auto simulator = StatevectorSimulator {};
    simulator.run(circuit, state, prng_seed);
}

// ---- density matrix (simulate_density_matrix.cpp:463-548) ---------------------------------------------

DensityMatrixSimulator::DensityMatrixSimulator(std::size_t n_qubits) : n_qubits_ {n_qubits}
{
    if (n_qubits == 0) {
        throw std::runtime_error {"ERROR: cannot perform a DensityMatrix simulation with 0 qubits.\n"};
    }
}

void DensityMatrixSimulator::run(const QuantumCircuit& circuit, DensityMatrix& state, std::optional<int> prng_seed)
{
    check_valid_number_of_qubits(circuit.n_qubits(), state.n_qubits());
    if (state.n_qubits() != n_qubits_) {
        throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
    }
    const auto flat = circuit.flatten();
    std::vector<int8_t> creg(std::max<std::size_t>(circuit.n_bits(), 1), -1);
    kq_state* dev = state.device_state();
    const int status = kq_simulate(dev, static_cast<uint32_t>(circuit.n_qubits()), flat.ops.data(), flat.ops.size(),
                                   flat.predicates.data(), flat.predicates.size(), prng_seed ? 1 : 0,
                                   prng_seed.value_or(0), creg.data(), circuit.n_bits(), nullptr, nullptr);
    state.mark_device_modified();
    ki::kq_check(status);
    cregister_ = to_register(creg, circuit.n_bits());
    has_been_run_ = true;
}

auto DensityMatrixSimulator::has_been_run() const -> bool { return has_been_run_; }

auto DensityMatrixSimulator::classical_register() const -> const ClassicalRegister&
{
    if (!cregister_) {
        throw std::runtime_error {"ERROR: Cannot access classical register; no simulation has been run\n"};
    }
    return *cregister_;
}

void simulate(const QuantumCircuit& circuit, DensityMatrix& state, std::optional<int> prng_seed)
{
    auto simulator = DensityMatrixSimulator {state.n_qubits()};
    simulator.run(circuit, state, prng_seed);
}

// ---- Pauli strings (simulate_pauli.cpp:104-128) ----------------------------------------------------------

void StatevectorPauliStringSimulator::run(const SparsePauliString& pauli_string, Statevector& state)
{
    if (pauli_string.n_qubits() != state.n_qubits()) {
        throw std::runtime_error {"Invalid simulation; SparsePauliString and state have different number of qubits."};
    }
    if (pauli_string.n_qubits() == 0) {
        throw std::runtime_error {"Cannot simulate a SparsePauliString or state with zero qubits."};
    }
    std::vector<uint32_t> qs;
    std::vector<uint8_t> ks;
    for (const auto& [q, t] : pauli_string.terms()) {
        qs.push_back(static_cast<uint32_t>(q));
        ks.push_back(static_cast<uint8_t>(t));
    }
    kq_state* dev = state.device_state();
    const int status = kq_apply_pauli_string(dev, qs.size(), qs.data(), ks.data());
    state.mark_device_modified();
    ki::kq_check(status);
    has_been_run_ = true;
}

void simulate(const SparsePauliString& pauli_string, Statevector& state)
{
    auto simulator = StatevectorPauliStringSimulator {};
    simulator.run(pauli_string, state);
}

}  // namespace ket
