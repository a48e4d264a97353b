// n_local ansatz builder (optimize/n_local.cpp:57-305 in the reference, primitive gates only) and
// the adjoint energy gradient over kq_energy_gradient.
#include <stdexcept>
#include <vector>

#include "kettle/gradient/energy_gradient.hpp"
#include "kettle/optimize/n_local.hpp"

#include "ket_internal.hpp"

namespace ket
{

namespace ki = ket::internal;

namespace
{

auto is_1t(Gate g) -> bool { return g <= Gate::SXDAG; }
auto is_1t1a(Gate g) -> bool { return g >= Gate::RX && g <= Gate::P; }
auto is_1c1t(Gate g) -> bool { return g >= Gate::CH && g <= Gate::CSXDAG; }
auto is_1c1t1a(Gate g) -> bool { return g >= Gate::CRX && g <= Gate::CP; }

void add_1t(QuantumCircuit& c, Gate g, std::size_t t)
{
    switch (g) {
        case Gate::H: c.add_h_gate(t); break;
        case Gate::X: c.add_x_gate(t); break;
        case Gate::Y: c.add_y_gate(t); break;
        case Gate::Z: c.add_z_gate(t); break;
        case Gate::S: c.add_s_gate(t); break;
        case Gate::SDAG: c.add_sdag_gate(t); break;
        case Gate::T: c.add_t_gate(t); break;
        case Gate::TDAG: c.add_tdag_gate(t); break;
        case Gate::SX: c.add_sx_gate(t); break;
        case Gate::SXDAG: c.add_sxdag_gate(t); break;
        default: throw std::runtime_error {"DEV ERROR: invalid `Gate` found when applying rotation gates.\n"};
    }
}

auto add_1t1a(QuantumCircuit& c, Gate g, std::size_t t) -> ket::param::ParameterID
{
    constexpr auto key = ket::param::parameterized {};
    switch (g) {
        case Gate::RX: return c.add_rx_gate(t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        case Gate::RY: return c.add_ry_gate(t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        case Gate::RZ: return c.add_rz_gate(t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        case Gate::P: return c.add_p_gate(t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        default: throw std::runtime_error {"DEV ERROR: invalid `Gate` found when applying rotation gates.\n"};
    }
}

void add_1c1t(QuantumCircuit& c, Gate g, std::size_t ctrl, std::size_t t)
{
    switch (g) {
        case Gate::CH: c.add_ch_gate(ctrl, t); break;
        case Gate::CX: c.add_cx_gate(ctrl, t); break;
        case Gate::CY: c.add_cy_gate(ctrl, t); break;
        case Gate::CZ: c.add_cz_gate(ctrl, t); break;
        case Gate::CS: c.add_cs_gate(ctrl, t); break;
        case Gate::CSDAG: c.add_csdag_gate(ctrl, t); break;
        case Gate::CT: c.add_ct_gate(ctrl, t); break;
        case Gate::CTDAG: c.add_ctdag_gate(ctrl, t); break;
        case Gate::CSX: c.add_csx_gate(ctrl, t); break;
        case Gate::CSXDAG: c.add_csxdag_gate(ctrl, t); break;
        default: throw std::runtime_error {"ERROR: entanglement gate must be controlled\n"};
    }
}

auto add_1c1t1a(QuantumCircuit& c, Gate g, std::size_t ctrl, std::size_t t) -> ket::param::ParameterID
{
    constexpr auto key = ket::param::parameterized {};
    switch (g) {
        case Gate::CRX: return c.add_crx_gate(ctrl, t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        case Gate::CRY: return c.add_cry_gate(ctrl, t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        case Gate::CRZ: return c.add_crz_gate(ctrl, t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        case Gate::CP: return c.add_cp_gate(ctrl, t, DEFAULT_NLOCAL_GATE_PARAMETER, key);
        default: throw std::runtime_error {"ERROR: entanglement gate must be controlled\n"};
    }
}

void rotations(QuantumCircuit& c, const std::vector<Gate>& blocks, std::vector<ket::param::ParameterID>& ids)
{
    const auto n = c.n_qubits();
    for (auto g : blocks) {
        if (is_1t(g)) {
            for (std::size_t q = 0; q < n; ++q) { add_1t(c, g, q); }
        }
        else if (is_1t1a(g)) {
            for (std::size_t q = 0; q < n; ++q) { ids.push_back(add_1t1a(c, g, q)); }
        }
        else if (is_1c1t(g)) {
            for (std::size_t i = 0; i < n / 2; ++i) { add_1c1t(c, g, 2 * i, 2 * i + 1); }
        }
        else if (is_1c1t1a(g)) {
            for (std::size_t i = 0; i < n / 2; ++i) { ids.push_back(add_1c1t1a(c, g, 2 * i, 2 * i + 1)); }
        }
        else {
            throw std::runtime_error {"DEV ERROR: invalid `Gate` found when applying rotation gates.\n"};
        }
    }
}

void entanglement(QuantumCircuit& c, const std::vector<Gate>& blocks, NLocalEntangelement kind,
                  std::vector<ket::param::ParameterID>& ids)
{
    const auto n = c.n_qubits();
    std::vector<std::pair<std::size_t, std::size_t>> pairs;
    if (kind == NLocalEntangelement::LINEAR) {
        for (std::size_t q = 0; q + 1 < n; ++q) { pairs.emplace_back(q, q + 1); }
    }
    else {
        for (std::size_t a = 0; a + 1 < n; ++a) {
            for (std::size_t b = a + 1; b < n; ++b) { pairs.emplace_back(a, b); }
        }
    }
    for (auto g : blocks) {
        for (const auto& [a, b] : pairs) {
            if (is_1c1t(g)) {
                add_1c1t(c, g, a, b);
            }
            else if (is_1c1t1a(g)) {
                ids.push_back(add_1c1t1a(c, g, a, b));
            }
            else {
                throw std::runtime_error {"ERROR: entanglement gate must be controlled\n"};
            }
        }
    }
}

}  // namespace

auto n_local(std::size_t n_qubits, const std::vector<Gate>& rotation_blocks, const std::vector<Gate>& entanglement_blocks,
             NLocalEntangelement entanglement_kind, std::size_t n_repetitions, SkipLastRotationLayerFlag flag)
    -> std::tuple<QuantumCircuit, std::vector<ket::param::ParameterID>>
{
    for (auto g : entanglement_blocks) {
        if (!is_1c1t(g) && !is_1c1t1a(g)) {
            throw std::runtime_error {"ERROR: entanglement gate must be controlled\n"};
        }
    }
    auto circuit = QuantumCircuit {n_qubits};
    auto ids = std::vector<ket::param::ParameterID> {};
    for (std::size_t i = 0; i < n_repetitions; ++i) {
        rotations(circuit, rotation_blocks, ids);
        entanglement(circuit, entanglement_blocks, entanglement_kind, ids);
    }
    if (flag == SkipLastRotationLayerFlag::FALSE) {
        rotations(circuit, rotation_blocks, ids);
    }
    return {std::move(circuit), std::move(ids)};
}

// ---- adjoint energy gradient -------------------------------------------------------------------------

auto energy_and_gradient(const QuantumCircuit& circuit, const std::vector<ket::param::ParameterID>& parameter_ids,
                         const PauliOperator& hamiltonian, const AdjointOptions& options) -> EnergyGradient
{
    if (hamiltonian.n_qubits() != circuit.n_qubits()) {
        throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
    }
    const auto flat = circuit.flatten(&parameter_ids);
    for (const auto& op : flat.ops) {
        if (op.gate >= KQ_M) {
            throw std::runtime_error {"ERROR: the adjoint gradient needs a circuit without measurements, control flow or loggers.\n"};
        }
    }
    std::vector<uint32_t> offsets {0}, qubits;
    std::vector<uint8_t> kinds, phases;
    std::vector<double> coeffs;
    for (const auto& [c, s] : hamiltonian.weighted_pauli_strings()) {
        for (const auto& [q, t] : s.terms()) {
            qubits.push_back(static_cast<uint32_t>(q));
            kinds.push_back(static_cast<uint8_t>(t));
        }
        offsets.push_back(static_cast<uint32_t>(qubits.size()));
        phases.push_back(static_cast<uint8_t>(s.phase()));
        coeffs.push_back(c.real());
        coeffs.push_back(c.imag());
    }
    auto state = Statevector {circuit.n_qubits(), options.precision};
    auto out = EnergyGradient {0.0, std::vector<double>(parameter_ids.size(), 0.0), 0};
    kq_state* dev = state.device_state();
    ki::kq_check(kq_energy_gradient(dev, flat.ops.data(), flat.ops.size(), phases.size(), offsets.data(), qubits.data(),
                                    kinds.data(), phases.data(), coeffs.data(), parameter_ids.size(), &out.energy,
                                    out.gradient.data(), static_cast<uint32_t>(options.checkpoint_every),
                                    options.drift_tolerance, &out.checkpoint_restores));
    return out;
}

}  // namespace ket
