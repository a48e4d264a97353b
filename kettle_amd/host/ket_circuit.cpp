// QuantumCircuit builders and flattening (circuit.cpp in the reference: builder semantics,
// validation messages, parameter bookkeeping and the QFT / swap expansions).
#include <algorithm>
#include <cmath>
#include <sstream>
#include <stdexcept>
#include <string>

#include "kettle/circuit/circuit.hpp"
#include "kettle/parameter/parameter.hpp"

namespace ket
{

namespace
{

auto gate_name(Gate g) -> std::string_view
{
    static constexpr std::string_view names[] = {
        "H", "X", "Y", "Z", "S", "SDAG", "T", "TDAG", "SX", "SXDAG", "RX", "RY", "RZ", "P",
        "CH", "CX", "CY", "CZ", "CS", "CSDAG", "CT", "CTDAG", "CSX", "CSXDAG", "CRX", "CRY", "CRZ", "CP",
        "U", "CU", "M"};
    return names[static_cast<int>(g)];
}

auto default_parameter_name(std::size_t n) -> std::string
{
    return "theta" + std::to_string(n);
}

auto make_gate(Gate g, std::size_t a0, std::size_t a1, double angle) -> GateInfo
{
    return GateInfo {g, a0, a1, angle, std::nullopt, std::nullopt};
}

}  // namespace

void QuantumCircuit::check_qubit_range_(std::size_t index, std::string_view qubit_name, std::string_view gname) const
{
    if (index >= n_qubits_) {
        auto err = std::stringstream {};
        err << "The index for the " << qubit_name << " at which the '" << gname << "' gate is applied, is out of bounds.\n";
        err << "n_qubits             = " << n_qubits_ << '\n';
        err << "provided qubit index = " << index << '\n';
        throw std::runtime_error {err.str()};
    }
}

void QuantumCircuit::check_bit_range_(std::size_t bit_index) const
{
    if (bit_index >= n_bits_) {
        auto err = std::stringstream {};
        err << "The bit index at which the 'M' gate is applied, is out of bounds.\n";
        err << "n_bits             = " << n_bits_ << '\n';
        err << "provided bit index = " << bit_index << '\n';
        throw std::runtime_error {err.str()};
    }
}

void QuantumCircuit::pop_back()
{
    if (elements_.empty()) {
        throw std::runtime_error {"ERROR: called `pop_back()` on a `QuantumCircuit` instance with no circuit elements.\n"};
    }
    elements_.pop_back();
}

void QuantumCircuit::set_parameter_value(const ket::param::ParameterID& id, double angle)
{
    if (!parameter_data_.contains(id)) {
        throw std::out_of_range {"ERROR: no parameter found with the provided id.\n"};
    }
    parameter_data_[id].value = angle;
}

auto QuantumCircuit::new_parameter_(std::optional<double> value, const ket::param::ParameterID* id) -> ket::param::ParameterID
{
    const auto name = default_parameter_name(parameter_count_++);
    const auto param = id ? ket::param::Parameter {name, *id} : ket::param::Parameter {name};
    parameter_data_[param.id()] = ket::param::ParameterData {value, param.name(), 1};
    parameter_order_.push_back(param.id());
    return param.id();
}

void QuantumCircuit::use_parameter_(const ket::param::ParameterID& id)
{
    if (parameter_data_.contains(id)) {
        ++parameter_data_.at(id).count;
    }
    else {
        new_parameter_(std::nullopt, &id);
    }
}

void QuantumCircuit::add_one_target_(Gate g, std::size_t target)
{
    check_qubit_range_(target, "qubit", gate_name(g));
    elements_.emplace_back(make_gate(g, target, 0, 0.0));
}

void QuantumCircuit::add_one_target_angle_(Gate g, std::size_t target, double angle)
{
    check_qubit_range_(target, "qubit", gate_name(g));
    elements_.emplace_back(make_gate(g, target, 0, angle));
}

auto QuantumCircuit::add_one_target_new_param_(Gate g, std::size_t target, double angle) -> ket::param::ParameterID
{
    check_qubit_range_(target, "qubit", gate_name(g));
    const auto id = new_parameter_(angle, nullptr);
    auto info = make_gate(g, target, 0, angle);
    info.parameter = id;
    elements_.emplace_back(std::move(info));
    return id;
}

void QuantumCircuit::add_one_target_param_(Gate g, std::size_t target, const ket::param::ParameterID& id)
{
    check_qubit_range_(target, "qubit", gate_name(g));
    use_parameter_(id);
    auto info = make_gate(g, target, 0, 0.0);
    info.parameter = id;
    elements_.emplace_back(std::move(info));
}

void QuantumCircuit::add_ctrl_(Gate g, std::size_t control, std::size_t target)
{
    check_qubit_range_(control, "control qubit", gate_name(g));
    check_qubit_range_(target, "target qubit", gate_name(g));
    elements_.emplace_back(make_gate(g, control, target, 0.0));
}

void QuantumCircuit::add_ctrl_angle_(Gate g, std::size_t control, std::size_t target, double angle)
{
    check_qubit_range_(control, "control qubit", gate_name(g));
    check_qubit_range_(target, "target qubit", gate_name(g));
    elements_.emplace_back(make_gate(g, control, target, angle));
}

auto QuantumCircuit::add_ctrl_new_param_(Gate g, std::size_t control, std::size_t target, double angle) -> ket::param::ParameterID
{
    check_qubit_range_(control, "control qubit", gate_name(g));
    check_qubit_range_(target, "target qubit", gate_name(g));
    const auto id = new_parameter_(angle, nullptr);
    auto info = make_gate(g, control, target, angle);
    info.parameter = id;
    elements_.emplace_back(std::move(info));
    return id;
}

void QuantumCircuit::add_ctrl_param_(Gate g, std::size_t control, std::size_t target, const ket::param::ParameterID& id)
{
    check_qubit_range_(control, "control qubit", gate_name(g));
    check_qubit_range_(target, "target qubit", gate_name(g));
    use_parameter_(id);
    auto info = make_gate(g, control, target, 0.0);
    info.parameter = id;
    elements_.emplace_back(std::move(info));
}

void QuantumCircuit::add_h_gate(std::size_t t) { add_one_target_(Gate::H, t); }
void QuantumCircuit::add_x_gate(std::size_t t) { add_one_target_(Gate::X, t); }
void QuantumCircuit::add_y_gate(std::size_t t) { add_one_target_(Gate::Y, t); }
void QuantumCircuit::add_z_gate(std::size_t t) { add_one_target_(Gate::Z, t); }
void QuantumCircuit::add_s_gate(std::size_t t) { add_one_target_(Gate::S, t); }
void QuantumCircuit::add_sdag_gate(std::size_t t) { add_one_target_(Gate::SDAG, t); }
void QuantumCircuit::add_t_gate(std::size_t t) { add_one_target_(Gate::T, t); }
void QuantumCircuit::add_tdag_gate(std::size_t t) { add_one_target_(Gate::TDAG, t); }
void QuantumCircuit::add_sx_gate(std::size_t t) { add_one_target_(Gate::SX, t); }
void QuantumCircuit::add_sxdag_gate(std::size_t t) { add_one_target_(Gate::SXDAG, t); }

void QuantumCircuit::add_rx_gate(std::size_t t, double a) { add_one_target_angle_(Gate::RX, t, a); }
void QuantumCircuit::add_ry_gate(std::size_t t, double a) { add_one_target_angle_(Gate::RY, t, a); }
void QuantumCircuit::add_rz_gate(std::size_t t, double a) { add_one_target_angle_(Gate::RZ, t, a); }
void QuantumCircuit::add_p_gate(std::size_t t, double a) { add_one_target_angle_(Gate::P, t, a); }
auto QuantumCircuit::add_rx_gate(std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_one_target_new_param_(Gate::RX, t, a); }
auto QuantumCircuit::add_ry_gate(std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_one_target_new_param_(Gate::RY, t, a); }
auto QuantumCircuit::add_rz_gate(std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_one_target_new_param_(Gate::RZ, t, a); }
auto QuantumCircuit::add_p_gate(std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_one_target_new_param_(Gate::P, t, a); }
void QuantumCircuit::add_rx_gate(std::size_t t, const ket::param::ParameterID& id) { add_one_target_param_(Gate::RX, t, id); }
void QuantumCircuit::add_ry_gate(std::size_t t, const ket::param::ParameterID& id) { add_one_target_param_(Gate::RY, t, id); }
void QuantumCircuit::add_rz_gate(std::size_t t, const ket::param::ParameterID& id) { add_one_target_param_(Gate::RZ, t, id); }
void QuantumCircuit::add_p_gate(std::size_t t, const ket::param::ParameterID& id) { add_one_target_param_(Gate::P, t, id); }

void QuantumCircuit::add_ch_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CH, c, t); }
void QuantumCircuit::add_cx_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CX, c, t); }
void QuantumCircuit::add_cy_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CY, c, t); }
void QuantumCircuit::add_cz_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CZ, c, t); }
void QuantumCircuit::add_cs_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CS, c, t); }
void QuantumCircuit::add_csdag_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CSDAG, c, t); }
void QuantumCircuit::add_ct_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CT, c, t); }
void QuantumCircuit::add_ctdag_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CTDAG, c, t); }
void QuantumCircuit::add_csx_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CSX, c, t); }
void QuantumCircuit::add_csxdag_gate(std::size_t c, std::size_t t) { add_ctrl_(Gate::CSXDAG, c, t); }

void QuantumCircuit::add_crx_gate(std::size_t c, std::size_t t, double a) { add_ctrl_angle_(Gate::CRX, c, t, a); }
void QuantumCircuit::add_cry_gate(std::size_t c, std::size_t t, double a) { add_ctrl_angle_(Gate::CRY, c, t, a); }
void QuantumCircuit::add_crz_gate(std::size_t c, std::size_t t, double a) { add_ctrl_angle_(Gate::CRZ, c, t, a); }
void QuantumCircuit::add_cp_gate(std::size_t c, std::size_t t, double a) { add_ctrl_angle_(Gate::CP, c, t, a); }
auto QuantumCircuit::add_crx_gate(std::size_t c, std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_ctrl_new_param_(Gate::CRX, c, t, a); }
auto QuantumCircuit::add_cry_gate(std::size_t c, std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_ctrl_new_param_(Gate::CRY, c, t, a); }
auto QuantumCircuit::add_crz_gate(std::size_t c, std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_ctrl_new_param_(Gate::CRZ, c, t, a); }
auto QuantumCircuit::add_cp_gate(std::size_t c, std::size_t t, double a, ket::param::parameterized) -> ket::param::ParameterID { return add_ctrl_new_param_(Gate::CP, c, t, a); }
void QuantumCircuit::add_crx_gate(std::size_t c, std::size_t t, const ket::param::ParameterID& id) { add_ctrl_param_(Gate::CRX, c, t, id); }
void QuantumCircuit::add_cry_gate(std::size_t c, std::size_t t, const ket::param::ParameterID& id) { add_ctrl_param_(Gate::CRY, c, t, id); }
void QuantumCircuit::add_crz_gate(std::size_t c, std::size_t t, const ket::param::ParameterID& id) { add_ctrl_param_(Gate::CRZ, c, t, id); }
void QuantumCircuit::add_cp_gate(std::size_t c, std::size_t t, const ket::param::ParameterID& id) { add_ctrl_param_(Gate::CP, c, t, id); }

void QuantumCircuit::add_u_gate(const Matrix2X2& gate, std::size_t target)
{
    check_qubit_range_(target, "qubit", "U");
    auto info = make_gate(Gate::U, target, 0, 0.0);
    info.unitary = gate;
    elements_.emplace_back(std::move(info));
}

void QuantumCircuit::add_cu_gate(const Matrix2X2& gate, std::size_t control, std::size_t target)
{
    check_qubit_range_(control, "control qubit", "CU");
    check_qubit_range_(target, "target qubit", "CU");
    auto info = make_gate(Gate::CU, control, target, 0.0);
    info.unitary = gate;
    elements_.emplace_back(std::move(info));
}

void QuantumCircuit::add_m_gate(std::size_t target) { add_m_gate(target, target); }

void QuantumCircuit::add_m_gate(std::size_t target, std::size_t bit)
{
    check_qubit_range_(target, "qubit", "M");
    check_bit_range_(bit);
    elements_.emplace_back(make_gate(Gate::M, target, bit, 0.0));
}

void QuantumCircuit::add_swap_gate(std::size_t t0, std::size_t t1)
{
    if (t0 == t1) {
        throw std::runtime_error {"Cannot swap a index with itself"};
    }
    add_cx_gate(t0, t1);
    add_cx_gate(t1, t0);
    add_cx_gate(t0, t1);
}

void QuantumCircuit::add_qft_gate(const std::vector<std::size_t>& indices)
{
    const auto size = indices.size();
    for (std::size_t i = 0; i < size; ++i) {
        add_h_gate(indices[i]);
        auto denom = std::size_t {2};
        for (std::size_t c = i + 1; c < size; ++c) {
            add_cp_gate(indices[c], indices[i], 2.0 * M_PI / static_cast<double>(std::size_t {1} << denom));
            ++denom;
        }
    }
    if (size > 1) {
        for (std::size_t l = 0, r = size - 1; r > l; ++l, --r) {
            add_swap_gate(indices[l], indices[r]);
        }
    }
}

void QuantumCircuit::add_iqft_gate(const std::vector<std::size_t>& indices)
{
    const auto size = indices.size();
    if (size > 1) {
        for (std::size_t l = 0, r = size - 1; r > l; ++l, --r) {
            add_swap_gate(indices[l], indices[r]);
        }
    }
    for (std::size_t ii = size; ii-- > 0;) {
        auto denom = size - ii;
        for (std::size_t c = size; c-- > ii + 1;) {
            add_cp_gate(indices[c], indices[ii], -2.0 * M_PI / static_cast<double>(std::size_t {1} << denom));
            --denom;
        }
        add_h_gate(indices[ii]);
    }
}

void QuantumCircuit::merge_subcircuit_parameters_(const QuantumCircuit& sub, double tolerance)
{
    for (const auto& id : sub.parameter_order_) {
        const auto& data = sub.parameter_data_.at(id);
        auto [it, inserted] = parameter_data_.insert({id, data});
        if (inserted) {
            parameter_order_.push_back(id);
            continue;
        }
        auto& cur = it->second;
        cur.count += data.count;
        if (!cur.value && data.value) {
            cur.value = data.value;
        }
        else if (cur.value && data.value && std::fabs(*cur.value - *data.value) > tolerance) {
            throw std::runtime_error {"ERROR: found two parameter instances with same id but different non-optional values\n"};
        }
    }
}

void QuantumCircuit::add_if_statement(ControlFlowPredicate predicate, QuantumCircuit circuit, double tolerance)
{
    for (auto b : predicate.bit_indices_to_check()) {
        check_bit_range_(b);
    }
    merge_subcircuit_parameters_(circuit, tolerance);
    elements_.emplace_back(ClassicalIfStatement {std::move(predicate), std::make_shared<const QuantumCircuit>(std::move(circuit))});
}

void QuantumCircuit::add_if_statement(std::size_t bit, QuantumCircuit sub)
{
    add_if_statement(ControlFlowPredicate {{bit}, {1}, ControlFlowBooleanKind::IF}, std::move(sub));
}

void QuantumCircuit::add_if_not_statement(std::size_t bit, QuantumCircuit sub)
{
    add_if_statement(ControlFlowPredicate {{bit}, {0}, ControlFlowBooleanKind::IF}, std::move(sub));
}

void QuantumCircuit::add_if_else_statement(ControlFlowPredicate predicate, QuantumCircuit if_sub, QuantumCircuit else_sub)
{
    for (auto b : predicate.bit_indices_to_check()) {
        check_bit_range_(b);
    }
    merge_subcircuit_parameters_(if_sub, MATCHING_PARAMETER_VALUE_TOLERANCE);
    merge_subcircuit_parameters_(else_sub, MATCHING_PARAMETER_VALUE_TOLERANCE);
    elements_.emplace_back(ClassicalIfElseStatement {std::move(predicate), std::make_shared<const QuantumCircuit>(std::move(if_sub)),
                                                     std::make_shared<const QuantumCircuit>(std::move(else_sub))});
}

void QuantumCircuit::add_if_else_statement(std::size_t bit, QuantumCircuit if_sub, QuantumCircuit else_sub)
{
    add_if_else_statement(ControlFlowPredicate {{bit}, {1}, ControlFlowBooleanKind::IF}, std::move(if_sub), std::move(else_sub));
}

void QuantumCircuit::add_if_not_else_statement(std::size_t bit, QuantumCircuit if_sub, QuantumCircuit else_sub)
{
    add_if_else_statement(ControlFlowPredicate {{bit}, {0}, ControlFlowBooleanKind::IF}, std::move(if_sub), std::move(else_sub));
}

void QuantumCircuit::add_classical_register_circuit_logger() { elements_.emplace_back(CircuitLoggerKind::CLASSICAL_REGISTER); }
void QuantumCircuit::add_statevector_circuit_logger() { elements_.emplace_back(CircuitLoggerKind::STATEVECTOR); }

// ---- flattening ---------------------------------------------------------------------------------------

auto QuantumCircuit::flatten(const std::vector<ket::param::ParameterID>* order) const -> FlatCircuit
{
    // create_parameter_values_map (parameter_expression.cpp:84-97): every parameter needs a value
    for (const auto& [id, data] : parameter_data_) {
        if (!data.value) {
            throw std::runtime_error {"ERROR: cannot perform simulation with an uninitialized parameter value.\n"};
        }
    }
    FlatCircuit out;
    out.parameter_order = order ? *order : parameter_order_;
    flatten_into_(out, parameter_data_, out.parameter_order);
    return out;
}

void QuantumCircuit::flatten_into_(FlatCircuit& out, const ket::param::ParameterDataMap& values,
                                   const std::vector<ket::param::ParameterID>& order) const
{
    for (const auto& el : elements_) {
        if (el.is_gate()) {
            const auto& g = el.get_gate();
            kq_op op {};
            op.gate = static_cast<int32_t>(g.gate);
            op.q0 = static_cast<int32_t>(g.arg0);
            op.q1 = static_cast<int32_t>(g.arg1);
            op.aux = -1;
            op.angle = g.arg2;
            if (g.parameter) {
                const auto it = values.find(*g.parameter);
                if (it == values.end() || !it->second.value) {
                    throw std::runtime_error {"DEV ERROR: unable to find parameter during expression evaluation.\n"};
                }
                op.angle = *it->second.value;
                const auto pos = std::find(order.begin(), order.end(), *g.parameter);
                op.aux = pos == order.end() ? -1 : static_cast<int32_t>(pos - order.begin());
            }
            if (g.unitary) {
                const auto& m = *g.unitary;
                const std::complex<double> e[4] = {m.elem00, m.elem01, m.elem10, m.elem11};
                for (int i = 0; i < 4; ++i) {
                    op.u[2 * i] = e[i].real();
                    op.u[2 * i + 1] = e[i].imag();
                }
            }
            out.ops.push_back(op);
        }
        else if (el.is_control_flow()) {
            const ControlFlowPredicate& p = el.is_if_statement() ? el.get_if_statement().predicate
                                                                 : el.get_if_else_statement().predicate;
            const auto off = static_cast<int32_t>(out.predicates.size());
            out.predicates.push_back(p.control_kind() == ControlFlowBooleanKind::IF ? 0 : 1);
            out.predicates.push_back(static_cast<int32_t>(p.bit_indices_to_check().size()));
            for (std::size_t i = 0; i < p.bit_indices_to_check().size(); ++i) {
                out.predicates.push_back(static_cast<int32_t>(p.bit_indices_to_check()[i]));
                out.predicates.push_back(p.expected_bits()[i]);
            }
            const auto head = out.ops.size();
            out.ops.push_back(kq_op {});
            if (el.is_if_statement()) {
                el.get_if_statement().circuit->flatten_into_(out, values, order);
                auto& h = out.ops[head];
                h.gate = KQ_OP_IF;
                h.q0 = off;
                h.q1 = static_cast<int32_t>(out.ops.size() - head - 1);
            }
            else {
                const auto& s = el.get_if_else_statement();
                s.if_circuit->flatten_into_(out, values, order);
                const auto len_if = out.ops.size() - head - 1;
                s.else_circuit->flatten_into_(out, values, order);
                auto& h = out.ops[head];
                h.gate = KQ_OP_IF_ELSE;
                h.q0 = off;
                h.q1 = static_cast<int32_t>(len_if);
                h.aux = static_cast<int32_t>(out.ops.size() - head - 1 - len_if);
            }
        }
        else {
            kq_op op {};
            op.gate = el.get_circuit_logger() == CircuitLoggerKind::STATEVECTOR ? KQ_OP_LOG_STATE : KQ_OP_LOG_CREG;
            out.ops.push_back(op);
        }
    }
}

void extend_circuit(QuantumCircuit& left, const QuantumCircuit& right)
{
    if (left.n_qubits() != right.n_qubits()) {
        throw std::runtime_error {"ERROR: cannot extend circuits with different numbers of qubits.\n"};
    }
    left.merge_subcircuit_parameters_(right, MATCHING_PARAMETER_VALUE_TOLERANCE);
    left.elements_.insert(left.elements_.end(), right.elements_.begin(), right.elements_.end());
}

auto append_circuits(QuantumCircuit left, const QuantumCircuit& right) -> QuantumCircuit
{
    extend_circuit(left, right);
    return left;
}

}  // namespace ket
