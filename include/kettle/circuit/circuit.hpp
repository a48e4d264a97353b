#pragma once

#include <cstddef>
#include <ranges>
#include <initializer_list>
#include <memory>
#include <string_view>
#include <tuple>
#include <utility>
#include <variant>
#include <vector>

#include "kettle/circuit/classical_register.hpp"
#include "kettle/circuit/control_flow_predicate.hpp"
#include "kettle/common/matrix2x2.hpp"
#include "kettle/common/tolerance.hpp"
#include "kettle/gates/primitive_gate.hpp"
#include "kettle/parameter/parameter.hpp"

#include "kq.h"

/*
    kettle's circuit builder (include/kettle/circuit/circuit.hpp:20-419 and
    source/kettle_internal/circuit/circuit.cpp in the reference): the same builder names, index
    conventions (1-qubit gates take the target, controlled gates (control, target), M takes
    (qubit, bit)), validation messages and parameter bookkeeping.  `flatten()` turns the
    element tree into libkq's flat op array (include/kq.h).
*/
namespace ket
{

class QuantumCircuit;

struct ClassicalIfStatement
{
    ControlFlowPredicate predicate;
    std::shared_ptr<const QuantumCircuit> circuit;
};

struct ClassicalIfElseStatement
{
    ControlFlowPredicate predicate;
    std::shared_ptr<const QuantumCircuit> if_circuit;
    std::shared_ptr<const QuantumCircuit> else_circuit;
};

enum class CircuitLoggerKind : std::uint8_t
{
    STATEVECTOR,
    CLASSICAL_REGISTER
};

class CircuitElement
{
public:
    CircuitElement(GateInfo gate) : element_ {std::move(gate)} {}                        // NOLINT
    CircuitElement(ClassicalIfStatement s) : element_ {std::move(s)} {}                  // NOLINT
    CircuitElement(ClassicalIfElseStatement s) : element_ {std::move(s)} {}              // NOLINT
    CircuitElement(CircuitLoggerKind k) : element_ {k} {}                                // NOLINT

    [[nodiscard]] auto is_gate() const -> bool { return std::holds_alternative<GateInfo>(element_); }
    [[nodiscard]] auto is_if_statement() const -> bool { return std::holds_alternative<ClassicalIfStatement>(element_); }
    [[nodiscard]] auto is_if_else_statement() const -> bool { return std::holds_alternative<ClassicalIfElseStatement>(element_); }
    [[nodiscard]] auto is_control_flow() const -> bool { return is_if_statement() || is_if_else_statement(); }
    [[nodiscard]] auto is_circuit_logger() const -> bool { return std::holds_alternative<CircuitLoggerKind>(element_); }
    [[nodiscard]] auto get_gate() const -> const GateInfo& { return std::get<GateInfo>(element_); }
    [[nodiscard]] auto get_if_statement() const -> const ClassicalIfStatement& { return std::get<ClassicalIfStatement>(element_); }
    [[nodiscard]] auto get_if_else_statement() const -> const ClassicalIfElseStatement& { return std::get<ClassicalIfElseStatement>(element_); }
    [[nodiscard]] auto get_circuit_logger() const -> CircuitLoggerKind { return std::get<CircuitLoggerKind>(element_); }

private:
    std::variant<GateInfo, ClassicalIfStatement, ClassicalIfElseStatement, CircuitLoggerKind> element_;
};

// A circuit flattened for libkq: ops, predicate table and the parameter slot order.
struct FlatCircuit
{
    std::vector<kq_op> ops;
    std::vector<std::int32_t> predicates;
    std::vector<ket::param::ParameterID> parameter_order;
};

class QuantumCircuit
{
public:
    explicit QuantumCircuit(std::size_t n_qubits, std::size_t n_bits) : n_qubits_ {n_qubits}, n_bits_ {n_bits} {}
    explicit QuantumCircuit(std::size_t n_qubits) : n_qubits_ {n_qubits}, n_bits_ {n_qubits} {}

    [[nodiscard]] auto n_qubits() const noexcept -> std::size_t { return n_qubits_; }
    [[nodiscard]] auto n_bits() const noexcept -> std::size_t { return n_bits_; }
    [[nodiscard]] auto n_circuit_elements() const noexcept -> std::size_t { return elements_.size(); }
    [[nodiscard]] auto operator[](std::size_t i) const noexcept -> const CircuitElement& { return elements_[i]; }
    [[nodiscard]] auto begin() const noexcept { return elements_.begin(); }
    [[nodiscard]] auto end() const noexcept { return elements_.end(); }
    [[nodiscard]] auto circuit_elements() const noexcept -> const std::vector<CircuitElement>& { return elements_; }
    [[nodiscard]] auto parameter_data_map() const noexcept -> const ket::param::ParameterDataMap& { return parameter_data_; }
    // parameter ids in creation order (the default order of gradients)
    [[nodiscard]] auto parameter_ids() const -> const std::vector<ket::param::ParameterID>& { return parameter_order_; }

    void pop_back();
    void set_parameter_value(const ket::param::ParameterID& id, double angle);

    // 1-target gates
    void add_h_gate(std::size_t target);
    void add_x_gate(std::size_t target);
    void add_y_gate(std::size_t target);
    void add_z_gate(std::size_t target);
    void add_s_gate(std::size_t target);
    void add_sdag_gate(std::size_t target);
    void add_t_gate(std::size_t target);
    void add_tdag_gate(std::size_t target);
    void add_sx_gate(std::size_t target);
    void add_sxdag_gate(std::size_t target);
    template <std::ranges::range C> void add_h_gate(const C& ts) { for (auto t : ts) { add_h_gate(t); } }
    template <std::ranges::range C> void add_x_gate(const C& ts) { for (auto t : ts) { add_x_gate(t); } }
    template <std::ranges::range C> void add_y_gate(const C& ts) { for (auto t : ts) { add_y_gate(t); } }
    template <std::ranges::range C> void add_z_gate(const C& ts) { for (auto t : ts) { add_z_gate(t); } }
    template <std::ranges::range C> void add_s_gate(const C& ts) { for (auto t : ts) { add_s_gate(t); } }
    template <std::ranges::range C> void add_sdag_gate(const C& ts) { for (auto t : ts) { add_sdag_gate(t); } }
    template <std::ranges::range C> void add_t_gate(const C& ts) { for (auto t : ts) { add_t_gate(t); } }
    template <std::ranges::range C> void add_tdag_gate(const C& ts) { for (auto t : ts) { add_tdag_gate(t); } }
    template <std::ranges::range C> void add_sx_gate(const C& ts) { for (auto t : ts) { add_sx_gate(t); } }
    template <std::ranges::range C> void add_sxdag_gate(const C& ts) { for (auto t : ts) { add_sxdag_gate(t); } }
    void add_h_gate(std::initializer_list<std::size_t> ts) { for (auto t : ts) { add_h_gate(t); } }
    void add_x_gate(std::initializer_list<std::size_t> ts) { for (auto t : ts) { add_x_gate(t); } }

    // 1-target angle gates: literal angle, new parameter (returns its id), or existing parameter
    void add_rx_gate(std::size_t target, double angle);
    void add_ry_gate(std::size_t target, double angle);
    void add_rz_gate(std::size_t target, double angle);
    void add_p_gate(std::size_t target, double angle);
    auto add_rx_gate(std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    auto add_ry_gate(std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    auto add_rz_gate(std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    auto add_p_gate(std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    void add_rx_gate(std::size_t target, const ket::param::ParameterID& id);
    void add_ry_gate(std::size_t target, const ket::param::ParameterID& id);
    void add_rz_gate(std::size_t target, const ket::param::ParameterID& id);
    void add_p_gate(std::size_t target, const ket::param::ParameterID& id);

    // 1-control-1-target gates
    void add_ch_gate(std::size_t control, std::size_t target);
    void add_cx_gate(std::size_t control, std::size_t target);
    void add_cy_gate(std::size_t control, std::size_t target);
    void add_cz_gate(std::size_t control, std::size_t target);
    void add_cs_gate(std::size_t control, std::size_t target);
    void add_csdag_gate(std::size_t control, std::size_t target);
    void add_ct_gate(std::size_t control, std::size_t target);
    void add_ctdag_gate(std::size_t control, std::size_t target);
    void add_csx_gate(std::size_t control, std::size_t target);
    void add_csxdag_gate(std::size_t control, std::size_t target);
    void add_cx_gate(std::initializer_list<std::pair<std::size_t, std::size_t>> pairs)
    {
        for (auto [c, t] : pairs) { add_cx_gate(c, t); }
    }

    void add_crx_gate(std::size_t control, std::size_t target, double angle);
    void add_cry_gate(std::size_t control, std::size_t target, double angle);
    void add_crz_gate(std::size_t control, std::size_t target, double angle);
    void add_cp_gate(std::size_t control, std::size_t target, double angle);
    auto add_crx_gate(std::size_t control, std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    auto add_cry_gate(std::size_t control, std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    auto add_crz_gate(std::size_t control, std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    auto add_cp_gate(std::size_t control, std::size_t target, double initial_angle, ket::param::parameterized key) -> ket::param::ParameterID;
    void add_crx_gate(std::size_t control, std::size_t target, const ket::param::ParameterID& id);
    void add_cry_gate(std::size_t control, std::size_t target, const ket::param::ParameterID& id);
    void add_crz_gate(std::size_t control, std::size_t target, const ket::param::ParameterID& id);
    void add_cp_gate(std::size_t control, std::size_t target, const ket::param::ParameterID& id);

    void add_u_gate(const Matrix2X2& gate, std::size_t target);
    void add_cu_gate(const Matrix2X2& gate, std::size_t control, std::size_t target);

    // measurement (circuit.cpp:688-718)
    void add_m_gate(std::size_t target);
    void add_m_gate(std::size_t target, std::size_t bit);
    void add_m_gate(std::initializer_list<std::size_t> targets) { for (auto t : targets) { add_m_gate(t); } }

    // compound gates expanded to primitive gates (circuit.cpp:801-903)
    void add_swap_gate(std::size_t target0, std::size_t target1);
    void add_qft_gate(const std::vector<std::size_t>& indices);
    void add_iqft_gate(const std::vector<std::size_t>& indices);

    // control flow (circuit.cpp:905-985)
    void add_if_statement(ControlFlowPredicate predicate, QuantumCircuit circuit,
                          double tolerance = MATCHING_PARAMETER_VALUE_TOLERANCE);
    void add_if_statement(std::size_t bit_index, QuantumCircuit subcircuit);
    void add_if_not_statement(std::size_t bit_index, QuantumCircuit subcircuit);
    void add_if_else_statement(ControlFlowPredicate predicate, QuantumCircuit if_subcircuit, QuantumCircuit else_subcircuit);
    void add_if_else_statement(std::size_t bit_index, QuantumCircuit if_subcircuit, QuantumCircuit else_subcircuit);
    void add_if_not_else_statement(std::size_t bit_index, QuantumCircuit if_subcircuit, QuantumCircuit else_subcircuit);

    void add_classical_register_circuit_logger();
    void add_statevector_circuit_logger();

    // libkq form; `order` fixes the parameter slots (default: creation order).  Throws kettle's
    // error for an uninitialised parameter (parameter_expression.cpp:84-97).
    [[nodiscard]] auto flatten(const std::vector<ket::param::ParameterID>* order = nullptr) const -> FlatCircuit;

    friend void extend_circuit(QuantumCircuit& left, const QuantumCircuit& right);

private:
    std::size_t n_qubits_;
    std::size_t n_bits_;
    std::vector<CircuitElement> elements_;
    ket::param::ParameterDataMap parameter_data_;
    std::vector<ket::param::ParameterID> parameter_order_;
    std::size_t parameter_count_ {0};

    void check_qubit_range_(std::size_t index, std::string_view qubit_name, std::string_view gate_name) const;
    void check_bit_range_(std::size_t bit_index) const;
    void add_one_target_(Gate gate, std::size_t target);
    void add_one_target_angle_(Gate gate, std::size_t target, double angle);
    auto add_one_target_new_param_(Gate gate, std::size_t target, double angle) -> ket::param::ParameterID;
    void add_one_target_param_(Gate gate, std::size_t target, const ket::param::ParameterID& id);
    void add_ctrl_(Gate gate, std::size_t control, std::size_t target);
    void add_ctrl_angle_(Gate gate, std::size_t control, std::size_t target, double angle);
    auto add_ctrl_new_param_(Gate gate, std::size_t control, std::size_t target, double angle) -> ket::param::ParameterID;
    void add_ctrl_param_(Gate gate, std::size_t control, std::size_t target, const ket::param::ParameterID& id);
    auto new_parameter_(std::optional<double> value, const ket::param::ParameterID* id) -> ket::param::ParameterID;
    void use_parameter_(const ket::param::ParameterID& id);
    void merge_subcircuit_parameters_(const QuantumCircuit& subcircuit, double tolerance);
    void flatten_into_(FlatCircuit& out, const ket::param::ParameterDataMap& values,
                       const std::vector<ket::param::ParameterID>& order) const;
};

void extend_circuit(QuantumCircuit& left, const QuantumCircuit& right);
auto append_circuits(QuantumCircuit left, const QuantumCircuit& right) -> QuantumCircuit;

}  // namespace ket
