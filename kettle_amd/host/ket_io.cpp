// Text formats at the boundary of the simulator: the Hamiltonian reader (io/read_pauli_operator.cpp:29-69
// in the reference), which feeds expectation_value() and energy_and_gradient(), and the tangelo-style
// circuit reader and writer (io/read_tangelo_file.cpp, io/write_tangelo_file.cpp, io/io_control_flow.cpp),
// and statevector snapshots (io/statevector.cpp, io/numpy_statevector.cpp).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <ios>
#include <sstream>
#include <stdexcept>
#include <string>

#include "kettle/io/numpy_statevector.hpp"
#include "kettle/io/read_pauli_operator.hpp"
#include "kettle/io/read_tangelo_file.hpp"
#include "kettle/io/statevector.hpp"
#include "kettle/io/write_tangelo_file.hpp"
#include "kettle/state/qubit_state_conversion.hpp"

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

// ---- tangelo-style circuits -------------------------------------------------------------------------

namespace
{

constexpr std::size_t CONTROL_FLOW_INDENT = 4;   // io_control_flow.hpp:19

constexpr const char* GATE_NAMES[] = {
    "H", "X", "Y", "Z", "S", "SDAG", "T", "TDAG", "SX", "SXDAG", "RX", "RY", "RZ", "P",
    "CH", "CX", "CY", "CZ", "CS", "CSDAG", "CT", "CTDAG", "CSX", "CSXDAG", "CRX", "CRY", "CRZ", "CP",
    "U", "CU", "M"};

auto gate_of_name(const std::string& name) -> Gate
{
    // tangelo's own names (read_tangelo_file.cpp:23-37)
    const std::string local = name == "CPHASE" ? "CP" : name == "CNOT" ? "CX" : name == "PHASE" ? "P" : name;
    for (int g = 0; g < static_cast<int>(std::size(GATE_NAMES)); ++g) {
        if (local == GATE_NAMES[g]) {
            return static_cast<Gate>(g);
        }
    }
    throw std::runtime_error {"Unknown gate found in `read_tangelo_file()` : " + local + "\n"};
}

// "<word> : [" then an index then the closing bracket, as the reference's stream extraction reads them
auto read_bracketed_index(std::istream& in) -> std::size_t
{
    std::string word;
    char ch = 0;
    std::size_t value = 0;
    in >> word >> word >> ch >> value >> ch;
    return value;
}

auto read_angle(std::istream& in) -> double
{
    std::string word;
    double value = 0.0;
    in >> word >> word >> value;
    return value;
}

auto read_complex(std::istream& in) -> std::complex<double>
{
    char ch = 0;
    double re = 0.0;
    double im = 0.0;
    in >> ch >> re >> ch >> im >> ch;
    return {re, im};
}

auto read_matrix(std::istream& stream) -> Matrix2X2
{
    std::string first;
    std::string second;
    std::getline(stream, first);
    std::getline(stream, second);
    auto a = std::istringstream {first};
    auto b = std::istringstream {second};
    Matrix2X2 m {};
    m.elem00 = read_complex(a);
    m.elem01 = read_complex(a);
    m.elem10 = read_complex(b);
    m.elem11 = read_complex(b);
    return m;
}

// "[i, j, ...]" (io_control_flow.cpp:75-98): an empty list is "[]"
template <typename Int>
auto read_index_list(std::istream& in) -> std::vector<Int>
{
    std::vector<Int> out;
    char ch = 0;
    in >> ch;
    if (in.peek() == ']') {
        return out;
    }
    std::string sep;
    do {
        Int v {};
        in >> v >> sep;
        out.push_back(v);
    } while (sep != "]" && in);
    return out;
}

void skip_to(std::istream& in, char delimiter)
{
    while (in.peek() != std::char_traits<char>::eof() && in.peek() != delimiter) {
        in.get();
    }
}

// "BITS[...] == [...]" or "!=" (io_control_flow.cpp:126-139)
auto read_predicate(std::istream& in) -> ControlFlowPredicate
{
    skip_to(in, '[');
    auto bits = read_index_list<std::size_t>(in);
    std::string sign;
    in >> sign;
    ControlFlowBooleanKind kind {};
    if (sign == "==") {
        kind = ControlFlowBooleanKind::IF;
    }
    else if (sign == "!=") {
        kind = ControlFlowBooleanKind::IF_NOT;
    }
    else {
        throw std::runtime_error {"Invalid comparison sign found; must be '==' or '!='\n"};
    }
    skip_to(in, '[');
    auto expected = read_index_list<int>(in);
    return ControlFlowPredicate {std::move(bits), std::move(expected), kind};
}

void add_one_target(QuantumCircuit& c, Gate g, std::size_t t)
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
        default: c.add_sxdag_gate(t); break;
    }
}

void add_one_target_angle(QuantumCircuit& c, Gate g, std::size_t t, double a)
{
    switch (g) {
        case Gate::RX: c.add_rx_gate(t, a); break;
        case Gate::RY: c.add_ry_gate(t, a); break;
        case Gate::RZ: c.add_rz_gate(t, a); break;
        default: c.add_p_gate(t, a); break;
    }
}

void add_controlled(QuantumCircuit& c, Gate g, std::size_t ctrl, std::size_t t)
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
        default: c.add_csxdag_gate(ctrl, t); break;
    }
}

void add_controlled_angle(QuantumCircuit& c, Gate g, std::size_t ctrl, std::size_t t, double a)
{
    switch (g) {
        case Gate::CRX: c.add_crx_gate(ctrl, t, a); break;
        case Gate::CRY: c.add_cry_gate(ctrl, t, a); break;
        case Gate::CRZ: c.add_crz_gate(ctrl, t, a); break;
        default: c.add_cp_gate(ctrl, t, a); break;
    }
}

auto gate_index(Gate g) -> int { return static_cast<int>(g); }

auto padded_name(Gate g) -> std::string
{
    std::string name = GATE_NAMES[gate_index(g)];
    if (name.size() < 10) {
        name.resize(10, ' ');
    }
    return name;
}

auto fixed(double x, const char* fmt) -> std::string
{
    char buf[400];
    std::snprintf(buf, sizeof buf, fmt, x);
    return buf;
}

// format_double_ (write_tangelo_file.cpp:24-32): "{: .17f}" cut to 16 characters
auto matrix_element(double x) -> std::string
{
    auto s = fixed(x, "% .17f");
    if (s.size() > 16) {
        s.resize(16);
    }
    return s;
}

auto matrix_lines(const Matrix2X2& m) -> std::string
{
    const auto pair = [](std::complex<double> z) {
        return "[" + matrix_element(z.real()) + ", " + matrix_element(z.imag()) + "]";
    };
    return "    " + pair(m.elem00) + "   " + pair(m.elem01) + "\n" + "    " + pair(m.elem10) + "   " + pair(m.elem11)
         + "\n";
}

auto index_list(const auto& values) -> std::string
{
    std::string out = "[";
    for (std::size_t i = 0; i < values.size(); ++i) {
        out += (i ? ", " : "") + std::to_string(values[i]);
    }
    return out + "]";
}

auto predicate_text(const ControlFlowPredicate& p) -> std::string
{
    return "IF BITS" + index_list(p.bit_indices_to_check())
         + (p.control_kind() == ControlFlowBooleanKind::IF ? " == " : " != ") + index_list(p.expected_bits());
}

}  // namespace

auto read_tangelo_circuit(std::size_t n_qubits, std::istream& stream, std::size_t n_skip_lines,
                          std::optional<std::size_t> line_starts_with_spaces) -> QuantumCircuit
{
    auto circuit = QuantumCircuit {n_qubits};
    std::string line;
    for (std::size_t i = 0; i < n_skip_lines; ++i) {
        std::getline(stream, line);
    }
    auto pos = stream.tellg();
    while (std::getline(stream, line)) {
        if (line_starts_with_spaces && line.rfind(std::string(*line_starts_with_spaces, ' '), 0) != 0) {
            stream.clear();
            stream.seekg(pos);
            return circuit;
        }
        auto in = std::istringstream {line};
        std::string name;
        in >> name;
        if (name.empty()) {
            continue;
        }
        if (name == "IF") {
            auto predicate = read_predicate(in);
            auto body = read_tangelo_circuit(n_qubits, stream, 0, CONTROL_FLOW_INDENT);
            circuit.add_if_statement(std::move(predicate), std::move(body));
            continue;
        }
        if (name == "ELSE") {
            const auto n = circuit.n_circuit_elements();
            if (n == 0 || !circuit[n - 1].is_if_statement()) {
                throw std::runtime_error {
                    "ERROR: encountered an 'ELSE' statement, but no previous matching 'IF' statement was found.\n"};
            }
            const auto stmt = circuit[n - 1].get_if_statement();
            circuit.pop_back();
            auto else_body = read_tangelo_circuit(n_qubits, stream, 0, CONTROL_FLOW_INDENT);
            circuit.add_if_else_statement(stmt.predicate, *stmt.circuit, std::move(else_body));
            continue;
        }
        if (name == "SWAP") {
            std::string word;
            char ch = 0;
            std::size_t q0 = 0;
            std::size_t q1 = 0;
            in >> word >> word >> ch >> q0 >> ch >> q1 >> ch;
            circuit.add_swap_gate(q0, q1);
            pos = stream.tellg();
            continue;
        }
        const Gate g = gate_of_name(name);
        const int k = gate_index(g);
        if (k <= gate_index(Gate::SXDAG)) {
            add_one_target(circuit, g, read_bracketed_index(in));
        }
        else if (k <= gate_index(Gate::P)) {
            const auto t = read_bracketed_index(in);
            add_one_target_angle(circuit, g, t, read_angle(in));
        }
        else if (k <= gate_index(Gate::CSXDAG)) {
            const auto t = read_bracketed_index(in);
            add_controlled(circuit, g, read_bracketed_index(in), t);
        }
        else if (k <= gate_index(Gate::CP)) {
            const auto t = read_bracketed_index(in);
            const auto ctrl = read_bracketed_index(in);
            add_controlled_angle(circuit, g, ctrl, t, read_angle(in));
        }
        else if (g == Gate::M) {
            const auto q = read_bracketed_index(in);
            circuit.add_m_gate(q, read_bracketed_index(in));
        }
        else if (g == Gate::U) {
            const auto t = read_bracketed_index(in);
            circuit.add_u_gate(read_matrix(stream), t);
        }
        else {
            const auto t = read_bracketed_index(in);
            const auto ctrl = read_bracketed_index(in);
            circuit.add_cu_gate(read_matrix(stream), ctrl, t);
        }
        pos = stream.tellg();
    }
    return circuit;
}

auto read_tangelo_circuit(std::size_t n_qubits, const std::filesystem::path& filepath, std::size_t n_skip_lines)
    -> QuantumCircuit
{
    auto in = std::ifstream {filepath};
    if (!in.is_open()) {
        throw std::ios::failure {"ERROR: unable to read tangelo circuit from : '" + filepath.string() + "'\n"};
    }
    return read_tangelo_circuit(n_qubits, in, n_skip_lines);
}

void write_tangelo_circuit(const QuantumCircuit& circuit, std::ostream& stream, std::size_t n_leading_whitespace)
{
    const auto indent = std::string(n_leading_whitespace, ' ');
    for (const auto& el : circuit) {
        if (el.is_circuit_logger()) {
            continue;
        }
        if (el.is_if_statement()) {
            const auto& s = el.get_if_statement();
            stream << predicate_text(s.predicate) << '\n';
            write_tangelo_circuit(*s.circuit, stream, CONTROL_FLOW_INDENT);
            continue;
        }
        if (el.is_if_else_statement()) {
            const auto& s = el.get_if_else_statement();
            stream << predicate_text(s.predicate) << '\n';
            write_tangelo_circuit(*s.if_circuit, stream, CONTROL_FLOW_INDENT);
            stream << "ELSE\n";
            write_tangelo_circuit(*s.else_circuit, stream, CONTROL_FLOW_INDENT);
            continue;
        }
        const GateInfo& g = el.get_gate();
        const int k = gate_index(g.gate);
        const double angle = g.parameter ? 0.0 : g.arg2;
        stream << indent << padded_name(g.gate);
        if (k <= gate_index(Gate::SXDAG)) {
            stream << "target : [" << g.arg0 << "]\n";
        }
        else if (k <= gate_index(Gate::P)) {
            stream << "target : [" << g.arg0 << "]   parameter : " << fixed(angle, "%.16f") << '\n';
        }
        else if (k <= gate_index(Gate::CSXDAG)) {
            stream << "target : [" << g.arg1 << "]   control : [" << g.arg0 << "]\n";
        }
        else if (k <= gate_index(Gate::CP)) {
            stream << "target : [" << g.arg1 << "]   control : [" << g.arg0 << "]   parameter : " << fixed(angle, "%.16f")
                   << '\n';
        }
        else if (g.gate == Gate::M) {
            stream << "target : [" << g.arg0 << "]   bit : [" << g.arg1 << "]\n";
        }
        else if (g.gate == Gate::U) {
            stream << "target : [" << g.arg0 << "]\n" << matrix_lines(*g.unitary);
        }
        else {
            stream << "target : [" << g.arg1 << "]   control : [" << g.arg0 << "]\n" << matrix_lines(*g.unitary);
        }
    }
}

void write_tangelo_circuit(const QuantumCircuit& circuit, const std::filesystem::path& filepath)
{
    auto out = std::ofstream {filepath};
    if (!out.is_open()) {
        throw std::ios::failure {"ERROR: unable to open file to write tangelo-style circuit: '" + filepath.string()
                                 + "'\n"};
    }
    write_tangelo_circuit(circuit, out);
}

void print_tangelo_circuit(const QuantumCircuit& circuit) { write_tangelo_circuit(circuit, std::cout); }

// ---- statevector snapshots ---------------------------------------------------------------------------

namespace
{

// "%.14f" with a space in front of non-negative values (statevector.cpp:45-66)
auto snapshot_number(double x) -> std::string
{
    char buf[400];
    std::snprintf(buf, sizeof buf, "%.14f", x);
    return (x >= 0.0 ? " " : "") + std::string {buf};
}

auto endian_name(Endian e) -> const char* { return e == Endian::BIG ? "BIG" : "LITTLE"; }

auto endian_of(const std::string& name) -> Endian
{
    if (name == "BIG") {
        return Endian::BIG;
    }
    if (name == "LITTLE") {
        return Endian::LITTLE;
    }
    throw std::runtime_error {"UNREACHABLE: dev error in `string_to_endian()`"};
}

}  // namespace

void save_statevector(std::ostream& outstream, const Statevector& state, Endian endian)
{
    outstream << "ENDIANNESS: " << endian_name(endian) << '\n';
    outstream << "NUMBER OF STATES: " << state.n_states() << '\n';
    const std::size_t n = state.n_qubits();
    std::string line;
    for (std::size_t i = 0; i < state.n_states(); ++i) {
        const auto& z = state[endian == Endian::LITTLE ? i : endian_flip(i, n)];
        line = snapshot_number(z.real());
        line += "   ";
        line += snapshot_number(z.imag());
        line += '\n';
        outstream << line;
    }
}

void save_statevector(const std::filesystem::path& filepath, const Statevector& state, Endian endian)
{
    auto out = std::ofstream {filepath};
    if (!out.is_open()) {
        throw std::ios::failure {"ERROR: unable to open file to save statevector: \n'" + filepath.string() + "'\n"};
    }
    save_statevector(out, state, endian);
}

auto load_statevector(std::istream& instream) -> Statevector
{
    std::string word;
    std::string endian;
    instream >> word >> endian;
    const Endian e = endian_of(endian);
    std::size_t n_states = 0;
    instream >> word >> word >> word >> n_states;
    std::vector<std::complex<double>> amps;
    amps.reserve(n_states);
    for (std::size_t i = 0; i < n_states; ++i) {
        double re = 0.0;
        double im = 0.0;
        instream >> re >> im;
        amps.emplace_back(re, im);
    }
    return Statevector {std::move(amps), e};
}

auto load_statevector(const std::filesystem::path& filepath) -> Statevector
{
    auto in = std::ifstream {filepath};
    if (!in.is_open()) {
        throw std::ios::failure {"ERROR: unable to open file to load statevector: \n'" + filepath.string() + "'\n"};
    }
    return load_statevector(in);
}

auto read_numpy_statevector(std::istream& instream, Endian input_endian) -> Statevector
{
    std::size_t n_qubits = 0;
    instream >> n_qubits;
    const std::size_t n_states = std::size_t {1} << n_qubits;
    std::vector<std::complex<double>> amps;
    amps.reserve(n_states);
    for (std::size_t i = 0; i < n_states; ++i) {
        // "(re+imj)": the sign of the imaginary part starts the second number
        char ch = 0;
        double re = 0.0;
        double im = 0.0;
        instream >> ch >> re >> im >> ch >> ch;
        amps.emplace_back(re, im);
    }
    return Statevector {std::move(amps), input_endian};
}

auto read_numpy_statevector(const std::filesystem::path& filepath, Endian input_endian) -> Statevector
{
    auto in = std::ifstream {filepath};
    if (!in.is_open()) {
        throw std::ios::failure {"ERROR: unable to open file for statevector: \n'" + filepath.string() + "'\n"};
    }
    return read_numpy_statevector(in, input_endian);
}

}  // namespace ket
