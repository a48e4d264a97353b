// Host-only tests of the kettle-compatible C++ API (no device needed): the Hamiltonian reader,
// Pauli-string bookkeeping, circuit builders, n_local and flattening.  They mirror the reference's
// test/source/io/read_pauli_operator_test.cpp, operator/pauli/sparse_pauli_string_test.cpp,
// circuit/circuit_test.cpp, optimize/n_local_test.cpp, io/{read,write}_tangelo_file_test.cpp,
// io/statevector_test.cpp and io/numpy_statevector_test.cpp.
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <kettle/kettle.hpp>

namespace
{

int g_failed = 0;
int g_passed = 0;

struct Failure : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

#define REQUIRE(cond)                                                                             \
    do {                                                                                          \
        if (!(cond)) {                                                                            \
            throw Failure {std::string {__FILE__} + ":" + std::to_string(__LINE__) + ": " #cond}; \
        }                                                                                         \
    } while (0)

template <typename E, typename F>
void require_throws(F&& f, const std::string& fragment = "")
{
    try {
        f();
    }
    catch (const E& e) {
        if (!fragment.empty() && std::string {e.what()}.find(fragment) == std::string::npos) {
            throw Failure {std::string {"wrong message: "} + e.what()};
        }
        return;
    }
    throw Failure {"expected exception not thrown"};
}

void run(const char* name, const std::function<void()>& f)
{
    std::printf("[case] %s\n", name);
    std::fflush(stdout);
    try {
        f();
        ++g_passed;
        std::printf("PASS %s\n", name);
    }
    catch (const std::exception& e) {
        ++g_failed;
        std::printf("FAIL %s: %s\n", name, e.what());
    }
    std::fflush(stdout);
}

using T = ket::PauliTerm;

auto same_matrix(const ket::Matrix2X2& a, const ket::Matrix2X2& b, double tol) -> bool
{
    return std::abs(a.elem00 - b.elem00) <= tol && std::abs(a.elem01 - b.elem01) <= tol
        && std::abs(a.elem10 - b.elem10) <= tol && std::abs(a.elem11 - b.elem11) <= tol;
}

auto same_predicate(const ket::ControlFlowPredicate& a, const ket::ControlFlowPredicate& b) -> bool
{
    return a.bit_indices_to_check() == b.bit_indices_to_check() && a.expected_bits() == b.expected_bits()
        && a.control_kind() == b.control_kind();
}

// element-by-element equality of two circuits (gates, arguments, angles and matrices within tol)
auto same_circuit(const ket::QuantumCircuit& a, const ket::QuantumCircuit& b, double tol = 1.0e-6) -> bool
{
    if (a.n_qubits() != b.n_qubits() || a.n_circuit_elements() != b.n_circuit_elements()) {
        return false;
    }
    for (std::size_t i = 0; i < a.n_circuit_elements(); ++i) {
        const auto& x = a[i];
        const auto& y = b[i];
        if (x.is_gate() != y.is_gate() || x.is_if_statement() != y.is_if_statement()
            || x.is_if_else_statement() != y.is_if_else_statement()) {
            return false;
        }
        if (x.is_gate()) {
            const auto& g = x.get_gate();
            const auto& h = y.get_gate();
            if (g.gate != h.gate || g.arg0 != h.arg0 || g.arg1 != h.arg1 || std::abs(g.arg2 - h.arg2) > tol
                || g.unitary.has_value() != h.unitary.has_value()
                || (g.unitary && !same_matrix(*g.unitary, *h.unitary, tol))) {
                return false;
            }
        }
        else if (x.is_if_statement()) {
            if (!same_predicate(x.get_if_statement().predicate, y.get_if_statement().predicate)
                || !same_circuit(*x.get_if_statement().circuit, *y.get_if_statement().circuit, tol)) {
                return false;
            }
        }
        else if (x.is_if_else_statement()) {
            const auto& s = x.get_if_else_statement();
            const auto& t = y.get_if_else_statement();
            if (!same_predicate(s.predicate, t.predicate) || !same_circuit(*s.if_circuit, *t.if_circuit, tol)
                || !same_circuit(*s.else_circuit, *t.else_circuit, tol)) {
                return false;
            }
        }
    }
    return true;
}

auto written(const ket::QuantumCircuit& c) -> std::string
{
    auto out = std::ostringstream {};
    ket::write_tangelo_circuit(c, out);
    return out.str();
}

}  // namespace

int main()
{
    run("read_pauli_operator known answer", [] {
        auto in = std::stringstream {
            " 1.100000000000e+01    0.000000000000e+00   :                                                         \n"
            "-1.875000000000e-01    0.000000000000e+00   :   (0, X)   (1, X)   (2, X)   (3, X)   (4, X)   (5, X)   \n"
            " 1.875000000000e-01    0.000000000000e+00   :   (0, X)   (1, X)   (2, X)   (3, Z)                     \n"};
        const auto op = ket::read_pauli_operator(in, 6);
        const auto expected = ket::PauliOperator {
            {{11.0, 0.0}, ket::SparsePauliString {T::I, T::I, T::I, T::I, T::I, T::I}},
            {{-0.1875, 0.0}, ket::SparsePauliString {T::X, T::X, T::X, T::X, T::X, T::X}},
            {{0.1875, 0.0}, ket::SparsePauliString {T::X, T::X, T::X, T::Z, T::I, T::I}},
        };
        REQUIRE(ket::almost_eq(op, expected));
        REQUIRE(op.n_qubits() == 6);
    });
    run("read_pauli_operator complex weights, Y terms, errors", [] {
        auto in = std::stringstream {"0.5 -0.25 : (2, Y) (0, Z)\n"};
        const auto op = ket::read_pauli_operator(in, 3);
        REQUIRE(op.size() == 1);
        REQUIRE(op.at(0).coefficient == std::complex<double>(0.5, -0.25));
        REQUIRE(op.at(0).pauli_string.at(2) == T::Y);
        REQUIRE(op.at(0).pauli_string.at(0) == T::Z);
        auto dup = std::stringstream {"1.0 0.0 : (1, X) (1, Z)\n"};
        require_throws<std::runtime_error>([&] { (void) ket::read_pauli_operator(dup, 2); }, "already present");
        auto bad = std::stringstream {"1.0 0.0 : (0, Q)\n"};
        require_throws<std::out_of_range>([&] { (void) ket::read_pauli_operator(bad, 2); });
        require_throws<std::ios::failure>([&] { (void) ket::read_pauli_operator(std::filesystem::path {"/nonexistent/h.txt"}, 2); });
    });
    run("SparsePauliString bookkeeping", [] {
        auto s = ket::SparsePauliString {4};
        s.add(2, T::X);
        s.add(0, T::Z);
        REQUIRE(s.contains_index(2) && !s.contains_index(1));
        require_throws<std::runtime_error>([&] { s.add(2, T::Y); }, "already present");
        require_throws<std::runtime_error>([&] { s.add(7, T::Y); }, "beyond qubit range");
        s.overwrite(2, T::Y);
        REQUIRE(s.at(2) == T::Y);
        s.remove(2);
        REQUIRE(!s.contains_index(2));
        require_throws<std::runtime_error>([&] { (void) s.at(2); }, "no Pauli term");
        const auto a = ket::SparsePauliString {{T::Z, T::I, T::I, T::I}};
        REQUIRE(a == s);
        require_throws<std::runtime_error>([] { (void) ket::SparsePauliString {0}; }, "0 qubits");
    });
    run("PauliOperator construction errors", [] {
        require_throws<std::runtime_error>([] {
            (void) ket::PauliOperator {std::vector<ket::WeightedPauliString> {}};
        }, "non-empty");
        require_throws<std::runtime_error>([] {
            (void) ket::PauliOperator {{{1.0, 0.0}, ket::SparsePauliString {T::X}},
                                       {{1.0, 0.0}, ket::SparsePauliString {T::X, T::Y}}};
        }, "same number of qubits");
    });
    run("n_local assigns parameters like a manual circuit", [] {
        auto [c, ids] = ket::n_local(2, {ket::Gate::RX}, {ket::Gate::CX}, ket::NLocalEntangelement::LINEAR, 2,
                                     ket::SkipLastRotationLayerFlag::FALSE);
        const std::vector<double> values {1.1, 2.2, 3.3, 4.4, 5.5, 6.6};
        REQUIRE(ids.size() == values.size());
        for (std::size_t i = 0; i < ids.size(); ++i) {
            c.set_parameter_value(ids[i], values[i]);
        }
        auto manual = ket::QuantumCircuit {2};
        manual.add_rx_gate(0, values[0]);
        manual.add_rx_gate(1, values[1]);
        manual.add_cx_gate(0, 1);
        manual.add_rx_gate(0, values[2]);
        manual.add_rx_gate(1, values[3]);
        manual.add_cx_gate(0, 1);
        manual.add_rx_gate(0, values[4]);
        manual.add_rx_gate(1, values[5]);
        const auto a = c.flatten();
        const auto b = manual.flatten();
        REQUIRE(a.ops.size() == b.ops.size());
        for (std::size_t i = 0; i < a.ops.size(); ++i) {
            REQUIRE(a.ops[i].gate == b.ops[i].gate);
            REQUIRE(a.ops[i].q0 == b.ops[i].q0);
            REQUIRE(a.ops[i].q1 == b.ops[i].q1);
            REQUIRE(a.ops[i].angle == b.ops[i].angle);
        }
    });
    run("n_local skip-last and entanglement checks", [] {
        auto [c, ids] = ket::n_local(4, {ket::Gate::RY, ket::Gate::RZ}, {ket::Gate::CX}, ket::NLocalEntangelement::LINEAR, 2,
                                     ket::SkipLastRotationLayerFlag::TRUE);
        REQUIRE(ids.size() == 16);
        REQUIRE(c.n_circuit_elements() == 16 + 6);
        REQUIRE(c[8].get_gate().gate == ket::Gate::CX);
        require_throws<std::runtime_error>([] {
            (void) ket::n_local(3, {ket::Gate::RX}, {ket::Gate::H}, ket::NLocalEntangelement::LINEAR, 1,
                                ket::SkipLastRotationLayerFlag::FALSE);
        }, "must be controlled");
    });
    run("circuit builders validate qubit indices", [] {
        auto c = ket::QuantumCircuit {3};
        c.add_x_gate(std::vector<std::size_t> {0, 1, 2});
        REQUIRE(c.n_circuit_elements() == 3);
        require_throws<std::runtime_error>([&] { c.add_h_gate(3); }, "is out of bounds");
        require_throws<std::runtime_error>([&] { c.add_cx_gate(0, 5); }, "target qubit");
        c.pop_back();
        REQUIRE(c.n_circuit_elements() == 2);
        auto sub = ket::QuantumCircuit {3};
        sub.add_x_gate(1);
        c.add_m_gate(0);
        c.add_if_statement(0, sub);
        const auto flat = c.flatten();
        REQUIRE(flat.ops.size() == 5);
        REQUIRE(flat.ops[3].gate == KQ_OP_IF);
        REQUIRE(flat.ops[3].q1 == 1);
    });

    // read_tangelo_file_test.cpp:28-213
    run("read_tangelo_circuit gate lines", [] {
        auto one = std::stringstream {"H         target : [4]   \n"};
        auto c = ket::read_tangelo_circuit(9, one, 0);
        REQUIRE(c.n_circuit_elements() == 1);
        REQUIRE(c[0].get_gate().gate == ket::Gate::H && c[0].get_gate().arg0 == 4);

        auto many = std::stringstream {
            "Circuit object. Size 339200                            \n"
            "                                                       \n"
            "H         target : [4]                                 \n"
            "RX        target : [5]   parameter : 1.5707963267948966\n"
            "CNOT      target : [4]   control : [2]                 \n"
            "RZ        target : [5]   parameter : 12.533816585267923\n"};
        c = ket::read_tangelo_circuit(9, many, 2);
        REQUIRE(c.n_circuit_elements() == 4);
        REQUIRE(c[1].get_gate().gate == ket::Gate::RX && c[1].get_gate().arg0 == 5);
        REQUIRE(c[1].get_gate().arg2 == 1.5707963267948966);
        REQUIRE(c[2].get_gate().gate == ket::Gate::CX && c[2].get_gate().arg0 == 2 && c[2].get_gate().arg1 == 4);
        REQUIRE(c[3].get_gate().gate == ket::Gate::RZ && c[3].get_gate().arg2 == 12.533816585267923);

        auto swap = std::stringstream {"SWAP      target : [12, 9]\n"};
        c = ket::read_tangelo_circuit(13, swap, 0);
        REQUIRE(c.n_circuit_elements() == 3);
        for (std::size_t i = 0; i < 3; ++i) {
            REQUIRE(c[i].get_gate().gate == ket::Gate::CX);
        }
        REQUIRE(c[0].get_gate().arg0 == c[1].get_gate().arg1 && c[1].get_gate().arg1 == c[2].get_gate().arg0);
        REQUIRE(c[0].get_gate().arg1 == c[1].get_gate().arg0 && c[1].get_gate().arg0 == c[2].get_gate().arg1);

        auto phase = std::stringstream {"PHASE     target : [11]   parameter : -1.3474016644659843\n"};
        c = ket::read_tangelo_circuit(13, phase, 0);
        REQUIRE(c[0].get_gate().gate == ket::Gate::P && c[0].get_gate().arg0 == 11);
        REQUIRE(c[0].get_gate().arg2 == -1.3474016644659843);

        const std::pair<const char*, ket::Gate> controlled[] = {
            {"CPHASE    target : [9]   control : [12]   parameter : -0.39269908169872414\n", ket::Gate::CP},
            {"CRX    target : [9]   control : [12]   parameter : -0.39269908169872414\n", ket::Gate::CRX},
            {"CRY    target : [9]   control : [12]   parameter : -0.39269908169872414\n", ket::Gate::CRY},
            {"CRZ    target : [9]   control : [12]   parameter : -0.39269908169872414\n", ket::Gate::CRZ}};
        for (const auto& [text, gate] : controlled) {
            auto in = std::stringstream {text};
            c = ket::read_tangelo_circuit(13, in, 0);
            REQUIRE(c.n_circuit_elements() == 1);
            const auto& g = c[0].get_gate();
            REQUIRE(g.gate == gate && g.arg0 == 12 && g.arg1 == 9 && g.arg2 == -0.39269908169872414);
        }

        auto m = std::stringstream {"M         target : [1]   bit : [4]\n"};
        c = ket::read_tangelo_circuit(13, m, 0);
        REQUIRE(c[0].get_gate().gate == ket::Gate::M && c[0].get_gate().arg0 == 1 && c[0].get_gate().arg1 == 4);

        const auto expected = ket::Matrix2X2 {{1.234, -4.321}, {2.345, -5.432}, {3.456, -6.543}, {4.567, -7.654}};
        auto u = std::stringstream {
            "U         target : [1]\n"
            "    [1.234, -4.321]   [2.345, -5.432]\n"
            "    [3.456, -6.543]   [4.567, -7.654]\n"};
        c = ket::read_tangelo_circuit(10, u, 0);
        REQUIRE(c.n_circuit_elements() == 1 && c[0].get_gate().gate == ket::Gate::U && c[0].get_gate().arg0 == 1);
        REQUIRE(same_matrix(*c[0].get_gate().unitary, expected, 1.0e-12));
        auto cu = std::stringstream {
            "CU        target : [1]   control : [2]\n"
            "    [1.234, -4.321]   [2.345, -5.432]\n"
            "    [3.456, -6.543]   [4.567, -7.654]\n"};
        c = ket::read_tangelo_circuit(10, cu, 0);
        const auto& g = c[0].get_gate();
        REQUIRE(g.gate == ket::Gate::CU && g.arg0 == 2 && g.arg1 == 1);
        REQUIRE(same_matrix(*g.unitary, expected, 1.0e-12));

        auto bad = std::stringstream {"FOO       target : [1]\n"};
        require_throws<std::runtime_error>([&] { (void) ket::read_tangelo_circuit(3, bad, 0); }, "Unknown gate");
        auto orphan = std::stringstream {"ELSE\n    X         target : [0]\n"};
        require_throws<std::runtime_error>([&] { (void) ket::read_tangelo_circuit(3, orphan, 0); }, "'ELSE'");
    });

    // write_tangelo_file_test.cpp:14-179
    run("write_tangelo_circuit line formats", [] {
        auto c = ket::QuantumCircuit {6};
        c.add_h_gate(5);
        c.add_sx_gate(0);
        c.add_csx_gate(3, 0);
        c.add_ch_gate(3, 5);
        const double value = 1.1234567812345677;
        c.add_rx_gate(1, value);
        c.add_p_gate(0, -value);
        c.add_crx_gate(3, 1, value);
        c.add_cp_gate(1, 0, -value);
        c.add_m_gate(1, 3);
        REQUIRE(written(c)
                == "H         target : [5]\n"
                   "SX        target : [0]\n"
                   "CSX       target : [0]   control : [3]\n"
                   "CH        target : [5]   control : [3]\n"
                   "RX        target : [1]   parameter : 1.1234567812345677\n"
                   "P         target : [0]   parameter : -1.1234567812345677\n"
                   "CRX       target : [1]   control : [3]   parameter : 1.1234567812345677\n"
                   "CP        target : [0]   control : [1]   parameter : -1.1234567812345677\n"
                   "M         target : [1]   bit : [3]\n");

        auto u = ket::QuantumCircuit {2};
        u.add_u_gate(ket::Matrix2X2 {{1.234, -4.321}, {2.345, -5.432}, {3.456, -6.543}, {4.567, -7.654}}, 1);
        REQUIRE(written(u)
                == "U         target : [1]\n"
                   "    [ 1.2339999999999, -4.3209999999999]   [ 2.3450000000000, -5.4320000000000]\n"
                   "    [ 3.4559999999999, -6.5430000000000]   [ 4.5670000000000, -7.6539999999999]\n");

        auto without_logger = ket::QuantumCircuit {2};
        without_logger.add_x_gate(0);
        without_logger.add_h_gate(1);
        auto with_logger = ket::QuantumCircuit {2};
        with_logger.add_x_gate(0);
        with_logger.add_classical_register_circuit_logger();
        with_logger.add_h_gate(1);
        REQUIRE(written(without_logger) == written(with_logger));
    });

    // read_tangelo_file_test.cpp:215-267: write then read back, with control flow
    run("tangelo round trip with control flow", [] {
        auto x_and_x = ket::QuantumCircuit {3};
        x_and_x.add_x_gate({0, 2});
        auto x_and_h = ket::QuantumCircuit {3};
        x_and_h.add_x_gate(1);
        x_and_h.add_h_gate(2);
        auto cx_and_h = ket::QuantumCircuit {3};
        cx_and_h.add_cx_gate(1, 2);
        cx_and_h.add_h_gate(2);
        auto original = ket::QuantumCircuit {3};
        original.add_x_gate({0, 1});
        original.add_h_gate({0, 1, 2});
        original.add_m_gate({0, 1});
        original.add_if_statement(0, x_and_x);
        original.add_y_gate(0);
        original.add_z_gate(1);
        original.add_if_else_statement(1, x_and_h, cx_and_h);
        original.add_y_gate(0);
        original.add_z_gate(1);
        original.add_u_gate(ket::Matrix2X2 {{0.6, 0.0}, {0.0, 0.8}, {0.0, 0.8}, {0.6, 0.0}}, 2);
        original.add_crz_gate(2, 0, 0.375);
        auto text = std::stringstream {written(original)};
        const auto back = ket::read_tangelo_circuit(3, text, 0);
        // matrix elements are cut (not rounded) to 13 decimals, so the read-back matrix is within 1e-12
        REQUIRE(same_circuit(original, back, 1.0e-12));
    });

    // numpy_statevector_test.cpp:12-49
    run("read_numpy_statevector known answer", [] {
        auto in = std::stringstream {
            "3                                                    \n"
            " (4.950737714883371443e-02+4.950737714883371443e-02j)\n"
            " (9.901475429766742886e-02+9.901475429766742886e-02j)\n"
            " (1.485221314465011433e-01+1.485221314465011433e-01j)\n"
            " (1.980295085953348577e-01+1.980295085953348577e-01j)\n"
            " (2.475368857441685444e-01+2.475368857441685444e-01j)\n"
            " (2.970442628930022866e-01+2.970442628930022866e-01j)\n"
            " (3.465516400418360288e-01+3.465516400418360288e-01j)\n"
            " (3.960590171906697154e-01+3.960590171906697154e-01j)\n"};
        const auto actual = ket::read_numpy_statevector(in);
        std::vector<std::complex<double>> want;
        double norm_sq = 0.0;
        for (int i = 1; i < 9; ++i) {
            want.emplace_back(1.1 * i, 1.1 * i);
            norm_sq += std::norm(want.back());
        }
        for (auto& z : want) {
            z /= std::sqrt(norm_sq);
        }
        REQUIRE(actual.n_qubits() == 3);
        for (std::size_t i = 0; i < 8; ++i) {
            REQUIRE(std::abs(actual[i] - want[i]) < 1.0e-15);
        }
        auto big = std::stringstream {"1\n (6.0e-01+0.0e+00j)\n (0.0e+00-8.0e-01j)\n"};
        const auto b = ket::read_numpy_statevector(big, ket::Endian::BIG);
        REQUIRE(std::abs(b[1] - std::complex<double> {0.0, -0.8}) < 1.0e-15);
    });

    // statevector_test.cpp:11-33 (the state comes from amplitudes here; the device suite saves a simulated one)
    run("save_statevector / load_statevector", [] {
        std::vector<std::complex<double>> amps;
        double norm_sq = 0.0;
        for (int i = 0; i < 8; ++i) {
            amps.emplace_back(0.1 * (i + 1), i % 3 == 0 ? -0.05 * i : 0.07 * i);
            norm_sq += std::norm(amps.back());
        }
        for (auto& z : amps) {
            z /= std::sqrt(norm_sq);
        }
        const auto state = ket::Statevector {amps};
        auto out = std::stringstream {};
        ket::save_statevector(out, state);
        const std::string text = out.str();
        REQUIRE(text.rfind("ENDIANNESS: LITTLE\nNUMBER OF STATES: 8\n", 0) == 0);
        char first[128];
        std::snprintf(first, sizeof first, " %.14f    %.14f\n", amps[0].real(), amps[0].imag());
        REQUIRE(text.find(first) != std::string::npos);
        const auto back = ket::load_statevector(out);
        REQUIRE(ket::almost_eq(state, back));

        auto big = std::stringstream {};
        ket::save_statevector(big, state, ket::Endian::BIG);
        REQUIRE(big.str().rfind("ENDIANNESS: BIG\n", 0) == 0);
        const auto back_big = ket::load_statevector(big);
        for (std::size_t i = 0; i < 8; ++i) {
            REQUIRE(std::abs(back_big[i] - state[i]) < 1.0e-13);
        }
        // line 1 of a BIG file is amplitude 0b100 (index 1 with its 3 bits reversed)
        auto lines = std::stringstream {big.str()};
        std::string line;
        std::getline(lines, line);
        std::getline(lines, line);
        std::getline(lines, line);
        std::getline(lines, line);
        std::snprintf(first, sizeof first, " %.14f    %.14f", amps[4].real(), amps[4].imag());
        REQUIRE(line == first);
    });

    std::printf("%d passed, %d failed\n", g_passed, g_failed);
    return g_failed == 0 ? 0 : 1;
}
