// Host-only tests of the kettle-compatible C++ API (no device needed): the Hamiltonian reader,
// Pauli-string bookkeeping, circuit builders, n_local and flattening.  They mirror the reference's
// test/source/io/read_pauli_operator_test.cpp, operator/pauli/sparse_pauli_string_test.cpp,
// circuit/circuit_test.cpp and optimize/n_local_test.cpp.
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

    std::printf("%d passed, %d failed\n", g_passed, g_failed);
    return g_failed == 0 ? 0 : 1;
}
