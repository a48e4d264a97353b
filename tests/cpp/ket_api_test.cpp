// Tests of the kettle-compatible C++ API (include/kettle/*) running on the GPU through libkq.
// They mirror the reference's own unit tests (test/source/simulation/simulate_test.cpp,
// control_flow_test.cpp, measure_test.cpp, operator/pauli/pauli_operator_test.cpp,
// state/density_matrix_test.cpp, optimize/n_local_test.cpp) written against the same API.
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include <kettle/kettle.hpp>
#include <kq.h>

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

using cd = std::complex<double>;
const double R2 = M_SQRT1_2;

auto sv(std::vector<cd> c) -> ket::Statevector { return ket::Statevector {std::move(c)}; }

auto max_abs_diff(const ket::Statevector& a, const ket::Statevector& b) -> double
{
    double m = 0.0;
    for (std::size_t i = 0; i < a.n_states(); ++i) {
        m = std::max(m, std::abs(a[i] - b[i]));
    }
    return m;
}

}  // namespace

int main()
{
    // ---- simulate_test.cpp known answers ---------------------------------------------------------
    run("X gate", [] {
        auto c = ket::QuantumCircuit {2};
        c.add_x_gate(1);
        auto s = ket::Statevector {2};
        ket::simulate(c, s);
        REQUIRE(ket::almost_eq(s, ket::Statevector {"01"}));
    });
    run("H gate two qubits", [] {
        auto c = ket::QuantumCircuit {2};
        c.add_h_gate({0, 1});
        auto s = ket::Statevector {2};
        ket::simulate(c, s);
        REQUIRE(ket::almost_eq(s, sv({0.5, 0.5, 0.5, 0.5})));
    });
    run("RX(pi/4) RX(pi/8)", [] {
        const auto c8 = std::cos(M_PI / 8), s8 = std::sin(M_PI / 8), c16 = std::cos(M_PI / 16), s16 = std::sin(M_PI / 16);
        auto c = ket::QuantumCircuit {2};
        c.add_rx_gate(0, M_PI / 4.0);
        c.add_rx_gate(1, M_PI / 8.0);
        auto s = ket::Statevector {2};
        ket::simulate(c, s);
        const auto expected = sv({{c8 * c16, 0.0}, {0.0, -s8 * c16}, {0.0, -c8 * s16}, {-s8 * s16, 0.0}});
        REQUIRE(max_abs_diff(s, expected) < 1e-14);
    });
    run("CX truth table", [] {
        const std::map<std::string, std::string> table = {{"00", "00"}, {"10", "11"}, {"01", "01"}, {"11", "10"}};
        for (const auto& [in, out] : table) {
            auto c = ket::QuantumCircuit {2};
            c.add_cx_gate(0, 1);
            auto s = ket::Statevector {in};
            ket::simulate(c, s);
            REQUIRE(ket::almost_eq(s, ket::Statevector {out}));
        }
    });
    run("CRX and CP on basis states", [] {
        for (double t : {0.0, M_PI / 6.0, M_PI / 3.0, 0.99 * M_PI}) {
            auto c = ket::QuantumCircuit {2};
            c.add_crx_gate(0, 1, t);
            auto s = ket::Statevector {"10"};
            ket::simulate(c, s);
            REQUIRE(max_abs_diff(s, sv({0.0, std::cos(t / 2), 0.0, {0.0, -std::sin(t / 2)}})) < 1e-14);
            auto c2 = ket::QuantumCircuit {2};
            c2.add_cp_gate(0, 1, t);
            auto s2 = ket::Statevector {"11"};
            ket::simulate(c2, s2);
            REQUIRE(max_abs_diff(s2, sv({0.0, 0.0, 0.0, {std::cos(t), std::sin(t)}})) < 1e-14);
        }
    });
    run("U gate mimics named gates", [] {
        const std::vector<std::pair<ket::Gate, ket::Matrix2X2>> gates = {
            {ket::Gate::H, ket::h_gate()}, {ket::Gate::Y, ket::y_gate()}, {ket::Gate::SX, ket::sx_gate()},
            {ket::Gate::T, ket::t_gate()}};
        for (const auto& [g, m] : gates) {
            auto named = ket::QuantumCircuit {3};
            auto mimic = ket::QuantumCircuit {3};
            named.add_h_gate(0);
            mimic.add_h_gate(0);
            named.add_rx_gate(1, 0.7);
            mimic.add_rx_gate(1, 0.7);
            if (g == ket::Gate::H) { named.add_h_gate(1); }
            if (g == ket::Gate::Y) { named.add_y_gate(1); }
            if (g == ket::Gate::SX) { named.add_sx_gate(1); }
            if (g == ket::Gate::T) { named.add_t_gate(1); }
            mimic.add_u_gate(m, 1);
            auto a = ket::Statevector {3};
            auto b = ket::Statevector {3};
            ket::simulate(named, a);
            ket::simulate(mimic, b);
            REQUIRE(max_abs_diff(a, b) < 1e-14);
        }
    });
    run("QFT of |1> basis state is a phase ramp", [] {
        const std::size_t n = 5;
        auto c = ket::QuantumCircuit {n};
        c.add_qft_gate({0, 1, 2, 3, 4});
        auto s = ket::Statevector {"10000"};
        ket::simulate(c, s);
        const double norm = 1.0 / std::sqrt(32.0);
        for (std::size_t k = 0; k < 32; ++k) {
            REQUIRE(std::abs(std::abs(s[k]) - norm) < 1e-12);
        }
    });

    // ---- parameters (simulate_with_parameter_test.cpp) -------------------------------------------
    run("parameterised gate follows set_parameter_value", [] {
        auto c = ket::QuantumCircuit {1};
        const auto id = c.add_ry_gate(0, 0.0, ket::param::parameterized {});
        c.set_parameter_value(id, M_PI);
        auto s = ket::Statevector {1};
        ket::simulate(c, s);
        REQUIRE(max_abs_diff(s, sv({0.0, 1.0})) < 1e-15);
        require_throws<std::out_of_range>([&] { c.set_parameter_value(ket::param::ParameterID {}, 1.0); });
        auto c2 = ket::QuantumCircuit {1};
        c2.add_rx_gate(0, ket::param::ParameterID {});
        auto s2 = ket::Statevector {1};
        require_throws<std::runtime_error>([&] { ket::simulate(c2, s2); }, "uninitialized parameter");
    });

    // ---- control flow (control_flow_test.cpp) -----------------------------------------------------
    run("if statement", [] {
        for (bool flip : {true, false}) {
            auto c = ket::QuantumCircuit {2};
            if (flip) {
                c.add_x_gate(0);
            }
            c.add_m_gate(0);
            auto sub = ket::QuantumCircuit {2};
            sub.add_x_gate(1);
            c.add_if_statement(0, sub);
            auto s = ket::Statevector {"00"};
            ket::simulate(c, s);
            REQUIRE(ket::almost_eq(s, ket::Statevector {flip ? "11" : "00"}));
        }
    });
    run("if-not and if-else statements", [] {
        auto c = ket::QuantumCircuit {2};
        c.add_m_gate(0);
        auto sub = ket::QuantumCircuit {2};
        sub.add_x_gate(0);
        c.add_if_not_statement(0, sub);
        auto a = ket::QuantumCircuit {2};
        a.add_x_gate(1);
        auto b = ket::QuantumCircuit {2};
        b.add_h_gate(1);
        c.add_m_gate(0, 1);
        c.add_if_else_statement(1, a, b);
        auto s = ket::Statevector {"00"};
        auto sim = ket::StatevectorSimulator {};
        sim.run(c, s, 3);
        REQUIRE(ket::almost_eq(s, ket::Statevector {"11"}));
        REQUIRE(sim.classical_register().get(0) == 0);
        REQUIRE(sim.classical_register().get(1) == 1);
    });
    run("predicate on an unmeasured bit", [] {
        auto c = ket::QuantumCircuit {2};
        auto sub = ket::QuantumCircuit {2};
        sub.add_x_gate(1);
        c.add_if_statement(1, sub);
        auto s = ket::Statevector {2};
        require_throws<std::runtime_error>([&] { ket::simulate(c, s); }, "There is no measured bit at classical register 1");
    });
    run("loggers", [] {
        auto c = ket::QuantumCircuit {2};
        c.add_h_gate(0);
        c.add_statevector_circuit_logger();
        c.add_m_gate(0);
        c.add_classical_register_circuit_logger();
        auto s = ket::Statevector {2};
        auto sim = ket::StatevectorSimulator {};
        sim.run(c, s, 7);
        const auto& logs = sim.circuit_loggers();
        REQUIRE(logs.size() == 2);
        REQUIRE(logs[0].is_statevector_circuit_logger());
        REQUIRE(ket::almost_eq(logs[0].get_statevector_circuit_logger().statevector(), sv({R2, R2, 0.0, 0.0})));
        REQUIRE(logs[1].get_classical_register_circuit_logger().classical_register().get(0)
                == sim.classical_register().get(0));
    });
    run("qubit count mismatch", [] {
        auto c = ket::QuantumCircuit {3};
        auto s = ket::Statevector {2};
        require_throws<std::runtime_error>([&] { ket::simulate(c, s); }, "different number of qubits");
    });

    // ---- measurements (measurements / probabilities tests) ---------------------------------------
    run("Bell state counts", [] {
        auto c = ket::QuantumCircuit {2};
        c.add_h_gate(0);
        c.add_cx_gate(0, 1);
        auto s = ket::Statevector {2};
        ket::simulate(c, s);
        const auto counts = ket::perform_measurements_as_counts(s, 4000, nullptr, 11);
        REQUIRE(counts.size() == 2);
        REQUIRE(counts.at("00") + counts.at("11") == 4000);
        REQUIRE(counts.at("00") > 1800 && counts.at("11") > 1800);
        const auto again = ket::perform_measurements_as_counts(s, 4000, nullptr, 11);
        REQUIRE(again == counts);
        const auto probs = ket::calculate_probabilities_raw(s);
        REQUIRE(std::abs(probs[0] - 0.5) < 1e-15 && std::abs(probs[3] - 0.5) < 1e-15);
        auto noise = ket::QuantumNoise {2};
        noise.set(0, 0.25);
        const auto pn = ket::calculate_probabilities_raw(s, &noise);
        REQUIRE(std::abs(pn[1] - 0.125) < 1e-15);
        require_throws<std::runtime_error>([&] { noise.set(1, 1.5); });
        const auto marg = ket::perform_measurements_as_counts_marginal(s, 100, {1}, nullptr, 3);
        for (const auto& [k, v] : marg) {
            REQUIRE(k[1] == 'x');
        }
    });

    // ---- Pauli operators (pauli_operator_test.cpp) ---------------------------------------------------
    run("Pauli expectation tables", [] {
        using T = ket::PauliTerm;
        auto s = sv({0.5, 0.5, 0.5, 0.5});
        const auto op = ket::PauliOperator {{{1.0, 0.0}, ket::SparsePauliString {T::X, T::I}},
                                            {{0.0, -1.0}, ket::SparsePauliString {T::I, T::X}}};
        REQUIRE(std::abs(ket::expectation_value(op, s) - cd {1.0, -1.0}) < 1e-14);
        auto bell = sv({R2, 0.0, 0.0, R2});
        const auto zz = ket::SparsePauliString {T::Z, T::Z};
        const auto yy = ket::SparsePauliString {T::Y, T::Y};
        REQUIRE(std::abs(ket::expectation_value(zz, bell) - 1.0) < 1e-14);
        REQUIRE(std::abs(ket::expectation_value(yy, bell) + 1.0) < 1e-14);
        auto str = ket::SparsePauliString {3};
        str.add(0, T::X);
        require_throws<std::runtime_error>([&] { str.add(0, T::Z); }, "already present");
        str.overwrite(0, T::Z);
        REQUIRE(str.at(0) == T::Z);
        auto ket_state = ket::Statevector {"110"};
        ket::simulate(ket::SparsePauliString {T::X, T::I, T::X}, ket_state);
        REQUIRE(ket::almost_eq(ket_state, ket::Statevector {"011"}));
    });

    // ---- density matrix (simulate_density_matrix_test.cpp, density_matrix_test.cpp) --------------------
    run("density matrix follows the statevector", [] {
        auto c = ket::QuantumCircuit {3};
        c.add_h_gate(0);
        c.add_cx_gate(0, 1);
        c.add_ry_gate(2, 0.4);
        c.add_crz_gate(2, 0, 1.1);
        auto s = ket::Statevector {3};
        ket::simulate(c, s);
        auto dm = ket::DensityMatrix {3};
        ket::simulate(c, dm);
        const auto& m = dm.matrix();
        for (std::size_t r = 0; r < 8; ++r) {
            for (std::size_t k = 0; k < 8; ++k) {
                REQUIRE(std::abs(m(r, k) - s[r] * std::conj(s[k])) < 1e-14);
            }
        }
        REQUIRE(std::abs(dm.trace() - 1.0) < 1e-14);
        const auto reduced = ket::partial_trace(dm, {2});
        REQUIRE(reduced.n_qubits() == 2);
        // tracing out qubit 2 (bit 2 of the index): reduced(r, k) = sum_b psi[r + 4b] conj(psi[k + 4b])
        for (std::size_t r = 0; r < 4; ++r) {
            for (std::size_t k = 0; k < 4; ++k) {
                const auto expected = s[r] * std::conj(s[k]) + s[r + 4] * std::conj(s[k + 4]);
                REQUIRE(std::abs(reduced.matrix()(r, k) - expected) < 1e-14);
            }
        }
        auto from_sv = ket::statevector_to_density_matrix(s);
        REQUIRE(ket::almost_eq(from_sv, dm));
        REQUIRE(std::abs(dm.trace_of_square() - 1.0) < 1e-13);
        REQUIRE(dm.is_pure());
        // the 2-qubit reduced state of an entangled 3-qubit state is mixed: tr(rho^2) = sum |rho(r,k)|^2 < 1
        auto purity = std::complex<double> {};
        for (std::size_t r = 0; r < 4; ++r) {
            for (std::size_t k = 0; k < 4; ++k) {
                purity += reduced.matrix()(r, k) * reduced.matrix()(k, r);
            }
        }
        REQUIRE(std::abs(reduced.trace_of_square() - purity) < 1e-13);
        REQUIRE(reduced.is_pure() == (std::norm(purity - 1.0) <= ket::DENSITY_MATRIX_TRACE_TOLERANCE));
    });
    run("density matrix measurement and loggers", [] {
        auto c = ket::QuantumCircuit {2};
        c.add_x_gate(0);
        c.add_m_gate(0);
        auto dm = ket::DensityMatrix {"00"};
        auto sim = ket::DensityMatrixSimulator {2};
        sim.run(c, dm, 1);
        REQUIRE(sim.classical_register().get(0) == 1);
        REQUIRE(std::abs(dm.matrix()(1, 1) - 1.0) < 1e-15);
        auto c2 = ket::QuantumCircuit {2};
        c2.add_statevector_circuit_logger();
        require_throws<std::runtime_error>([&] { ket::simulate(c2, dm); }, "loggers");
    });

    // ---- n_local and the adjoint gradient ------------------------------------------------------------
    run("n_local structure", [] {
        auto [c, ids] = ket::n_local(4, {ket::Gate::RY, ket::Gate::RZ}, {ket::Gate::CX}, ket::NLocalEntangelement::LINEAR, 2,
                                     ket::SkipLastRotationLayerFlag::TRUE);
        REQUIRE(ids.size() == 16);
        REQUIRE(c.n_circuit_elements() == 16 + 6);
        REQUIRE(c[8].get_gate().gate == ket::Gate::CX);
    });
    run("adjoint gradient matches central differences", [] {
        const std::size_t n = 6;
        auto [c, ids] = ket::n_local(n, {ket::Gate::RY, ket::Gate::RZ}, {ket::Gate::CX}, ket::NLocalEntangelement::LINEAR, 3,
                                     ket::SkipLastRotationLayerFlag::FALSE);
        auto prng = std::mt19937 {5};
        auto uni = std::uniform_real_distribution<double> {-M_PI, M_PI};
        std::vector<double> theta(ids.size());
        for (std::size_t i = 0; i < ids.size(); ++i) {
            theta[i] = uni(prng);
            c.set_parameter_value(ids[i], theta[i]);
        }
        using T = ket::PauliTerm;
        auto h = ket::PauliOperator {n};
        for (std::size_t i = 0; i < n; ++i) {
            h.add(-1.0, ket::SparsePauliString {{{i, T::Z}, {(i + 1) % n, T::Z}}, n});
            h.add(-0.7, ket::SparsePauliString {{{i, T::X}}, n});
        }
        h.add(0.3, ket::SparsePauliString {{{0, T::Y}, {3, T::Y}}, n});
        const auto energy = [&](const std::vector<double>& th) {
            for (std::size_t i = 0; i < ids.size(); ++i) {
                c.set_parameter_value(ids[i], th[i]);
            }
            auto s = ket::Statevector {n};
            ket::simulate(c, s);
            return ket::expectation_value(h, s).real();
        };
        const auto e0 = energy(theta);
        auto opts = ket::AdjointOptions {};
        opts.precision = ket::Precision::C128;
        const auto g128 = ket::energy_and_gradient(c, ids, h, opts);
        REQUIRE(std::abs(g128.energy - e0) < 1e-12);
        const double eps = 1e-5;
        for (std::size_t i = 0; i < ids.size(); ++i) {
            auto tp = theta, tm = theta;
            tp[i] += eps;
            tm[i] -= eps;
            const auto fd = (energy(tp) - energy(tm)) / (2 * eps);
            REQUIRE(std::abs(fd - g128.gradient[i]) < 1e-8);
        }
        energy(theta);
        opts.precision = ket::Precision::C64;
        const auto g64 = ket::energy_and_gradient(c, ids, h, opts);
        REQUIRE(std::abs(g64.energy - e0) < 1e-5);
        for (std::size_t i = 0; i < ids.size(); ++i) {
            REQUIRE(std::abs(g64.gradient[i] - g128.gradient[i]) < 1e-4);
        }
    });

    // measurements.cpp:178-208: per-shot re-simulation of a circuit with a mid-circuit measurement
    run("marginal counts re-simulate the circuit per shot", [] {
        auto c = ket::QuantumCircuit {2};
        c.add_h_gate(0);
        c.add_m_gate(0);
        auto flip = ket::QuantumCircuit {2};
        flip.add_x_gate(1);
        c.add_if_statement(0, flip);
        const auto state = ket::Statevector {2};
        const auto counts = ket::perform_measurements_as_counts_marginal(c, state, 4000, {}, nullptr, std::nullopt);
        // outcome 0 leaves |00>, outcome 1 gives |11>: only "00" and "11", about half each
        std::size_t total = 0;
        for (const auto& [bits, n] : counts) {
            REQUIRE(bits == "00" || bits == "11");
            total += n;
        }
        REQUIRE(total == 4000);
        REQUIRE(counts.at("00") > 1700 && counts.at("11") > 1700);
        const auto marginal = ket::perform_measurements_as_counts_marginal(c, state, 100, {1}, nullptr, 3);
        for (const auto& [bits, n] : marginal) {
            REQUIRE(bits == "0x" || bits == "1x");
        }
        REQUIRE(std::abs(state[0] - std::complex<double> {1.0, 0.0}) < 1e-15);
    });

    // ---- failures crossing the C-ABI (kq_debug_set_faults) ---------------------------------------------
    run("injected faults reach kettle's exception types", [] {
        // a cap on single device allocations below one 20-qubit state: simulate fails with std::runtime_error
        REQUIRE(kq_debug_set_faults(size_t {1} << 20, 0, 0) == KQ_OK);
        auto big = ket::QuantumCircuit {20};
        big.add_h_gate(0);
        auto s = ket::Statevector {20};
        require_throws<std::runtime_error>([&] { ket::simulate(big, s); }, "exceeds the cap");
        REQUIRE(kq_debug_set_faults(0, 0, 0) == KQ_OK);
        // a forced status on the next libkq call: out-of-range stays std::out_of_range, others runtime_error
        auto c = ket::QuantumCircuit {2};
        c.add_x_gate(0);
        auto s2 = ket::Statevector {2};
        REQUIRE(kq_debug_set_faults(0, KQ_ERR_OUT_OF_RANGE, 0) == KQ_OK);
        require_throws<std::out_of_range>([&] { ket::simulate(c, s2); }, "injected fault");
        REQUIRE(kq_debug_set_faults(0, KQ_ERR_HIP, 0) == KQ_OK);
        require_throws<std::runtime_error>([&] { ket::simulate(c, s2); }, "injected fault");
        // the fault fires once: the same call then succeeds
        ket::simulate(c, s2);
        REQUIRE(ket::almost_eq(s2, ket::Statevector {"10"}));
    });

    std::printf("%d passed, %d failed\n", g_passed, g_failed);
    return g_failed == 0 ? 0 : 1;
}
