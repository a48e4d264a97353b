/*
 * CPU oracle for kettle's simulate() hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a single-threaded restatement of kettle's own loops, used as the correctness checker
 * for the HIP engine (libkq) and as the timed CPU baseline.  Only the tests, the smoke check
 * and the benchmark's CPU-baseline leg load it; the product never calls it.
 *
 * It follows the reference operation for operation:
 *   pair generators      source/kettle_internal/simulation/gate_pair_generator.hpp:24-160
 *                        source/kettle_internal/common/mathtools.cpp:49-74
 *   per-pair formulas    source/kettle_internal/simulation/operations.cpp:13-185 (out of line,
 *                        cos/sin evaluated per pair, exactly as there)
 *   gate dispatch        source/kettle_internal/simulation/simulate.cpp:39-415
 *   interpreter          source/kettle_internal/simulation/simulate.cpp:417-597
 *   measurement          source/kettle_internal/simulation/measure.{hpp,cpp}, common/prng.cpp
 *   probabilities/noise  source/kettle_internal/calculations/probabilities.cpp:49-124
 *   sampler              source/kettle_internal/calculations/measurements.cpp:235-312
 *   Pauli expectation    source/kettle_internal/operator/pauli/pauli_operator.cpp:76-105,
 *                        source/kettle_internal/simulation/simulate_pauli.cpp:27-128,
 *                        source/kettle_internal/state/state.cpp:159-172
 *   density matrix       source/kettle_internal/simulation/operations_density_matrix.{hpp,cpp},
 *                        simulate_density_matrix.cpp, measure_density_matrix.cpp,
 *                        state/density_matrix.cpp:71-278
 * Build: see oracle/Makefile (g++ -O3 -DNDEBUG, kettle's release flags, no -ffast-math).
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../include/kq.h"

using cd = std::complex<double>;
using State = std::vector<cd>;

namespace
{

thread_local std::string g_err;

/* ---------------- pair generators (gate_pair_generator.hpp:24-160) ---------------- */

struct OnePairGen
{
    std::size_t i0_max, i1_max, i0 = 0, i1 = 0;
    OnePairGen(std::size_t target, std::size_t n)
        : i0_max {std::size_t {1} << target}, i1_max {std::size_t {1} << (n - target - 1)} {}
    void set_state(std::size_t k) { i0 = k / i1_max; i1 = k - i0 * i1_max; }
    std::size_t size() const { return i0_max * i1_max; }
    std::pair<std::size_t, std::size_t> next()
    {
        const auto s0 = i0 + 2 * i1 * i0_max;
        const auto s1 = s0 + i0_max;
        if (++i1 == i1_max) { ++i0; i1 = 0; }
        return {s0, s1};
    }
};

struct TwoPairGen
{
    std::size_t lo, hi, lo_shift, hi_shift, c_shift, t_shift, i0_max, i1_max, i2_max;
    std::size_t i0 = 0, i1 = 0, i2 = 0;
    TwoPairGen(std::size_t c, std::size_t t, std::size_t n)
        : lo {std::min(c, t)}, hi {std::max(c, t)}
        , lo_shift {std::size_t {1} << (lo + 1)}, hi_shift {std::size_t {1} << (hi + 1)}
        , c_shift {std::size_t {1} << c}, t_shift {std::size_t {1} << t}
        , i0_max {std::size_t {1} << lo}, i1_max {std::size_t {1} << (hi - lo - 1)}
        , i2_max {std::size_t {1} << (n - hi - 1)} {}
    void set_state(std::size_t k)
    {
        i0 = k / (i1_max * i2_max);
        k -= i0 * i1_max * i2_max;
        i1 = k / i2_max;
        k -= i1 * i2_max;
        i2 = k;
    }
    std::size_t size() const { return i0_max * i1_max * i2_max; }
    void advance()
    {
        if (++i2 == i2_max) {
            ++i1;
            i2 = 0;
            if (i1 == i1_max) { ++i0; i1 = 0; }
        }
    }
    std::pair<std::size_t, std::size_t> next()
    {
        const auto s0 = i0 + i1 * lo_shift + i2 * hi_shift + c_shift;
        advance();
        return {s0, s0 + t_shift};
    }
    std::tuple<std::size_t, std::size_t, std::size_t, std::size_t> next_unset_and_set()
    {
        const auto c0t0 = i0 + i1 * lo_shift + i2 * hi_shift;
        advance();
        return {c0t0, c0t0 + t_shift, c0t0 + c_shift, c0t0 + t_shift + c_shift};
    }
};

/* number_of_double_qubit_gate_pairs_ (simulate_utils.cpp:26-45): 0 pairs for one qubit */
std::size_t n_double_pairs(std::size_t n) { return n == 1 ? 0 : (std::size_t {1} << (n - 2)); }

/* ---------------- per-pair arithmetic (operations.cpp:13-185) ---------------- */

[[gnu::noinline]] void op_h(State& s, std::size_t a, std::size_t b)
{
    const auto x = s[a], y = s[b];
    s[a] = cd {M_SQRT1_2 * (x.real() + y.real()), M_SQRT1_2 * (x.imag() + y.imag())};
    s[b] = cd {M_SQRT1_2 * (x.real() - y.real()), M_SQRT1_2 * (x.imag() - y.imag())};
}
[[gnu::noinline]] void op_x(State& s, std::size_t a, std::size_t b) { std::swap(s[a], s[b]); }
[[gnu::noinline]] void op_y(State& s, std::size_t a, std::size_t b)
{
    const auto x = s[a], y = s[b];
    s[a] = cd {y.imag(), -y.real()};
    s[b] = cd {-x.imag(), x.real()};
}
[[gnu::noinline]] void op_z(State& s, std::size_t b) { s[b] *= -1.0; }
[[gnu::noinline]] void op_s(State& s, std::size_t b) { const auto y = s[b]; s[b] = {-y.imag(), y.real()}; }
[[gnu::noinline]] void op_sdag(State& s, std::size_t b) { const auto y = s[b]; s[b] = {y.imag(), -y.real()}; }
[[gnu::noinline]] void op_t(State& s, std::size_t b)
{
    const auto y = s[b];
    s[b] = {M_SQRT1_2 * (y.real() - y.imag()), M_SQRT1_2 * (y.real() + y.imag())};
}
[[gnu::noinline]] void op_tdag(State& s, std::size_t b)
{
    const auto y = s[b];
    s[b] = {M_SQRT1_2 * (y.real() + y.imag()), -M_SQRT1_2 * (y.real() - y.imag())};
}
[[gnu::noinline]] void op_sx(State& s, std::size_t a, std::size_t b)
{
    const auto x = s[a], y = s[b];
    s[a] = cd {0.5 * (x.real() - x.imag() + y.real() + y.imag()), 0.5 * (x.real() + x.imag() - y.real() + y.imag())};
    s[b] = cd {0.5 * (x.real() + x.imag() + y.real() - y.imag()), 0.5 * (-x.real() + x.imag() + y.real() + y.imag())};
}
[[gnu::noinline]] void op_sxdag(State& s, std::size_t a, std::size_t b)
{
    const auto x = s[a], y = s[b];
    s[a] = cd {0.5 * (x.real() + x.imag() + y.real() - y.imag()), 0.5 * (-x.real() + x.imag() + y.real() + y.imag())};
    s[b] = cd {0.5 * (x.real() - x.imag() + y.real() + y.imag()), 0.5 * (x.real() + x.imag() - y.real() + y.imag())};
}
[[gnu::noinline]] void op_rx(State& s, std::size_t a, std::size_t b, double th)
{
    const auto x = s[a], y = s[b];
    const auto c = std::cos(th / 2.0), sn = std::sin(th / 2.0);
    s[a] = cd {(x.real() * c) + (y.imag() * sn), (x.imag() * c) - (y.real() * sn)};
    s[b] = cd {(y.real() * c) + (x.imag() * sn), (y.imag() * c) - (x.real() * sn)};
}
[[gnu::noinline]] void op_ry(State& s, std::size_t a, std::size_t b, double th)
{
    const auto x = s[a], y = s[b];
    const auto c = std::cos(th / 2.0), sn = std::sin(th / 2.0);
    s[a] = cd {(x.real() * c) - (y.real() * sn), (x.imag() * c) - (y.imag() * sn)};
    s[b] = cd {(y.real() * c) + (x.real() * sn), (y.imag() * c) + (x.imag() * sn)};
}
[[gnu::noinline]] void op_rz(State& s, std::size_t a, std::size_t b, double th)
{
    const auto x = s[a], y = s[b];
    const auto c = std::cos(th / 2.0), sn = std::sin(th / 2.0);
    s[a] = cd {(x.real() * c) + (x.imag() * sn), (x.imag() * c) - (x.real() * sn)};
    s[b] = cd {(y.real() * c) - (y.imag() * sn), (y.imag() * c) + (y.real() * sn)};
}
[[gnu::noinline]] void op_p(State& s, std::size_t b, double th)
{
    const auto y = s[b];
    const auto c = std::cos(th), sn = std::sin(th);
    s[b] = cd {(y.real() * c) - (y.imag() * sn), (y.imag() * c) + (y.real() * sn)};
}
struct M2 { cd e00, e01, e10, e11; };
[[gnu::noinline]] void op_u(State& s, std::size_t a, std::size_t b, const M2& m)
{
    const auto x = s[a], y = s[b];
    s[a] = x * m.e00 + y * m.e01;
    s[b] = x * m.e10 + y * m.e11;
}

M2 op_matrix(const kq_op& op)
{
    return {cd {op.u[0], op.u[1]}, cd {op.u[2], op.u[3]}, cd {op.u[4], op.u[5]}, cd {op.u[6], op.u[7]}};
}

/* one gate over all pairs: simulate_gate_info_ (simulate.cpp:265-415) */
void sv_gate(State& s, std::size_t n, const kq_op& op)
{
    const int g = op.gate;
    if (g <= KQ_P || g == KQ_U) {
        OnePairGen gen {static_cast<std::size_t>(op.q0), n};
        gen.set_state(0);
        const auto np = std::size_t {1} << (n - 1);
        const auto m = op_matrix(op);
        for (std::size_t i = 0; i < np; ++i) {
            const auto [a, b] = gen.next();
            switch (g) {
                case KQ_H: op_h(s, a, b); break;
                case KQ_X: op_x(s, a, b); break;
                case KQ_Y: op_y(s, a, b); break;
                case KQ_Z: op_z(s, b); break;
                case KQ_S: op_s(s, b); break;
                case KQ_SDAG: op_sdag(s, b); break;
                case KQ_T: op_t(s, b); break;
                case KQ_TDAG: op_tdag(s, b); break;
                case KQ_SX: op_sx(s, a, b); break;
                case KQ_SXDAG: op_sxdag(s, a, b); break;
                case KQ_RX: op_rx(s, a, b, op.angle); break;
                case KQ_RY: op_ry(s, a, b, op.angle); break;
                case KQ_RZ: op_rz(s, a, b, op.angle); break;
                case KQ_P: op_p(s, b, op.angle); break;
                case KQ_U: op_u(s, a, b, m); break;
                default: break;
            }
        }
        return;
    }
    TwoPairGen gen {static_cast<std::size_t>(op.q0), static_cast<std::size_t>(op.q1), n};
    gen.set_state(0);
    const auto np = n_double_pairs(n);
    const auto m = op_matrix(op);
    for (std::size_t i = 0; i < np; ++i) {
        const auto [a, b] = gen.next();
        switch (g) {
            case KQ_CH: op_h(s, a, b); break;
            case KQ_CX: op_x(s, a, b); break;
            case KQ_CY: op_y(s, a, b); break;
            case KQ_CZ: op_z(s, b); break;
            case KQ_CS: op_s(s, b); break;
            case KQ_CSDAG: op_sdag(s, b); break;
            case KQ_CT: op_t(s, b); break;
            case KQ_CTDAG: op_tdag(s, b); break;
            case KQ_CSX: op_sx(s, a, b); break;
            case KQ_CSXDAG: op_sxdag(s, a, b); break;
            case KQ_CRX: op_rx(s, a, b, op.angle); break;
            case KQ_CRY: op_ry(s, a, b, op.angle); break;
            case KQ_CRZ: op_rz(s, a, b, op.angle); break;
            case KQ_CP: op_p(s, b, op.angle); break;
            case KQ_CU: op_u(s, a, b, m); break;
            default: break;
        }
    }
}

/* get_prng_ (common/prng.cpp:9-19) */
std::mt19937 get_prng(bool has_seed, int seed)
{
    if (has_seed) {
        return std::mt19937 {static_cast<std::mt19937::result_type>(seed)};
    }
    auto device = std::random_device {};
    return std::mt19937 {device()};
}

/* measure.hpp:33-57 + measure.cpp:22-72 */
int sv_measure(State& s, std::size_t n, std::size_t q, bool has_seed, int seed)
{
    OnePairGen gen {q, n};
    gen.set_state(0);
    double p0 = 0.0, p1 = 0.0;
    for (std::size_t i = 0; i < gen.size(); ++i) {
        const auto [a, b] = gen.next();
        p0 += std::norm(s[a]);
        p1 += std::norm(s[b]);
    }
    auto prng = get_prng(has_seed, seed);
    auto flip = std::discrete_distribution<int> {{p0, p1}};
    const int outcome = flip(prng);
    const double norm = outcome == 0 ? std::sqrt(1.0 / p0) : std::sqrt(1.0 / p1);
    OnePairGen g2 {q, n};
    g2.set_state(0);
    for (std::size_t i = 0; i < g2.size(); ++i) {
        const auto [a, b] = g2.next();
        if (outcome == 0) {
            s[a] *= norm;
            s[b] = {0.0, 0.0};
        }
        else {
            s[a] = {0.0, 0.0};
            s[b] *= norm;
        }
    }
    return outcome;
}

/* ControlFlowPredicate::operator() (control_flow_predicate.cpp:40-69) */
bool eval_predicate(const int32_t* preds, int32_t off, const int8_t* creg)
{
    const int kind = preds[off];
    const int n = preds[off + 1];
    bool out = true;
    for (int i = 0; i < n; ++i) {
        const int bit = preds[off + 2 + 2 * i];
        const int expected = preds[off + 3 + 2 * i];
        if (creg[bit] < 0) {
            throw std::runtime_error {"There is no measured bit at classical register " + std::to_string(bit) + "\n"};
        }
        if (creg[bit] != expected) {
            out = false;
            break;
        }
    }
    return kind == 0 ? out : !out;
}

/* ---------------- density matrix (operations_density_matrix.{hpp,cpp}) ---------------- */

struct DM
{
    std::size_t n, N;
    std::vector<cd> m;   // column-major: (r, c) at r + c*N
    cd& operator()(std::size_t r, std::size_t c) { return m[r + c * N]; }
};

constexpr cd plus_half {0.5, 0.5};
constexpr cd minus_half {0.5, -0.5};
constexpr cd plus_rt2 {M_SQRT1_2, M_SQRT1_2};
constexpr cd minus_rt2 {M_SQRT1_2, -M_SQRT1_2};

M2 mat_of(int g, double th)
{
    /* common_u_gates.cpp:10-172 */
    switch (g) {
        case KQ_H: case KQ_CH: return {{M_SQRT1_2, 0}, {M_SQRT1_2, 0}, {M_SQRT1_2, 0}, {-M_SQRT1_2, 0}};
        case KQ_X: case KQ_CX: return {{0, 0}, {1, 0}, {1, 0}, {0, 0}};
        case KQ_Y: case KQ_CY: return {{0, 0}, {0, -1}, {0, 1}, {0, 0}};
        case KQ_Z: case KQ_CZ: return {{1, 0}, {0, 0}, {0, 0}, {-1, 0}};
        case KQ_S: case KQ_CS: return {{1, 0}, {0, 0}, {0, 0}, {0, 1}};
        case KQ_SDAG: case KQ_CSDAG: return {{1, 0}, {0, 0}, {0, 0}, {0, -1}};
        case KQ_T: case KQ_CT: return {{1, 0}, {0, 0}, {0, 0}, {M_SQRT1_2, M_SQRT1_2}};
        case KQ_TDAG: case KQ_CTDAG: return {{1, 0}, {0, 0}, {0, 0}, {M_SQRT1_2, -M_SQRT1_2}};
        case KQ_SX: case KQ_CSX: return {{0.5, 0.5}, {0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}};
        case KQ_SXDAG: case KQ_CSXDAG: return {{0.5, -0.5}, {0.5, 0.5}, {0.5, 0.5}, {0.5, -0.5}};
        case KQ_RX: case KQ_CRX: {
            const auto c = std::cos(th / 2.0), s = std::sin(th / 2.0);
            return {{c, 0}, {0, -s}, {0, -s}, {c, 0}};
        }
        case KQ_RY: case KQ_CRY: {
            const auto c = std::cos(th / 2.0), s = std::sin(th / 2.0);
            return {{c, 0}, {-s, 0}, {s, 0}, {c, 0}};
        }
        case KQ_RZ: case KQ_CRZ: {
            const auto c = std::cos(th / 2.0), s = std::sin(th / 2.0);
            return {{c, -s}, {0, 0}, {0, 0}, {c, s}};
        }
        case KQ_P: case KQ_CP: {
            const auto c = std::cos(th), s = std::sin(th);
            return {{1, 0}, {0, 0}, {0, 0}, {c, s}};
        }
        default: break;
    }
    throw std::runtime_error {"oracle: no matrix for gate"};
}

M2 adjoint(const M2& m) { return {std::conj(m.e00), std::conj(m.e10), std::conj(m.e01), std::conj(m.e11)}; }

/* apply_1t_gate_first_/second_ and apply_1t1a_gate_first_/second_ and apply_u_gate_*
 * (operations_density_matrix.hpp:38-370, .cpp:18-77) */
void dm_gate_1t(DM& rho, DM& buf, const kq_op& op)
{
    const auto n = rho.n;
    const auto t = static_cast<std::size_t>(op.q0);
    const int g = op.gate;
    const double th = op.angle;
    const auto np = std::size_t {1} << (n - 1);
    OnePairGen outer {t, n}, inner {t, n};
    const M2 um = g == KQ_U ? op_matrix(op) : M2 {};
    /* first pass: buffer = U rho (rows) */
    outer.set_state(0);
    for (std::size_t io = 0; io < np; ++io) {
        const auto [r0, r1] = outer.next();
        inner.set_state(0);
        for (std::size_t ii = 0; ii < np; ++ii) {
            const auto [c0, c1] = inner.next();
            const auto p00 = rho(r0, c0), p10 = rho(r1, c0), p01 = rho(r0, c1), p11 = rho(r1, c1);
            cd b00, b10, b01, b11;
            switch (g) {
                case KQ_H: b00 = M_SQRT1_2 * (p00 + p10); b10 = M_SQRT1_2 * (p00 - p10); b01 = M_SQRT1_2 * (p01 + p11); b11 = M_SQRT1_2 * (p01 - p11); break;
                case KQ_X: b00 = p10; b10 = p00; b01 = p11; b11 = p01; break;
                case KQ_Y: b00 = {p10.imag(), -p10.real()}; b10 = {-p00.imag(), p00.real()}; b01 = {p11.imag(), -p11.real()}; b11 = {-p01.imag(), p01.real()}; break;
                case KQ_Z: b00 = p00; b10 = -p10; b01 = p01; b11 = -p11; break;
                case KQ_S: b00 = p00; b10 = {-p10.imag(), p10.real()}; b01 = p01; b11 = {-p11.imag(), p11.real()}; break;
                case KQ_SDAG: b00 = p00; b10 = {p10.imag(), -p10.real()}; b01 = p01; b11 = {p11.imag(), -p11.real()}; break;
                case KQ_T: b00 = p00; b10 = p10 * plus_rt2; b01 = p01; b11 = p11 * plus_rt2; break;
                case KQ_TDAG: b00 = p00; b10 = p10 * minus_rt2; b01 = p01; b11 = p11 * minus_rt2; break;
                case KQ_SX: b00 = (p00 * plus_half) + (p10 * minus_half); b10 = (p00 * minus_half) + (p10 * plus_half); b01 = (p01 * plus_half) + (p11 * minus_half); b11 = (p01 * minus_half) + (p11 * plus_half); break;
                case KQ_SXDAG: b00 = (p00 * minus_half) + (p10 * plus_half); b10 = (p00 * plus_half) + (p10 * minus_half); b01 = (p01 * minus_half) + (p11 * plus_half); b11 = (p01 * plus_half) + (p11 * minus_half); break;
                case KQ_RX: {
                    const auto c = std::cos(th / 2.0);
                    const auto nis = cd {0.0, -std::sin(th / 2.0)};
                    b00 = (p00 * c) + (p10 * nis); b10 = (p00 * nis) + (p10 * c); b01 = (p01 * c) + (p11 * nis); b11 = (p01 * nis) + (p11 * c);
                    break;
                }
                case KQ_RY: {
                    const auto c = std::cos(th / 2.0), s = std::sin(th / 2.0);
                    b00 = (p00 * c) - (p10 * s); b10 = (p00 * s) + (p10 * c); b01 = (p01 * c) - (p11 * s); b11 = (p01 * s) + (p11 * c);
                    break;
                }
                case KQ_RZ: {
                    const auto c = std::cos(th / 2.0), s = std::sin(th / 2.0);
                    const auto plus = cd {c, s}, minus = cd {c, -s};
                    b00 = p00 * minus; b10 = p10 * plus; b01 = p01 * minus; b11 = p11 * plus;
                    break;
                }
                case KQ_P: {
                    const auto plus = cd {std::cos(th), std::sin(th)};
                    b00 = p00; b10 = p10 * plus; b01 = p01; b11 = p11 * plus;
                    break;
                }
                case KQ_U:
                    b00 = (p00 * um.e00) + (p10 * um.e01); b10 = (p00 * um.e10) + (p10 * um.e11);
                    b01 = (p01 * um.e00) + (p11 * um.e01); b11 = (p01 * um.e10) + (p11 * um.e11);
                    break;
                default: break;
            }
            buf(r0, c0) = b00; buf(r1, c0) = b10; buf(r0, c1) = b01; buf(r1, c1) = b11;
        }
    }
    /* second pass: rho = buffer U^dagger (cols) */
    const M2 ua = adjoint(um);
    outer.set_state(0);
    for (std::size_t io = 0; io < np; ++io) {
        const auto [c0, c1] = outer.next();
        inner.set_state(0);
        for (std::size_t ii = 0; ii < np; ++ii) {
            const auto [r0, r1] = inner.next();
            const auto q00 = buf(r0, c0), q10 = buf(r1, c0), q01 = buf(r0, c1), q11 = buf(r1, c1);
            cd o00, o10, o01, o11;
            switch (g) {
                case KQ_H: o00 = M_SQRT1_2 * (q00 + q01); o10 = M_SQRT1_2 * (q10 + q11); o01 = M_SQRT1_2 * (q00 - q01); o11 = M_SQRT1_2 * (q10 - q11); break;
                case KQ_X: o00 = q01; o10 = q11; o01 = q00; o11 = q10; break;
                case KQ_Y: o00 = {-q01.imag(), q01.real()}; o10 = {-q11.imag(), q11.real()}; o01 = {q00.imag(), -q00.real()}; o11 = {q10.imag(), -q10.real()}; break;
                case KQ_Z: o00 = q00; o10 = q10; o01 = -q01; o11 = -q11; break;
                case KQ_S: o00 = q00; o10 = q10; o01 = {q01.imag(), -q01.real()}; o11 = {q11.imag(), -q11.real()}; break;
                case KQ_SDAG: o00 = q00; o10 = q10; o01 = {-q01.imag(), q01.real()}; o11 = {-q11.imag(), q11.real()}; break;
                case KQ_T: o00 = q00; o10 = q10; o01 = q01 * minus_rt2; o11 = q11 * minus_rt2; break;
                case KQ_TDAG: o00 = q00; o10 = q10; o01 = q01 * plus_rt2; o11 = q11 * plus_rt2; break;
                case KQ_SX: o00 = (q00 * minus_half) + (q01 * plus_half); o10 = (q10 * minus_half) + (q11 * plus_half); o01 = (q00 * plus_half) + (q01 * minus_half); o11 = (q10 * plus_half) + (q11 * minus_half); break;
                case KQ_SXDAG: o00 = (q00 * plus_half) + (q01 * minus_half); o10 = (q10 * plus_half) + (q11 * minus_half); o01 = (q00 * minus_half) + (q01 * plus_half); o11 = (q10 * minus_half) + (q11 * plus_half); break;
                case KQ_RX: {
                    const auto c = std::cos(th / 2.0);
                    const auto pis = cd {0.0, std::sin(th / 2.0)};
                    o00 = (q00 * c) + (q01 * pis); o10 = (q10 * c) + (q11 * pis); o01 = (q00 * pis) + (q01 * c); o11 = (q10 * pis) + (q11 * c);
                    break;
                }
                case KQ_RY: {
                    const auto c = std::cos(th / 2.0), s = std::sin(th / 2.0);
                    o00 = (q00 * c) - (q01 * s); o10 = (q10 * c) - (q11 * s); o01 = (q00 * s) + (q01 * c); o11 = (q10 * s) + (q11 * c);
                    break;
                }
                case KQ_RZ: {
                    const auto c = std::cos(th / 2.0), s = std::sin(th / 2.0);
                    const auto plus = cd {c, s}, minus = cd {c, -s};
                    o00 = q00 * plus; o10 = q10 * plus; o01 = q01 * minus; o11 = q11 * minus;
                    break;
                }
                case KQ_P: {
                    const auto minus = cd {std::cos(th), -std::sin(th)};
                    o00 = q00; o10 = q10; o01 = q01 * minus; o11 = q11 * minus;
                    break;
                }
                case KQ_U:
                    o00 = (q00 * ua.e00) + (q01 * ua.e10); o10 = (q10 * ua.e00) + (q11 * ua.e10);
                    o01 = (q00 * ua.e01) + (q01 * ua.e11); o11 = (q10 * ua.e01) + (q11 * ua.e11);
                    break;
                default: break;
            }
            rho(r0, c0) = o00; rho(r1, c0) = o10; rho(r0, c1) = o01; rho(r1, c1) = o11;
        }
    }
}

/* apply_cu_gate_first_/second_ (operations_density_matrix.cpp:80-176), used for every
 * controlled gate (operations_density_matrix.hpp:412-600) */
void dm_gate_cu(DM& rho, DM& buf, const kq_op& op)
{
    const auto n = rho.n;
    if (n < 2) {
        return;
    }
    const auto c = static_cast<std::size_t>(op.q0), t = static_cast<std::size_t>(op.q1);
    const int g = op.gate;
    M2 mat;
    M2 mat2;
    if (g == KQ_CU) {
        mat = op_matrix(op);
        mat2 = adjoint(mat);
    }
    else {
        mat = mat_of(g, op.angle);
        /* CH, CX, CY, CZ pass the gate itself to the second pass (Hermitian); the rest pass the
         * conjugate transpose (operations_density_matrix.hpp:481-600) */
        mat2 = (g == KQ_CH || g == KQ_CX || g == KQ_CY || g == KQ_CZ) ? mat : adjoint(mat);
    }
    const auto dot = [](cd a0, cd a1, cd b0, cd b1) { return (a0 * b0) + (a1 * b1); };
    const auto np = n_double_pairs(n);
    TwoPairGen outer {c, t, n}, inner {c, t, n};
    outer.set_state(0);
    for (std::size_t io = 0; io < np; ++io) {
        const auto [col0, col2, col1, col3] = outer.next_unset_and_set();
        inner.set_state(0);
        for (std::size_t ii = 0; ii < np; ++ii) {
            const auto [row0, row2, row1, row3] = inner.next_unset_and_set();
            for (const auto col : {col0, col1, col2, col3}) {
                const auto x1 = rho(row1, col), x3 = rho(row3, col);
                buf(row0, col) = rho(row0, col);
                buf(row1, col) = dot(mat.e00, mat.e01, x1, x3);
                buf(row2, col) = rho(row2, col);
                buf(row3, col) = dot(mat.e10, mat.e11, x1, x3);
            }
        }
    }
    outer.set_state(0);
    for (std::size_t io = 0; io < np; ++io) {
        const auto [row0, row2, row1, row3] = outer.next_unset_and_set();
        inner.set_state(0);
        for (std::size_t ii = 0; ii < np; ++ii) {
            const auto [col0, col2, col1, col3] = inner.next_unset_and_set();
            for (const auto row : {row0, row1, row2, row3}) {
                const auto y1 = buf(row, col1), y3 = buf(row, col3);
                rho(row, col0) = buf(row, col0);
                rho(row, col1) = dot(mat2.e00, mat2.e10, y1, y3);
                rho(row, col2) = buf(row, col2);
                rho(row, col3) = dot(mat2.e01, mat2.e11, y1, y3);
            }
        }
    }
}

/* measure_density_matrix.cpp:21-95 and measure_density_matrix.hpp:33-57 */
int dm_measure(DM& rho, std::size_t q, bool has_seed, int seed)
{
    const auto n = rho.n;
    OnePairGen outer {q, n}, inner {q, n};
    double p0 = 0.0, p1 = 0.0;
    outer.set_state(0);
    for (std::size_t io = 0; io < outer.size(); ++io) {
        const auto [o0, o1] = outer.next();
        inner.set_state(0);
        for (std::size_t ii = 0; ii < outer.size(); ++ii) {
            const auto [i0, i1] = inner.next();
            p0 += std::norm(rho(i0, o0));
            p1 += std::norm(rho(i1, o1));
        }
    }
    auto prng = get_prng(has_seed, seed);
    auto flip = std::discrete_distribution<int> {{p0, p1}};
    const int outcome = flip(prng);
    const double norm = outcome == 0 ? std::sqrt(1.0 / p0) : std::sqrt(1.0 / p1);
    outer.set_state(0);
    for (std::size_t io = 0; io < outer.size(); ++io) {
        const auto [o0, o1] = outer.next();
        inner.set_state(0);
        for (std::size_t ii = 0; ii < outer.size(); ++ii) {
            const auto [i0, i1] = inner.next();
            rho(i1, o0) = {0.0, 0.0};
            rho(i0, o1) = {0.0, 0.0};
            if (outcome == 0) {
                rho(i0, o0) *= norm;
                rho(i1, o1) = {0.0, 0.0};
            }
            else {
                rho(i0, o0) = {0.0, 0.0};
                rho(i1, o1) *= norm;
            }
        }
    }
    return outcome;
}

/* The interpreter (simulate.cpp:417-521 / simulate_density_matrix.cpp:354-461), over the
 * flattened form of kq.h: an explicit stack of (begin, end) op ranges. */
template <typename GateFn, typename MeasureFn, typename LogFn>
void interpret(const kq_op* ops, size_t n_ops, const int32_t* preds, int8_t* creg, GateFn&& gate,
               MeasureFn&& measure, LogFn&& log)
{
    std::vector<std::pair<size_t, size_t>> stack;
    stack.emplace_back(0, n_ops);
    while (!stack.empty()) {
        auto& [ip, end] = stack.back();
        if (ip >= end) {
            stack.pop_back();
            continue;
        }
        const kq_op& op = ops[ip];
        const size_t here = ip;
        if (op.gate == KQ_OP_IF) {
            ip += 1 + static_cast<size_t>(op.q1);
            if (eval_predicate(preds, op.q0, creg)) {
                stack.emplace_back(here + 1, here + 1 + static_cast<size_t>(op.q1));
            }
        }
        else if (op.gate == KQ_OP_IF_ELSE) {
            const auto len_if = static_cast<size_t>(op.q1), len_else = static_cast<size_t>(op.aux);
            ip += 1 + len_if + len_else;
            if (eval_predicate(preds, op.q0, creg)) {
                stack.emplace_back(here + 1, here + 1 + len_if);
            }
            else {
                stack.emplace_back(here + 1 + len_if, here + 1 + len_if + len_else);
            }
        }
        else if (op.gate == KQ_OP_LOG_STATE || op.gate == KQ_OP_LOG_CREG) {
            ++ip;
            log(op.gate);
        }
        else if (op.gate == KQ_M) {
            ++ip;
            creg[op.q1] = static_cast<int8_t>(measure(static_cast<std::size_t>(op.q0)));
        }
        else {
            ++ip;
            gate(op);
        }
    }
}

template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return KQ_OK;
    }
    catch (const std::out_of_range& e) {
        g_err = e.what();
        return KQ_ERR_OUT_OF_RANGE;
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return KQ_ERR_RUNTIME;
    }
}

State load(const double* amps, std::size_t N)
{
    State s(N);
    std::memcpy(static_cast<void*>(s.data()), amps, N * sizeof(cd));
    return s;
}

void store(const State& s, double* amps) { std::memcpy(amps, s.data(), s.size() * sizeof(cd)); }

/* calculate_cumulative_sum_, cumulative_end_offset_, ProbabilitySampler_
 * (measurements.cpp:235-312) */
struct Sampler
{
    std::vector<double> cdf;
    std::mt19937 prng;
    std::uniform_real_distribution<double> dist;
    Sampler(const std::vector<double>& probs, bool has_seed, int seed) : prng {get_prng(has_seed, seed)}
    {
        cdf.reserve(probs.size());
        std::partial_sum(probs.begin(), probs.end(), std::back_inserter(cdf));
        const auto last = cdf.back();
        double second_last = 0.0;
        for (std::size_t i = cdf.size() - 1; i > 0; --i) {
            if (cdf[i - 1] < last) {
                second_last = cdf[i - 1];
                break;
            }
        }
        const auto offset = (last - second_last) * 1.0e-4;
        dist = std::uniform_real_distribution<double> {0.0, last - offset};
    }
    std::size_t operator()()
    {
        const auto u = dist(prng);
        const auto it = std::lower_bound(cdf.begin(), cdf.end(), u);
        if (it == cdf.end()) {
            throw std::runtime_error {"LOGIC BUG: Ended up with measurement of state past end of cumulative\nprobability distribution, which shouldn't happen?"};
        }
        return static_cast<std::size_t>(it - cdf.begin());
    }
};

/* apply_noise_ (probabilities.cpp:110-124) */
void apply_noise(double noise, std::size_t q, std::size_t n, std::vector<double>& p)
{
    OnePairGen gen {q, n};
    gen.set_state(0);
    for (std::size_t i = 0; i < gen.size(); ++i) {
        const auto [a, b] = gen.next();
        const auto pa = p[a], pb = p[b];
        p[a] = ((1.0 - noise) * pa) + (noise * pb);
        p[b] = ((1.0 - noise) * pb) + (noise * pa);
    }
}

/* simulate_pauli_gates_ (simulate_pauli.cpp:58-86) */
void apply_pauli_terms(State& s, std::size_t n, const uint32_t* q, const uint8_t* k, std::size_t nt)
{
    for (std::size_t j = 0; j < nt; ++j) {
        if (k[j] == 0) {
            continue;
        }
        OnePairGen gen {q[j], n};
        gen.set_state(0);
        for (std::size_t i = 0; i < gen.size(); ++i) {
            const auto [a, b] = gen.next();
            if (k[j] == 1) { op_x(s, a, b); }
            else if (k[j] == 2) { op_y(s, a, b); }
            else { op_z(s, b); }
        }
    }
}

const cd PHASES[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};

cd expectation(const State& state, std::size_t n, size_t n_strings, const uint32_t* off,
               const uint32_t* q, const uint8_t* k, const uint8_t* phases, const double* coeffs)
{
    cd expval {};
    for (size_t s = 0; s < n_strings; ++s) {
        auto ket = state;
        apply_pauli_terms(ket, n, q + off[s], k + off[s], off[s + 1] - off[s]);
        cd inner {};
        for (std::size_t i = 0; i < state.size(); ++i) {
            inner += (std::conj(state[i]) * ket[i]);
        }
        const auto coeff = cd {coeffs[2 * s], coeffs[2 * s + 1]};
        expval += (coeff * PHASES[phases[s]] * inner);
    }
    return expval;
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

/* flat pair index k -> (s0, s1) in generator order, for every k */
int oracle_pair_map_1q(uint32_t n, uint32_t t, uint64_t* s0, uint64_t* s1)
{
    return guarded([&] {
        OnePairGen gen {t, n};
        gen.set_state(0);
        for (std::size_t i = 0; i < gen.size(); ++i) {
            const auto [a, b] = gen.next();
            s0[i] = a;
            s1[i] = b;
        }
    });
}

/* set_state(k) then next() for one k: checks the random-access form of the generator */
int oracle_pair_at_1q(uint32_t n, uint32_t t, uint64_t k, uint64_t* s0, uint64_t* s1)
{
    OnePairGen gen {t, n};
    gen.set_state(k);
    const auto [a, b] = gen.next();
    *s0 = a;
    *s1 = b;
    return KQ_OK;
}

int oracle_pair_map_2q(uint32_t n, uint32_t c, uint32_t t, int four_way, uint64_t* out)
{
    return guarded([&] {
        TwoPairGen gen {c, t, n};
        gen.set_state(0);
        const auto np = n_double_pairs(n);
        for (std::size_t i = 0; i < np; ++i) {
            if (four_way) {
                const auto [a, b, cc, d] = gen.next_unset_and_set();
                out[4 * i] = a; out[4 * i + 1] = b; out[4 * i + 2] = cc; out[4 * i + 3] = d;
            }
            else {
                const auto [a, b] = gen.next();
                out[2 * i] = a;
                out[2 * i + 1] = b;
            }
        }
    });
}

int oracle_pair_at_2q(uint32_t n, uint32_t c, uint32_t t, uint64_t k, uint64_t* s0, uint64_t* s1)
{
    TwoPairGen gen {c, t, n};
    gen.set_state(k);
    const auto [a, b] = gen.next();
    *s0 = a;
    *s1 = b;
    return KQ_OK;
}

/* ket::simulate(circuit, Statevector&, seed) over the flattened circuit of kq.h */
int oracle_sv_simulate(double* amps, uint32_t n, uint32_t circuit_n, const kq_op* ops, size_t n_ops,
                       const int32_t* preds, int has_seed, int32_t seed, int8_t* creg, size_t n_bits,
                       kq_logger_fn logger, void* user)
{
    return guarded([&] {
        if (circuit_n != n) {
            throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
        }
        if (n == 0) {
            throw std::runtime_error {"Cannot simulate a circuit or state with zero qubits."};
        }
        const std::size_t N = std::size_t {1} << n;
        auto s = load(amps, N);
        for (size_t i = 0; i < n_bits; ++i) {
            creg[i] = -1;
        }
        interpret(
            ops, n_ops, preds, creg, [&](const kq_op& op) { sv_gate(s, n, op); },
            [&](std::size_t q) { return sv_measure(s, n, q, has_seed != 0, seed); },
            [&](int kind) {
                if (logger) {
                    logger(user, kind, reinterpret_cast<const double*>(s.data()), creg, n_bits);
                }
            });
        store(s, amps);
    });
}

/* ket::simulate(circuit, DensityMatrix&, seed) (simulate_density_matrix.cpp:354-548) */
int oracle_dm_simulate(double* rho_cm, uint32_t n, uint32_t circuit_n, const kq_op* ops, size_t n_ops,
                       const int32_t* preds, int has_seed, int32_t seed, int8_t* creg, size_t n_bits)
{
    return guarded([&] {
        if (circuit_n != n) {
            throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
        }
        if (n == 0) {
            throw std::runtime_error {"Cannot simulate a circuit or state with zero qubits."};
        }
        const std::size_t N = std::size_t {1} << n;
        DM rho {n, N, load(rho_cm, N * N)};
        DM buf {n, N, std::vector<cd>(N * N)};
        for (size_t i = 0; i < n_bits; ++i) {
            creg[i] = -1;
        }
        interpret(
            ops, n_ops, preds, creg,
            [&](const kq_op& op) {
                if (op.gate <= KQ_P || op.gate == KQ_U) {
                    dm_gate_1t(rho, buf, op);
                }
                else {
                    dm_gate_cu(rho, buf, op);
                }
            },
            [&](std::size_t q) { return dm_measure(rho, q, has_seed != 0, seed); },
            [&](int) { throw std::runtime_error {"DEV ERROR: loggers haven't been implemented yet for the density matrix simulator.\n"}; });
        store(rho.m, rho_cm);
    });
}

/* CPU-baseline timer: applies the gates to |0...0> of n qubits and reports the seconds spent in
 * the gate loop only (std::chrono::steady_clock), excluding allocation. */
int oracle_time_gates(uint32_t n, const kq_op* ops, size_t n_ops, double* seconds)
{
    return guarded([&] {
        State s(std::size_t {1} << n, cd {0.0, 0.0});
        s[0] = {1.0, 0.0};
        const auto t0 = std::chrono::steady_clock::now();
        for (size_t i = 0; i < n_ops; ++i) {
            sv_gate(s, n, ops[i]);
        }
        const auto t1 = std::chrono::steady_clock::now();
        *seconds = std::chrono::duration<double>(t1 - t0).count();
    });
}

/* CPU baseline timing: each op applied `reps` times to one state of n qubits (|0...0> with a Hadamard
   on every qubit, so no amplitude is zero); seconds[i * reps + r] is the r-th run of op i */
int oracle_time_gate_reps(uint32_t n, const kq_op* ops, size_t n_ops, uint32_t reps, double* seconds)
{
    return guarded([&] {
        const std::size_t N = std::size_t {1} << n;
        State s(N, cd {1.0 / std::sqrt(static_cast<double>(N)), 0.0});
        for (size_t i = 0; i < n_ops; ++i) {
            for (uint32_t r = 0; r < reps; ++r) {
                const auto t0 = std::chrono::steady_clock::now();
                sv_gate(s, n, ops[i]);
                const auto t1 = std::chrono::steady_clock::now();
                seconds[i * reps + r] = std::chrono::duration<double>(t1 - t0).count();
            }
        }
    });
}

/* calculate_probabilities_raw (probabilities.cpp:49-71) */
int oracle_probabilities(const double* amps, uint32_t n, const double* noise, double* out)
{
    return guarded([&] {
        const std::size_t N = std::size_t {1} << n;
        const auto s = load(amps, N);
        std::vector<double> p;
        p.reserve(N);
        for (std::size_t i = 0; i < N; ++i) {
            p.push_back(std::norm(s[i]));
        }
        if (noise != nullptr) {
            for (std::size_t q = 0; q < n; ++q) {
                apply_noise(noise[q], q, n, p);
            }
        }
        std::memcpy(out, p.data(), N * sizeof(double));
    });
}

/* perform_measurements_as_memory(probabilities_raw, n_shots, seed) (measurements.cpp:79-97) */
int oracle_sample_probabilities(const double* probs, uint64_t n_probs, uint64_t n_shots, int has_seed,
                                int32_t seed, uint64_t* out)
{
    return guarded([&] {
        std::vector<double> p(probs, probs + n_probs);
        Sampler sampler {p, has_seed != 0, seed};
        for (uint64_t i = 0; i < n_shots; ++i) {
            out[i] = sampler();
        }
    });
}

/* the cumulative distribution and the sampler's upper bound, for tests of the device scan */
int oracle_cdf(const double* probs, uint64_t n_probs, double* cdf_out, double* upper)
{
    return guarded([&] {
        std::vector<double> p(probs, probs + n_probs);
        Sampler sampler {p, true, 0};
        std::memcpy(cdf_out, sampler.cdf.data(), n_probs * sizeof(double));
        *upper = sampler.dist.b();
    });
}

/* expectation_value(PauliOperator, Statevector) (pauli_operator.cpp:76-91) */
int oracle_pauli_expectation(const double* amps, uint32_t n, size_t n_strings, const uint32_t* off,
                             const uint32_t* q, const uint8_t* k, const uint8_t* phases,
                             const double* coeffs, double* out)
{
    return guarded([&] {
        const auto s = load(amps, std::size_t {1} << n);
        const auto e = expectation(s, n, n_strings, off, q, k, phases, coeffs);
        out[0] = e.real();
        out[1] = e.imag();
    });
}

/* ket::simulate(SparsePauliString, Statevector&) (simulate_pauli.cpp:104-128) */
int oracle_apply_pauli_string(double* amps, uint32_t n, size_t nt, const uint32_t* q, const uint8_t* k)
{
    return guarded([&] {
        auto s = load(amps, std::size_t {1} << n);
        apply_pauli_terms(s, n, q, k, nt);
        store(s, amps);
    });
}

/* statevector_to_density_matrix (density_matrix.cpp:170-185) */
int oracle_sv_to_dm(const double* amps, uint32_t n, double* rho_cm)
{
    return guarded([&] {
        const std::size_t N = std::size_t {1} << n;
        const auto s = load(amps, N);
        DM rho {n, N, std::vector<cd>(N * N)};
        for (std::size_t i0 = 0; i0 < N; ++i0) {
            for (std::size_t i1 = 0; i1 < N; ++i1) {
                rho(i0, i1) = s[i0] * std::conj(s[i1]);
            }
        }
        store(rho.m, rho_cm);
    });
}

/* DensityMatrix::trace (density_matrix.cpp:158-161; sum of the diagonal in index order) */
int oracle_dm_trace(const double* rho_cm, uint32_t n, double* out)
{
    return guarded([&] {
        const std::size_t N = std::size_t {1} << n;
        DM rho {n, N, load(rho_cm, N * N)};
        cd t {};
        for (std::size_t i = 0; i < N; ++i) {
            t += rho(i, i);
        }
        out[0] = t.real();
        out[1] = t.imag();
    });
}

/* DensityMatrix::trace_of_square (density_matrix.cpp:164-168): the trace of rho * rho, i.e.
   sum_i sum_k rho(i, k) rho(k, i), with the inner sum over k for each diagonal entry i */
int oracle_dm_trace_of_square(const double* rho_cm, uint32_t n, double* out)
{
    return guarded([&] {
        const std::size_t N = std::size_t {1} << n;
        DM rho {n, N, load(rho_cm, N * N)};
        cd t {};
        for (std::size_t i = 0; i < N; ++i) {
            cd d {};
            for (std::size_t k = 0; k < N; ++k) {
                d += rho(i, k) * rho(k, i);
            }
            t += d;
        }
        out[0] = t.real();
        out[1] = t.imag();
    });
}

/* partial_trace (density_matrix.cpp:71-112, 231-278) */
int oracle_dm_partial_trace(const double* rho_cm, uint32_t n, const uint32_t* qubits, size_t nq, double* out)
{
    return guarded([&] {
        std::vector<std::size_t> qs(qubits, qubits + nq);
        for (auto q : qs) {
            if (q >= n) {
                throw std::runtime_error {"ERROR: Cannot take partial trace; found qubit index out of bounds\n"};
            }
        }
        std::sort(qs.begin(), qs.end(), [](auto x, auto y) { return x > y; });
        std::size_t cur_n = n;
        std::vector<cd> cur = load(rho_cm, (std::size_t {1} << n) << n);
        for (std::size_t i = 0; i < qs.size(); ++i) {
            const auto q = qs[i];
            const std::size_t step = std::size_t {1} << q;
            const std::size_t Ncur = std::size_t {1} << cur_n;
            const std::size_t n_blocks = std::size_t {1} << (cur_n - q - 1);
            const std::size_t per_block = step;
            const std::size_t Nnext = Ncur / 2;
            std::vector<cd> red(Nnext * Nnext);
            const auto gen = [&](std::size_t j) { return (j / per_block) * (2 * step) + (j % per_block); };
            for (std::size_t ic = 0; ic < n_blocks * per_block; ++ic) {
                const auto icc = gen(ic);
                for (std::size_t ir = 0; ir < n_blocks * per_block; ++ir) {
                    const auto irc = gen(ir);
                    red[ir + ic * Nnext] = cur[irc + icc * Ncur] + cur[(irc + step) + (icc + step) * Ncur];
                }
            }
            cur = std::move(red);
            --cur_n;
        }
        std::memcpy(out, cur.data(), cur.size() * sizeof(cd));
    });
}

}  // extern "C"
