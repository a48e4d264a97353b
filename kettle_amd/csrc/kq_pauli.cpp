// Pauli strings (canonical form, grouping by flip mask, expectation and application) and the
// density-matrix read-outs (trace, partial trace, |psi><psi|).
#include <algorithm>
#include <unordered_map>
#include <functional>
#include <bit>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <optional>
#include <random>
#include <string>
#include <vector>

#include "kq_internal.hpp"

#include "kq_host.hpp"

namespace kq
{

// ---- Pauli strings ----------------------------------------------------------------------------------

struct Canon
{
    std::complex<double> c {1.0, 0.0};
    uint64_t x = 0;
    uint64_t z = 0;
};

// Product of the string's terms in kettle's application order (simulate_pauli.cpp:66-85):
// the operator is T_m ... T_1 with T_1 applied first.  Each term is (c, x, z) acting as
// (P psi)_i = c (-1)^{|i & z|} psi_{i ^ x}; X = (1, m, 0), Z = (1, 0, m), Y = (-i, m, m).
// Composing Q after P gives c_Q c_P (-1)^{|x_Q & z_P|}, x_Q ^ x_P, z_Q ^ z_P.
Canon canonical(const uint32_t* q, const uint8_t* k, size_t nt, uint32_t n)
{
    Canon p;
    for (size_t j = 0; j < nt; ++j) {
        if (k[j] == 0) {
            continue;
        }
        if (k[j] > 3) {
            throw InvalidArg {"Pauli term kind must be 0..3"};
        }
        if (q[j] >= n) {
            throw InvalidArg {"Pauli term qubit out of range"};
        }
        const uint64_t m = uint64_t {1} << q[j];
        std::complex<double> cq {1.0, 0.0};
        uint64_t xq = 0, zq = 0;
        if (k[j] == 1) {
            xq = m;
        }
        else if (k[j] == 2) {
            xq = m;
            zq = m;
            cq = {0.0, -1.0};
        }
        else {
            zq = m;
        }
        const double sign = (std::popcount(xq & p.z) & 1) ? -1.0 : 1.0;
        p.c = cq * p.c * sign;
        p.x ^= xq;
        p.z ^= zq;
    }
    return p;
}

const std::complex<double> PHASES[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};


void build_pauli(uint32_t n, size_t n_strings, const uint32_t* off, const uint32_t* q, const uint8_t* k,
                 const uint8_t* phases, const double* coeffs, std::vector<PauliGroup>& groups,
                 std::vector<PauliZ>& strings)
{
    if (n_strings > 0 && (off == nullptr || phases == nullptr || coeffs == nullptr)) {
        throw InvalidArg {"null Pauli operator arrays"};
    }
    std::map<uint64_t, std::vector<PauliZ>> by_x;
    for (size_t s = 0; s < n_strings; ++s) {
        if (off[s + 1] < off[s]) {
            throw InvalidArg {"term_offsets must be non-decreasing"};
        }
        if (phases[s] > 3) {
            throw InvalidArg {"Pauli phase must be 0..3"};
        }
        const Canon cn = canonical(q + off[s], k + off[s], off[s + 1] - off[s], n);
        const std::complex<double> w = std::complex<double> {coeffs[2 * s], coeffs[2 * s + 1]} * PHASES[phases[s]] * cn.c;
        by_x[cn.x].push_back(PauliZ {cn.z, {w.real(), w.imag()}});
    }
    groups.clear();
    strings.clear();
    for (auto& [x, zs] : by_x) {
        groups.push_back(PauliGroup {x, static_cast<uint32_t>(strings.size()), static_cast<uint32_t>(zs.size())});
        strings.insert(strings.end(), zs.begin(), zs.end());
    }
}

}  // namespace kq

using namespace kq;

extern "C" {

int kq_pauli_expectation(kq_state* s, size_t n_strings, const uint32_t* term_offsets, const uint32_t* term_qubits,
                         const uint8_t* term_kinds, const uint8_t* phases, const double* coeffs, double out[2])
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (s->is_dm) {
            throw InvalidArg {"Pauli expectation: statevector states only"};
        }
        ensure_init();
        std::vector<PauliGroup> groups;
        std::vector<PauliZ> strings;
        build_pauli(s->n, n_strings, term_offsets, term_qubits, term_kinds, phases, coeffs, groups, strings);
        if (groups.empty()) {
            out[0] = 0.0;
            out[1] = 0.0;
            return;
        }
        launch_pauli_expectation(s, groups, strings, out);
    });
}

int kq_apply_pauli_string(kq_state* s, size_t n_terms, const uint32_t* term_qubits, const uint8_t* term_kinds)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (s->is_dm) {
            throw InvalidArg {"Pauli string: statevector states only"};
        }
        ensure_init();
        const Canon cn = canonical(term_qubits, term_kinds, n_terms, s->n);
        const double w[2] = {cn.c.real(), cn.c.imag()};
        launch_pauli_apply_one(s, cn.x, cn.z, w);
        KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int kq_dm_trace(kq_state* s, double out[2])
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (!s->is_dm) {
            throw InvalidArg {"trace: density matrices only"};
        }
        ensure_init();
        launch_dm_trace(s, out);
    });
}

int kq_dm_trace_of_square(kq_state* s, double out[2])
{
    return guarded([&] {
        check_state(s);
        if (!s->is_dm) {
            throw InvalidArg {"trace of square: density matrices only"};
        }
        if (!out) {
            throw InvalidArg {"null output"};
        }
        ensure_init();
        launch_dm_trace_of_square(s, out);
    });
}

int kq_dm_partial_trace(kq_state* s, const uint32_t* qubits, size_t n, double* out)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (!s->is_dm) {
            throw InvalidArg {"partial trace: density matrices only"};
        }
        std::vector<uint32_t> qs(qubits, qubits + n);
        for (auto x : qs) {
            if (x >= s->n) {
                throw std::runtime_error {"ERROR: Cannot take partial trace; found qubit index out of bounds\n"};
            }
        }
        std::sort(qs.begin(), qs.end(), [](auto a, auto b) { return a > b; });
        if (std::adjacent_find(qs.begin(), qs.end()) != qs.end()) {
            throw InvalidArg {"partial trace: repeated qubit"};
        }
        ensure_init();
        launch_dm_partial_trace(s, qs, out);
    });
}

int kq_dm_from_statevector(kq_state* dm, const kq_state* sv)
{
    return guarded([&] {
        check_state(dm);
        check_state(sv);
        if (!dm->is_dm || sv->is_dm || dm->n != sv->n || dm->precision != sv->precision) {
            throw InvalidArg {"dm_from_statevector: shapes differ"};
        }
        ensure_init();
        normalise_layout(sv);
        dm->bit_of.clear();
        KQ_HIP_CHECK(hipStreamSynchronize(sv->stream));
        launch_dm_from_sv(dm, sv);
        KQ_HIP_CHECK(hipStreamSynchronize(dm->stream));
    });
}

}  // extern "C"
