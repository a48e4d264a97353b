// Mid-circuit measurement, the simulate() interpreter (control flow, loggers), probabilities and
// sampling, and the per-shot re-simulation over shared outcome histories.
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

// ---- measurement ---------------------------------------------------------------------------------

// get_prng_ (common/prng.cpp:9-19)
std::mt19937 get_prng(bool has_seed, int32_t seed)
{
    if (has_seed) {
        return std::mt19937 {static_cast<std::mt19937::result_type>(seed)};
    }
    auto device = std::random_device {};
    return std::mt19937 {device()};
}

// simulate_measurement_ (measure.hpp:33-57; measure_density_matrix.hpp:33-57): weights on the
// device, the coin flip with kettle's std::discrete_distribution<int> on the host, collapse on
// the device with scale sqrt(1 / w).
int measure(kq_state* s, uint32_t q, bool has_seed, int32_t seed)
{
    if (q >= s->n) {
        throw InvalidArg {"measured qubit out of range"};
    }
    normalise_layout(s);
    double w[2];
    launch_qubit_weights(s, q, w);
    auto prng = get_prng(has_seed, seed);
    auto flip = std::discrete_distribution<int> {{w[0], w[1]}};
    const int outcome = flip(prng);
    const double norm = outcome == 0 ? std::sqrt(1.0 / w[0]) : std::sqrt(1.0 / w[1]);
    launch_collapse(s, q, outcome, norm);
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    return outcome;
}

// ControlFlowPredicate::operator() (control_flow_predicate.cpp:40-69)
bool eval_predicate(const int32_t* preds, size_t n_preds, int32_t off, const int8_t* creg, size_t n_bits)
{
    if (off < 0 || static_cast<size_t>(off) + 2 > n_preds) {
        throw InvalidArg {"predicate offset out of range"};
    }
    const int kind = preds[off];
    const int n = preds[off + 1];
    if (n <= 0 || static_cast<size_t>(off) + 2 + 2 * static_cast<size_t>(n) > n_preds) {
        throw InvalidArg {"malformed predicate"};
    }
    bool out = true;
    for (int i = 0; i < n; ++i) {
        const int bit = preds[off + 2 + 2 * i];
        const int expected = preds[off + 3 + 2 * i];
        if (bit < 0 || static_cast<size_t>(bit) >= n_bits) {
            throw InvalidArg {"predicate bit out of range"};
        }
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

std::vector<double> download_host(const kq_state* s)
{
    std::vector<double> h(2 * s->n_amps);
    if (s->precision == KQ_C128) {
        KQ_HIP_CHECK(hipMemcpy(h.data(), s->amps, s->n_amps * 16, hipMemcpyDeviceToHost));
    }
    else {
        std::vector<float> f(2 * s->n_amps);
        KQ_HIP_CHECK(hipMemcpy(f.data(), s->amps, s->n_amps * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); ++i) {
            h[i] = f[i];
        }
    }
    return h;
}

// simulate_loop_body_iterative_ (simulate.cpp:417-521): an explicit stack of op ranges.
// Consecutive gates of one range are compiled, fused and launched together; the batch is
// flushed at every measurement, control-flow element and logger.
// Interpreter position in a flat program: a stack of [next, end) op ranges (bodies of taken
// branches are pushed, kettle's simulate_loop_body_iterative_ at simulate.cpp:417-521).
using Cursor = std::vector<std::pair<size_t, size_t>>;

// Runs the program from `cur` on `s`, fusing gate runs, until the next measurement (returned, with
// the cursor past it) or the end (nullptr).  Control flow reads `creg`; loggers are called when given.
const kq_op* run_to_measurement(kq_state* s, Cursor& cur, const kq_op* ops, const int32_t* preds, size_t n_preds,
                                const int8_t* creg, size_t n_bits, kq_logger_fn logger, void* user)
{
    std::vector<GateRec> batch;
    const auto flush = [&] {
        run_records(s, batch);
        batch.clear();
    };
    while (!cur.empty()) {
        auto& top = cur.back();
        if (top.first >= top.second) {
            cur.pop_back();
            continue;
        }
        const size_t here = top.first;
        const kq_op& op = ops[here];
        const auto body_end = [&](size_t len) {
            if (here + 1 + len > top.second) {
                throw InvalidArg {"control-flow body overruns its enclosing range"};
            }
        };
        if (op.gate == KQ_OP_IF) {
            const auto len = static_cast<size_t>(op.q1);
            body_end(len);
            top.first += 1 + len;
            flush();
            if (eval_predicate(preds, n_preds, op.q0, creg, n_bits)) {
                cur.emplace_back(here + 1, here + 1 + len);
            }
        }
        else if (op.gate == KQ_OP_IF_ELSE) {
            const auto len_if = static_cast<size_t>(op.q1), len_else = static_cast<size_t>(op.aux);
            body_end(len_if + len_else);
            top.first += 1 + len_if + len_else;
            flush();
            if (eval_predicate(preds, n_preds, op.q0, creg, n_bits)) {
                cur.emplace_back(here + 1, here + 1 + len_if);
            }
            else {
                cur.emplace_back(here + 1 + len_if, here + 1 + len_if + len_else);
            }
        }
        else if (op.gate == KQ_OP_LOG_STATE || op.gate == KQ_OP_LOG_CREG) {
            ++top.first;
            if (s->is_dm) {
                throw std::runtime_error {"DEV ERROR: loggers haven't been implemented yet for the density matrix simulator.\n"};
            }
            flush();
            if (logger != nullptr) {
                if (op.gate == KQ_OP_LOG_STATE) {
                    normalise_layout(s);
                    const auto h = download_host(s);
                    logger(user, op.gate, h.data(), creg, n_bits);
                }
                else {
                    logger(user, op.gate, nullptr, creg, n_bits);
                }
            }
        }
        else if (op.gate == KQ_M) {
            ++top.first;
            if (op.q1 < 0 || static_cast<size_t>(op.q1) >= n_bits) {
                throw InvalidArg {"measurement bit out of range"};
            }
            flush();
            return &op;
        }
        else {
            ++top.first;
            if (op.gate < 0 || op.gate > KQ_CU) {
                throw InvalidArg {"unknown op code " + std::to_string(op.gate)};
            }
            compile_gate(op, s->n, s->is_dm != 0, batch);
        }
    }
    flush();
    return nullptr;
}

void simulate(kq_state* s, uint32_t circuit_n, const kq_op* ops, size_t n_ops, const int32_t* preds, size_t n_preds,
              bool has_seed, int32_t seed, int8_t* creg, size_t n_bits, kq_logger_fn logger, void* user)
{
    if (circuit_n != s->n) {
        throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
    }
    if (n_bits > 0 && creg == nullptr) {
        throw InvalidArg {"null classical register"};
    }
    for (size_t i = 0; i < n_bits; ++i) {
        creg[i] = -1;
    }
    Cursor cur {{0, n_ops}};
    while (const kq_op* m = run_to_measurement(s, cur, ops, preds, n_preds, creg, n_bits, logger, user)) {
        creg[m->q1] = static_cast<int8_t>(measure(s, static_cast<uint32_t>(m->q0), has_seed, seed));
    }
}

// ---- sampling ---------------------------------------------------------------------------------------

// ProbabilitySampler_ (measurements.cpp:288-312) with the CDF on the device.  The uniform draws
// use kettle's std::uniform_real_distribution(0, last - offset) on std::mt19937 (libstdc++), so
// the draw sequence is kettle's; lower_bound runs per shot on the device.
// The sampler over device probabilities d_p (scanned into their CDF in place) with a caller-provided
// device workspace `sc` of sc_bytes (CDF partials, the uniform draws, the drawn indices).  The host
// draws are staged through `out` itself, which the indices then overwrite (same stream, in order).
void sample_with_workspace(hipStream_t st, double* d_p, uint64_t n, uint64_t shots, bool has_seed, int32_t seed,
                           uint64_t* out, void* sc, size_t sc_bytes)
{
    const size_t sb = (cdf_scratch_bytes(n) + 255) / 256 * 256;
    if (sc_bytes < sb + shots * (sizeof(double) + sizeof(uint64_t))) {
        throw InvalidArg {"sampler: workspace too small"};
    }
    double last = 0.0, second_last = 0.0;
    launch_cdf(st, d_p, n, sc, sb, &last, &second_last);
    const double offset = (last - second_last) * 1.0e-4;
    auto prng = get_prng(has_seed, seed);
    auto dist = std::uniform_real_distribution<double> {0.0, last - offset};
    for (uint64_t i = 0; i < shots; ++i) {
        const double u = dist(prng);
        std::memcpy(out + i, &u, sizeof u);
    }
    auto* d_u = reinterpret_cast<double*>(static_cast<char*>(sc) + sb);
    auto* d_out = reinterpret_cast<uint64_t*>(d_u + shots);
    if (shots) {
        KQ_HIP_CHECK(hipMemcpyAsync(d_u, out, shots * sizeof(double), hipMemcpyHostToDevice, st));
        launch_lower_bound(st, d_p, n, d_u, shots, sc, d_out);
        KQ_HIP_CHECK(hipMemcpyAsync(out, d_out, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    }
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    for (uint64_t i = 0; i < shots; ++i) {
        if (out[i] >= n) {
            throw std::runtime_error {"LOGIC BUG: Ended up with measurement of state past end of cumulative\nprobability distribution, which shouldn't happen?"};
        }
    }
}

void sample_from_device_probs(hipStream_t st, double* d_p, uint64_t n, uint64_t shots, bool has_seed, int32_t seed,
                              uint64_t* out)
{
    const size_t bytes = (cdf_scratch_bytes(n) + 255) / 256 * 256 + shots * (sizeof(double) + sizeof(uint64_t));
    void* sc = device_alloc(bytes);
    try {
        sample_with_workspace(st, d_p, n, shots, has_seed, seed, out, sc, bytes);
    }
    catch (...) {
        (void) hipFree(sc);
        throw;
    }
    KQ_HIP_CHECK(hipFree(sc));
}

void check_noise(const double* noise, uint32_t n)
{
    if (noise == nullptr) {
        return;
    }
    for (uint32_t q = 0; q < n; ++q) {
        if (!(0.0 <= noise[q] && noise[q] <= 1.0)) {
            throw std::runtime_error {"ERROR: Cannot set probability in QuantumNoise object outside of [0, 1]."};
        }
    }
}

// ---- per-shot re-simulation ----------------------------------------------------------------------

// perform_measurements_as_counts_marginal(circuit, state, ...) (measurements.cpp:178-208) runs the
// whole circuit once per shot.  Between measurements a run is deterministic given the classical
// register, so shots with the same outcome history reach the same state: a branch node holds the
// state at a measurement (or at the end) for one history, and every distinct history is simulated
// once while its state fits the cache (an evicted branch is rebuilt from its nearest cached ancestor).
struct ShotNode
{
    kq_state* state = nullptr;   // the state at this point (owned), or null once it did not fit the cache
    Cursor cursor;               // program position right after the measurement op (or the end)
    std::vector<int8_t> creg;
    int32_t m_qubit = -1;        // the measurement reached here (-1: the program ended)
    int32_t m_bit = -1;
    double w[2] = {0.0, 0.0};    // its outcome weights
    int64_t child[2] = {-1, -1};
    int64_t parent = -1;
    int outcome = -1;            // the parent's outcome that leads here
    uint64_t shots = 0;          // shots that end here
};

struct ShotTree
{
    const kq_op* ops;
    const int32_t* preds;
    size_t n_preds;
    size_t n_bits;
    uint32_t n;
    int precision;
    uint64_t cache_left;
    uint64_t state_bytes;
    std::vector<ShotNode> nodes;

    ~ShotTree()
    {
        for (auto& nd : nodes) {
            if (nd.state) {
                (void) kq_state_destroy(nd.state);
            }
        }
    }

    // runs from the node's cursor to the next measurement and records it (and its weights)
    void settle(ShotNode& nd, kq_state* st)
    {
        const kq_op* m = run_to_measurement(st, nd.cursor, ops, preds, n_preds, nd.creg.data(), n_bits, nullptr, nullptr);
        normalise_layout(st);   // branch states are kept and copied with qubit q on bit q
        if (m) {
            if (m->q0 < 0 || static_cast<uint32_t>(m->q0) >= n) {
                throw InvalidArg {"measured qubit out of range"};
            }
            nd.m_qubit = m->q0;
            nd.m_bit = m->q1;
            launch_qubit_weights(st, static_cast<uint32_t>(m->q0), nd.w);
        }
    }

    // keeps `st` as the node's state if the cache has room, else frees it
    void keep_or_free(ShotNode& nd, kq_state* st)
    {
        if (cache_left >= state_bytes) {
            cache_left -= state_bytes;
            nd.state = st;
        }
        else {
            (void) kq_state_destroy(st);
        }
    }

    // a fresh state holding node i's state (the caller owns it)
    kq_state* materialise(int64_t i)
    {
        kq_state* st = create_state(n, precision, false);
        try {
            const ShotNode& nd = nodes[static_cast<size_t>(i)];
            if (nd.state) {
                KQ_HIP_CHECK(hipStreamSynchronize(nd.state->stream));
                launch_copy_kernel(st, st->amps, nd.state->amps, st->n_amps);
                KQ_HIP_CHECK(hipStreamSynchronize(st->stream));
                return st;
            }
            (void) kq_state_destroy(st);
            st = materialise(nd.parent);
            step(nodes[static_cast<size_t>(nd.parent)], nd.outcome, st, nullptr);
            return st;
        }
        catch (...) {
            (void) kq_state_destroy(st);
            throw;
        }
    }

    // collapses `st` (holding parent p's state) on outcome b and runs it to the next measurement;
    // fills `out` with the child's bookkeeping when given
    void step(const ShotNode& p, int b, kq_state* st, ShotNode* out)
    {
        const double norm = std::sqrt(1.0 / p.w[b]);
        launch_collapse(st, static_cast<uint32_t>(p.m_qubit), b, norm);
        ShotNode tmp;
        ShotNode& nd = out ? *out : tmp;
        nd.cursor = p.cursor;
        nd.creg = p.creg;
        nd.creg[static_cast<size_t>(p.m_bit)] = static_cast<int8_t>(b);
        settle(nd, st);
    }

    int64_t child(int64_t i, int b)
    {
        if (nodes[static_cast<size_t>(i)].child[b] >= 0) {
            return nodes[static_cast<size_t>(i)].child[b];
        }
        kq_state* st = materialise(i);
        ShotNode nd;
        try {
            step(nodes[static_cast<size_t>(i)], b, st, &nd);
        }
        catch (...) {
            (void) kq_state_destroy(st);
            throw;
        }
        nd.parent = i;
        nd.outcome = b;
        keep_or_free(nd, st);
        nodes.push_back(std::move(nd));
        const auto id = static_cast<int64_t>(nodes.size() - 1);
        nodes[static_cast<size_t>(i)].child[b] = id;
        return id;
    }
};

void sample_circuit_shots(const kq_state* init, uint32_t circuit_n, const kq_op* ops, size_t n_ops, const int32_t* preds,
                          size_t n_preds, size_t n_bits, const double* noise, uint64_t n_shots, bool has_seed, int32_t seed,
                          uint64_t max_cached_bytes, uint64_t* out, uint64_t* n_branches)
{
    if (circuit_n != init->n) {
        throw std::runtime_error {"Invalid simulation; circuit and state have different number of qubits."};
    }
    if (init->is_dm) {
        throw InvalidArg {"per-shot sampling takes a statevector"};
    }
    check_noise(noise, init->n);
    ShotTree t {ops, preds, n_preds, n_bits, init->n, init->precision, 0,
                init->n_amps * amp_bytes(init->precision), {}};
    // the root state always stays
    {
        kq_state* st = create_state(init->n, init->precision, false);
        t.nodes.push_back(ShotNode {});
        t.nodes[0].state = st;
        KQ_HIP_CHECK(hipStreamSynchronize(init->stream));
        launch_copy_kernel(st, st->amps, init->amps, st->n_amps);
        t.nodes[0].cursor = Cursor {{0, n_ops}};
        t.nodes[0].creg.assign(n_bits, -1);
        t.settle(t.nodes[0], st);
    }
    t.cache_left = max_cached_bytes;
    // mid-circuit outcomes: independent draws for every shot, as kettle's simulate() without a seed
    // draws each measurement from a std::random_device-seeded generator
    std::mt19937_64 rng {std::random_device {}()};
    std::vector<int64_t> leaf(n_shots);
    for (uint64_t k = 0; k < n_shots; ++k) {
        int64_t i = 0;
        while (t.nodes[static_cast<size_t>(i)].m_qubit >= 0) {
            const ShotNode& nd = t.nodes[static_cast<size_t>(i)];
            auto flip = std::discrete_distribution<int> {{nd.w[0], nd.w[1]}};
            i = t.child(i, flip(rng));
        }
        leaf[k] = i;
        ++t.nodes[static_cast<size_t>(i)].shots;
    }
    // one sampler per shot over its final state (ProbabilitySampler_ seeded with `seed`, or from
    // std::random_device): per branch, the shots that end there
    const uint64_t N = uint64_t {1} << init->n;
    void* d = device_alloc(N * sizeof(double));
    try {
        std::vector<uint64_t> idx;
        std::vector<uint64_t> next(t.nodes.size(), 0);
        std::vector<std::vector<uint64_t>> drawn(t.nodes.size());
        for (size_t i = 0; i < t.nodes.size(); ++i) {
            const uint64_t m = t.nodes[i].shots;
            if (m == 0) {
                continue;
            }
            kq_state* st = t.nodes[i].state ? nullptr : t.materialise(static_cast<int64_t>(i));
            const kq_state* src = st ? st : t.nodes[i].state;
            try {
                launch_probabilities(const_cast<kq_state*>(src), static_cast<double*>(d), noise);
                KQ_HIP_CHECK(hipStreamSynchronize(src->stream));
                // a seeded sampler draws the same first number for every shot of a branch
                const uint64_t draws = has_seed ? 1 : m;
                idx.assign(draws, 0);
                sample_from_device_probs(src->stream, static_cast<double*>(d), N, draws, has_seed, seed, idx.data());
            }
            catch (...) {
                if (st) {
                    (void) kq_state_destroy(st);
                }
                throw;
            }
            if (st) {
                (void) kq_state_destroy(st);
            }
            drawn[i] = idx;
        }
        for (uint64_t k = 0; k < n_shots; ++k) {
            const auto i = static_cast<size_t>(leaf[k]);
            out[k] = has_seed ? drawn[i][0] : drawn[i][next[i]++];
        }
    }
    catch (...) {
        (void) hipFree(d);
        throw;
    }
    KQ_HIP_CHECK(hipFree(d));
    if (n_branches) {
        *n_branches = t.nodes.size();
    }
}



}  // namespace kq

using namespace kq;

extern "C" {

int kq_simulate(kq_state* s, uint32_t circuit_n_qubits, const kq_op* ops, size_t n_ops, const int32_t* preds,
                size_t n_preds, int has_seed, int32_t seed, int8_t* creg, size_t n_bits, kq_logger_fn logger,
                void* logger_user)
{
    return guarded([&] {
        check_state_lazy(s);
        if (n_ops > 0 && ops == nullptr) {
            throw InvalidArg {"null op array"};
        }
        ensure_init();
        last_stats() = PlanStats {};
        simulate(s, circuit_n_qubits, ops, n_ops, preds, n_preds, has_seed != 0, seed, creg, n_bits, logger, logger_user);
    });
}

int kq_qubit_weights(kq_state* s, uint32_t qubit, double out[2])
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (qubit >= s->n) {
            throw InvalidArg {"qubit out of range"};
        }
        ensure_init();
        launch_qubit_weights(s, qubit, out);
    });
}

int kq_collapse(kq_state* s, uint32_t qubit, int kept_bit, double scale)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (qubit >= s->n || (kept_bit != 0 && kept_bit != 1)) {
            throw InvalidArg {"collapse: bad qubit or bit"};
        }
        ensure_init();
        launch_collapse(s, qubit, kept_bit, scale);
        KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int kq_measure(kq_state* s, uint32_t qubit, int has_seed, int32_t seed, int* outcome)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        ensure_init();
        const int o = measure(s, qubit, has_seed != 0, seed);
        if (outcome) {
            *outcome = o;
        }
    });
}

int kq_probabilities(kq_state* s, const double* noise, double* out)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (!out) {
            throw InvalidArg {"null output"};
        }
        check_noise(noise, s->n);
        ensure_init();
        const uint64_t N = uint64_t {1} << s->n;
        void* d = device_alloc(N * sizeof(double));
        try {
            launch_probabilities(s, static_cast<double*>(d), noise);
            KQ_HIP_CHECK(hipMemcpyAsync(out, d, N * sizeof(double), hipMemcpyDeviceToHost, s->stream));
            KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
        }
        catch (...) {
            (void) hipFree(d);
            throw;
        }
        KQ_HIP_CHECK(hipFree(d));
    });
}

int kq_sample(kq_state* s, const double* noise, uint64_t n_shots, int has_seed, int32_t seed, uint64_t* out_indices)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (n_shots > 0 && !out_indices) {
            throw InvalidArg {"null output"};
        }
        check_noise(noise, s->n);
        ensure_init();
        const uint64_t N = uint64_t {1} << s->n;
        void* d = device_alloc(N * sizeof(double));
        try {
            launch_probabilities(s, static_cast<double*>(d), noise);
            sample_from_device_probs(s->stream, static_cast<double*>(d), N, n_shots, has_seed != 0, seed, out_indices);
        }
        catch (...) {
            (void) hipFree(d);
            throw;
        }
        KQ_HIP_CHECK(hipFree(d));
    });
}

int kq_sample_circuit_shots(const kq_state* initial, uint32_t circuit_n_qubits, const kq_op* ops, size_t n_ops,
                            const int32_t* preds, size_t n_preds, size_t n_bits, const double* noise, uint64_t n_shots,
                            int has_seed, int32_t seed, uint64_t max_cached_bytes, uint64_t* out_indices,
                            uint64_t* n_branches)
{
    return guarded([&] {
        check_state(initial);
        if ((n_shots > 0 && !out_indices) || (!ops && n_ops) || (!preds && n_preds)) {
            throw InvalidArg {"null argument"};
        }
        ensure_init();
        normalise_layout(initial);
        sample_circuit_shots(initial, circuit_n_qubits, ops, n_ops, preds, n_preds, n_bits, noise, n_shots, has_seed != 0,
                             seed, max_cached_bytes, out_indices, n_branches);
    });
}

int kq_sample_probabilities(const double* probs, uint64_t n_probs, uint64_t n_shots, int has_seed, int32_t seed,
                            uint64_t* out_indices)
{
    return guarded([&] {
        if (!probs || n_probs == 0) {
            throw InvalidArg {"empty probability array"};
        }
        if (n_shots > 0 && !out_indices) {
            throw InvalidArg {"null output"};
        }
        ensure_init();
        hipStream_t st;
        KQ_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        void* d = nullptr;
        try {
            d = device_alloc(n_probs * sizeof(double));
            KQ_HIP_CHECK(hipMemcpyAsync(d, probs, n_probs * sizeof(double), hipMemcpyHostToDevice, st));
            sample_from_device_probs(st, static_cast<double*>(d), n_probs, n_shots, has_seed != 0, seed, out_indices);
        }
        catch (...) {
            if (d) {
                (void) hipFree(d);
            }
            (void) hipStreamDestroy(st);
            throw;
        }
        KQ_HIP_CHECK(hipFree(d));
        KQ_HIP_CHECK(hipStreamDestroy(st));
    });
}

int kq_sample_probabilities_noalloc(const double* probs, uint64_t n_probs, uint64_t n_shots, int has_seed, int32_t seed,
                                    uint64_t* out_indices)
{
    return guarded([&] {
        if (!probs || n_probs == 0) {
            throw InvalidArg {"empty probability array"};
        }
        if (n_shots > 0 && !out_indices) {
            throw InvalidArg {"null output"};
        }
        if (n_probs > NOALLOC_MAX_PROBS || n_shots > NOALLOC_MAX_SHOTS) {
            throw InvalidArg {"allocation-free sampling takes at most " + std::to_string(NOALLOC_MAX_PROBS)
                              + " probabilities and " + std::to_string(NOALLOC_MAX_SHOTS) + " shots (got "
                              + std::to_string(n_probs) + " and " + std::to_string(n_shots) + ")"};
        }
        ensure_init();
        const NoallocWorkspace ws = noalloc_workspace();
        auto* base = static_cast<char*>(ws.ptr);
        auto* d_p = reinterpret_cast<double*>(base);
        KQ_HIP_CHECK(hipMemcpyAsync(d_p, probs, n_probs * sizeof(double), hipMemcpyHostToDevice, ws.stream));
        sample_with_workspace(ws.stream, d_p, n_probs, n_shots, has_seed != 0, seed, out_indices,
                              base + NOALLOC_MAX_PROBS * sizeof(double), NOALLOC_SCRATCH_BYTES - NOALLOC_MAX_PROBS * sizeof(double));
    });
}

}  // extern "C"
