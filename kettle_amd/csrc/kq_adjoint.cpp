// The adjoint energy gradient: forward sweep with checkpoints, lambda = H psi, and the reverse sweep
// through fused adjoint tile passes with one derivative reduction per parameterised gate.
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

// ---- adjoint gradient -------------------------------------------------------------------------------

GateRec inverse_of(const GateRec& g)
{
    GateRec r = g;
    if (g.kind == GK_SWAP) {
        return r;
    }
    if (g.kind == GK_DIAG || g.kind == GK_DIAG1) {
        r.m[1] = -g.m[1];
        r.m[7] = -g.m[7];
        return r;
    }
    // conjugate transpose
    r.m[0] = g.m[0]; r.m[1] = -g.m[1];
    r.m[2] = g.m[4]; r.m[3] = -g.m[5];
    r.m[4] = g.m[2]; r.m[5] = -g.m[3];
    r.m[6] = g.m[6]; r.m[7] = -g.m[7];
    return r;
}

// The generator Q of an angle gate's derivative, dG/dtheta = Q G, for kettle's angle-gate matrices (as
// written in gate_matrix of kq_plan.cpp): RX -> (-i/2) X, RY -> (-i/2) Y, RZ -> (-i/2) Z, P -> i |1><1|, and
// the same Q on the target of the controlled forms (their derivative vanishes where the control is 0).
// Q does not depend on the angle.  The adjoint sweep uses dE/dtheta = 2 Re <lambda| Q |psi> with psi and
// lambda as they are before the gate is undone.
bool derivative_of(const kq_op& op, const GateRec& g, GateRec& d)
{
    d = g;
    d.kind = GK_DENSE;
    std::fill(std::begin(d.m), std::end(d.m), 0.0);
    switch (op.gate) {
        case KQ_RX: case KQ_CRX:
            d.m[3] = -0.5; d.m[5] = -0.5;
            return true;
        case KQ_RY: case KQ_CRY:
            d.m[2] = -0.5; d.m[4] = 0.5;
            return true;
        case KQ_RZ: case KQ_CRZ:
            d.kind = GK_DIAG;
            d.m[1] = -0.5; d.m[7] = 0.5;
            return true;
        case KQ_P: case KQ_CP:
            d.kind = GK_DIAG;
            d.m[7] = 1.0;
            return true;
        default: break;
    }
    return false;
}

namespace
{

size_t align256b(size_t x) { return (x + 255) & ~size_t {255}; }

// Bytes of one lowered adjoint pass in the program blob: gates, steps, layouts, derivative records,
// merged-run terms, each 256-byte aligned.
size_t program_bytes(const TileProgram& tp)
{
    return align256b(tp.gates.size() * sizeof(TileGate)) + align256b(tp.steps.size() * sizeof(TileStep))
         + align256b(tp.layouts.size() * sizeof(TileLayout)) + align256b(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate))
         + align256b(std::max<size_t>(tp.terms.size(), 1) * sizeof(TileTerm));
}

void write_program(unsigned char* dst, const TileProgram& tp)
{
    size_t o = 0;
    std::memcpy(dst + o, tp.gates.data(), tp.gates.size() * sizeof(TileGate));
    o += align256b(tp.gates.size() * sizeof(TileGate));
    std::memcpy(dst + o, tp.steps.data(), tp.steps.size() * sizeof(TileStep));
    o += align256b(tp.steps.size() * sizeof(TileStep));
    std::memcpy(dst + o, tp.layouts.data(), tp.layouts.size() * sizeof(TileLayout));
    o += align256b(tp.layouts.size() * sizeof(TileLayout));
    std::memcpy(dst + o, tp.derivs.data(), tp.derivs.size() * sizeof(TileGate));
    o += align256b(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate));
    std::memcpy(dst + o, tp.terms.data(), tp.terms.size() * sizeof(TileTerm));
}

// one planned pass of the reverse sweep
struct ReversePass
{
    TileProgram tp;
    size_t blob_off = 0;      // its program (interpreted) or coefficients (compiled) in the blob
    uint64_t part_off = 0;    // its first derivative partial
    hipFunction_t fn = nullptr;   // the compiled pass kernel, or null: the interpreted tile kernel runs it
    AdjointCoefs coefs;           // its coefficient buffer and derivative factors (compiled kernels)
};

// Compiled reverse-pass kernels by pass structure (adjoint_structure_key): an optimiser's next call on the
// same circuit structure reuses them without generating their sources again.
class ReverseKernels
{
public:
    const hipFunction_t* find(const std::vector<uint64_t>& key)
    {
        std::lock_guard<std::mutex> lk(mu_);
        const auto it = map_.find(hash(key));
        return (it != map_.end() && it->second.first == key) ? &it->second.second : nullptr;
    }
    void add(std::vector<uint64_t>&& key, hipFunction_t fn)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (map_.size() >= 4096) {
            map_.clear();
        }
        const uint64_t h = hash(key);
        map_[h] = {std::move(key), fn};
    }

private:
    static uint64_t hash(const std::vector<uint64_t>& key)
    {
        uint64_t h = 1469598103934665603ull;
        for (const uint64_t v : key) {
            h = (h ^ v) * 1099511628211ull;
        }
        return h;
    }
    std::mutex mu_;
    std::unordered_map<uint64_t, std::pair<std::vector<uint64_t>, hipFunction_t>> map_;
};

ReverseKernels& reverse_kernels()
{
    static ReverseKernels c;
    return c;
}

// Gate-by-gate adjoint for states too small for a tile (fewer qubits than one thread's register block).
void energy_gradient_unfused(kq_state* s, const kq_op* ops, const std::vector<GateRec>& recs,
                             const std::vector<PauliGroup>& groups, const std::vector<PauliZ>& strings, size_t n_params,
                             double* energy, double* grad)
{
    const size_t G = recs.size();
    const size_t bytes = s->n_amps * amp_bytes(s->precision);
    const size_t acc_bytes = align256b(std::max<size_t>(G, 1) * 2 * sizeof(double));
    auto* ws = static_cast<unsigned char*>(grow(s->adj, s->adj_bytes, 2 * align256b(bytes) + acc_bytes));
    void* lam = ws;
    void* psi = ws + align256b(bytes);
    auto* accs = reinterpret_cast<double*>(ws + 2 * align256b(bytes));
    run_records(s, recs, false);
    launch_pauli_apply_sum(s, s->amps, lam, groups, strings);
    double e[2];
    launch_dot(s, s->amps, lam, e);
    *energy = e[0];
    launch_copy(s, psi, s->amps, bytes);
    std::fill(grad, grad + n_params, 0.0);
    std::vector<char> has_d(G, 0);
    std::vector<GateRec> derivs(G);
    for (size_t gi = G; gi-- > 0;) {
        has_d[gi] = ops[gi].aux >= 0 && derivative_of(ops[gi], recs[gi], derivs[gi]) ? 1 : 0;
        launch_adjoint_gate(s, psi, lam, inverse_of(recs[gi]), has_d[gi] ? &derivs[gi] : nullptr, has_d[gi] ? accs + 2 * gi : nullptr);
    }
    std::vector<double> h(2 * G);
    if (G) {
        KQ_HIP_CHECK(hipMemcpyAsync(h.data(), accs, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    }
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    for (size_t gi = 0; gi < G; ++gi) {
        if (has_d[gi]) {
            grad[ops[gi].aux] += 2.0 * h[2 * gi];
        }
    }
}

}  // namespace

// Adjoint energy gradient through fused passes (Jones & Gacon, arXiv:2009.02823).
//   - The forward sweep is one whole-circuit plan of tile passes; checkpoints of psi sit at pass
//     boundaries, every ceil(sqrt(P)) passes of the P (or, with checkpoint_every > 0, at the first pass
//     boundary past every checkpoint_every gates).
//   - lambda = H psi and E = <psi|lambda> (FP64).
//   - The reverse sweep runs segment by segment between checkpoints: each segment's gates, inverted and
//     in reverse execution order, are planned as adjoint tile passes that update psi and lambda in one
//     read and write of each and reduce Re <lambda| dG |psi> for every parameterised gate into per-wave
//     FP64 partials.  At each checkpoint the uncomputed psi is compared with the stored one on the
//     device and replaced by it when the max-abs drift exceeds drift_tol.
//   - Every buffer (lambda, the sweep's psi, checkpoints, partials, programs) lives in the state's
//     adjoint workspace, grown when a call needs more and kept; all reverse programs are planned on
//     the host and uploaded in one copy; the derivative rows are summed on the device in a fixed order
//     and read back once, so a call waits for the device three times (after the forward sweep, for the
//     energy, at the end).
void energy_gradient(kq_state* s, const kq_op* ops, size_t n_ops, size_t n_strings, const uint32_t* off,
                     const uint32_t* q, const uint8_t* k, const uint8_t* phases, const double* coeffs, size_t n_params,
                     double* energy, double* grad, uint32_t every, double drift_tol, size_t* n_restores)
{
    if (s->is_dm) {
        throw InvalidArg {"energy gradient: statevector states only"};
    }
    std::vector<GateRec> recs;
    recs.reserve(n_ops);
    for (size_t i = 0; i < n_ops; ++i) {
        if (ops[i].gate < 0 || ops[i].gate >= KQ_M) {
            throw InvalidArg {"energy gradient: op " + std::to_string(i) + " is not a unitary gate"};
        }
        if (ops[i].aux >= 0 && static_cast<size_t>(ops[i].aux) >= n_params) {
            throw InvalidArg {"energy gradient: parameter index out of range"};
        }
        compile_gate(ops[i], s->n, false, recs);
    }
    if (n_restores) {
        *n_restores = 0;
    }
    const size_t G = recs.size();
    const int max_k = fusion_max_k(s->precision, s->nvec);
    const int rq = register_qubits(s->precision);
    if (max_k < rq) {
        materialise_basis(s);   // the unfused sweep reads the state directly
        s->zero_mask = 0;
        std::vector<PauliGroup> groups;
        std::vector<PauliZ> strings;
        build_pauli(s->n, n_strings, off, q, k, phases, coeffs, groups, strings);
        sort_pauli_groups_by_tile(groups);
        energy_gradient_unfused(s, ops, recs, groups, strings, n_params, energy, grad);
        return;
    }
    const size_t bytes = s->n_amps * amp_bytes(s->precision);
    const size_t vec_bytes = align256b(bytes);

    // ---- forward plan (tile passes only: each has a mirror), checkpoint boundaries (after pass b) -------
    std::vector<uint8_t> bit_of;
    const std::vector<TilePass> fwd = plan_statevector_passes(s, recs, false, bit_of);
    const size_t P = fwd.size();
    // the operator acts on the state as the forward passes leave it: qubit q on index bit bit_of[q]
    if (n_strings > 0 && off == nullptr) {
        throw InvalidArg {"null Pauli operator arrays"};
    }
    std::vector<uint32_t> q_at(n_strings ? off[n_strings] : 0);
    if (!q_at.empty() && q == nullptr) {
        throw InvalidArg {"null Pauli operator arrays"};
    }
    for (size_t t = 0; t < q_at.size(); ++t) {
        if (q[t] >= s->n) {
            throw InvalidArg {"energy gradient: Pauli term qubit out of range"};
        }
        q_at[t] = bit_of[q[t]];
    }
    std::vector<PauliGroup> groups;
    std::vector<PauliZ> strings;
    build_pauli(s->n, n_strings, off, q_at.data(), k, phases, coeffs, groups, strings);
    sort_pauli_groups_by_tile(groups);
    std::vector<size_t> bounds;
    if (every == 0) {
        const size_t stride = std::max<size_t>(1, static_cast<size_t>(std::ceil(std::sqrt(static_cast<double>(P)))));
        for (size_t b = stride; b < P; b += stride) {
            bounds.push_back(b - 1);
        }
    }
    else {
        size_t done = 0, next = every;
        for (size_t p = 0; p + 1 < P; ++p) {
            done += fwd[p].src.size();
            if (done >= next) {
                bounds.push_back(p);
                while (next <= done) {
                    next += every;
                }
            }
        }
    }
    const size_t cap = std::max<size_t>(1, (size_t {32} << 30) / std::max<size_t>(vec_bytes, 1));
    if (bounds.size() > cap) {
        std::vector<size_t> thin;
        for (size_t i = 0; i < cap; ++i) {
            thin.push_back(bounds[(i * bounds.size()) / cap]);
        }
        bounds = thin;
    }
    const size_t C = bounds.size();
    std::vector<size_t> seg_first {0};   // first pass of each segment
    for (const size_t b : bounds) {
        seg_first.push_back(b + 1);
    }
    seg_first.push_back(P);

    // ---- workspace of the vectors: [lambda][psi][checkpoints][drift partials][drift][energy] ---------
    size_t o = 0;
    const size_t o_lam = o;
    o += vec_bytes;
    const size_t o_psi = o;
    o += vec_bytes;
    const size_t o_ckpt = o;
    o += C * vec_bytes;
    const size_t o_red = o;
    o += align256b(REDUCTION_BLOCKS * sizeof(double2));
    const size_t o_drift = o;
    o += align256b(std::max<size_t>(C, 1) * sizeof(double));
    const size_t o_energy = o;
    o += 256;
    auto* ws = static_cast<unsigned char*>(grow(s->adj, s->adj_bytes, o));
    void* lam = ws + o_lam;
    void* psi = ws + o_psi;
    const auto ckpt = [&](size_t i) { return static_cast<void*>(ws + o_ckpt + i * vec_bytes); };
    auto* d_red = reinterpret_cast<double2*>(ws + o_red);
    auto* d_drift = reinterpret_cast<double*>(ws + o_drift);
    auto* d_energy = reinterpret_cast<double2*>(ws + o_energy);

    // ---- forward sweep with checkpoint copies, lambda = H psi, E = <psi|lambda>: enqueued only, so the
    //      host plans the reverse sweep while the device runs them -------------------------------------
    std::vector<std::pair<size_t, std::function<void()>>> after;
    for (size_t i = 0; i < C; ++i) {
        after.emplace_back(bounds[i], [&, i] { launch_copy(s, ckpt(i), s->amps, bytes); });
    }
    execute_passes_with(s, fwd, after, false);
    last_stats().n_gates += G;
    {
        bool identity = true;
        for (size_t b = 0; b < bit_of.size(); ++b) {
            identity = identity && bit_of[b] == b;
        }
        s->bit_of = identity ? std::vector<uint8_t> {} : bit_of;
    }
    profiled(s, 4, 2.0 * static_cast<double>(amp_bytes(s->precision) * s->n_amps), strings.size(),
             [&] { launch_pauli_apply_sum(s, s->amps, lam, groups, strings); });
    launch_dot_to_device(s, s->amps, lam, d_red, d_energy);
    launch_copy(s, psi, s->amps, bytes);   // the reverse sweep runs on a copy: the state keeps psi(theta)

    // ---- reverse plan: the mirror of every forward pass, segment by segment ---------------------------
    std::vector<GateRec> derivs(G);
    std::vector<const GateRec*> deriv_of_rec(G, nullptr);
    for (size_t gi = 0; gi < G; ++gi) {
        if (ops[gi].aux >= 0 && derivative_of(ops[gi], recs[gi], derivs[gi])) {
            deriv_of_rec[gi] = &derivs[gi];
        }
    }
    std::vector<std::vector<ReversePass>> rev(C + 1);
    std::vector<uint32_t> row_gate;   // record index of each derivative row
    std::vector<uint64_t> row_off;
    std::vector<uint32_t> row_cnt;
    size_t blob_bytes = 0;
    uint64_t part_total = 0;
    for (size_t j = 0; j <= C; ++j) {
        for (size_t p = seg_first[j + 1]; p-- > seg_first[j];) {
            ReversePass rp;
            lower_adjoint_mirror(fwd[p], s->precision, s->nvec, deriv_of_rec, rp.tp);
            rp.part_off = part_total;
            const uint64_t n_part = uint64_t {tile_pass_blocks(rp.tp)} * tile_pass_waves(rp.tp);
            for (size_t d = 0; d < rp.tp.derivs.size(); ++d) {
                row_off.push_back(part_total + d * n_part);
                row_cnt.push_back(static_cast<uint32_t>(n_part));
                row_gate.push_back(rp.tp.deriv_src[d]);
            }
            part_total += rp.tp.derivs.size() * n_part;
            rev[j].push_back(std::move(rp));
        }
    }
    const size_t R = row_gate.size();
    // reverse passes as run-time compiled kernels where the state is large enough for them; the
    // programs of the others (and of any that fail to compile) go to the interpreted tile kernel.  A pass
    // whose structure was seen before takes its kernel from the cache and only its coefficients anew.
    if (jit_wanted(s) && jit_available()) {
        std::vector<std::string> srcs;
        std::vector<ReversePass*> who;
        std::vector<std::vector<uint64_t>> keys;
        ReverseKernels& cache = reverse_kernels();
        for (auto& seg : rev) {
            for (ReversePass& rp : seg) {
                adjoint_coefs(rp.tp, s->precision, rp.coefs);
                std::vector<uint64_t> key = adjoint_structure_key(rp.tp, s->precision, rp.coefs);
                if (const hipFunction_t* fn = cache.find(key)) {
                    rp.fn = *fn;
                    continue;
                }
                srcs.push_back(jit_adjoint_source(rp.tp, s->precision, rp.coefs));
                who.push_back(&rp);
                keys.push_back(std::move(key));
            }
        }
        const auto fns = jit_kernels(srcs);
        for (size_t i = 0; i < who.size(); ++i) {
            who[i]->fn = fns[i];
            if (fns[i]) {
                cache.add(std::move(keys[i]), fns[i]);
            }
        }
    }
    // each derivative row's factor, in row order (compiled passes divide the vectors by their pivots)
    std::vector<double> row_scale;
    row_scale.reserve(R);
    for (auto& seg : rev) {
        for (ReversePass& rp : seg) {
            rp.blob_off = blob_bytes;
            blob_bytes += rp.fn ? align256b(std::max<size_t>(rp.coefs.coefs.size(), 1) * sizeof(double)) : program_bytes(rp.tp);
            for (size_t d = 0; d < rp.tp.derivs.size(); ++d) {
                row_scale.push_back(rp.fn ? rp.coefs.dscale[d] : 1.0);
            }
        }
    }

    // ---- workspace of the reverse sweep: [partials][row sums][uploaded: row tables, restore counter,
    //      programs] -----------------------------------------------------------------------------------
    size_t ob = 0;
    const size_t o_part = ob;
    ob += align256b(std::max<uint64_t>(part_total, 1) * sizeof(double));
    const size_t o_sums = ob;
    ob += align256b(std::max<size_t>(R, 1) * sizeof(double));
    const size_t o_up = ob;
    const size_t u_off = 0;
    const size_t u_cnt = u_off + align256b(std::max<size_t>(R, 1) * sizeof(uint64_t));
    const size_t u_count = u_cnt + align256b(std::max<size_t>(R, 1) * sizeof(uint32_t));
    const size_t u_blob = u_count + 256;
    const size_t up_bytes = u_blob + std::max<size_t>(blob_bytes, 256);
    ob += up_bytes;
    auto* wb = static_cast<unsigned char*>(grow(s->aux, s->aux_bytes, ob));
    std::vector<unsigned char> up(up_bytes, 0);   // restore counter starts at 0
    std::memcpy(up.data() + u_off, row_off.data(), R * sizeof(uint64_t));
    std::memcpy(up.data() + u_cnt, row_cnt.data(), R * sizeof(uint32_t));
    for (const auto& seg : rev) {
        for (const ReversePass& rp : seg) {
            if (rp.fn) {
                std::memcpy(up.data() + u_blob + rp.blob_off, rp.coefs.coefs.data(), rp.coefs.coefs.size() * sizeof(double));
            }
            else {
                write_program(up.data() + u_blob + rp.blob_off, rp.tp);
            }
        }
    }
    auto* d_part = reinterpret_cast<double*>(wb + o_part);
    auto* d_sums = reinterpret_cast<double*>(wb + o_sums);
    unsigned char* d_up = wb + o_up;
    auto* d_count = reinterpret_cast<uint32_t*>(d_up + u_count);
    KQ_HIP_CHECK(hipMemcpyAsync(d_up, up.data(), up_bytes, hipMemcpyHostToDevice, s->stream));

    // ---- reverse sweep --------------------------------------------------------------------------------
    const double ab = static_cast<double>(amp_bytes(s->precision));
    for (size_t j = C + 1; j-- > 0;) {
        for (const ReversePass& rp : rev[j]) {
            const TileProgram& tp = rp.tp;
            if (rp.fn) {
                profiled(s, 3, 4.0 * ab * static_cast<double>(s->n_amps), tp.gates.size(), [&] {
                    launch_jit_adjoint_pass(s, rp.fn, tp, psi, lam, reinterpret_cast<const double*>(d_up + u_blob + rp.blob_off),
                                            d_part + rp.part_off);
                });
                continue;
            }
            const unsigned char* g = d_up + u_blob + rp.blob_off;
            const unsigned char* sp = g + align256b(tp.gates.size() * sizeof(TileGate));
            const unsigned char* ly = sp + align256b(tp.steps.size() * sizeof(TileStep));
            const unsigned char* dv = ly + align256b(tp.layouts.size() * sizeof(TileLayout));
            const unsigned char* tm = dv + align256b(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate));
            profiled(s, 3, 4.0 * ab * static_cast<double>(s->n_amps), tp.gates.size(),
                     [&] { launch_adjoint_tile_pass(s, psi, lam, tp, g, sp, ly, dv, tm, d_part + rp.part_off); });
        }
        // psi is now the forward state at the segment's start: the checkpoint taken there
        if (j >= 1) {
            launch_drift_restore(s, psi, ckpt(j - 1), drift_tol, d_red, d_drift + (j - 1), d_count);
        }
    }
    std::vector<double> sums(R);
    uint32_t count = 0;
    double2 e {0.0, 0.0};
    if (R) {
        launch_sum_rows(s->stream, d_part, reinterpret_cast<const uint64_t*>(d_up + u_off),
                        reinterpret_cast<const uint32_t*>(d_up + u_cnt), static_cast<uint32_t>(R), d_sums);
        KQ_HIP_CHECK(hipMemcpyAsync(sums.data(), d_sums, R * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    }
    KQ_HIP_CHECK(hipMemcpyAsync(&count, d_count, sizeof count, hipMemcpyDeviceToHost, s->stream));
    KQ_HIP_CHECK(hipMemcpyAsync(&e, d_energy, sizeof e, hipMemcpyDeviceToHost, s->stream));
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    profiler_collect();
    *energy = e.x;
    std::fill(grad, grad + n_params, 0.0);
    for (size_t r = 0; r < R; ++r) {
        grad[ops[row_gate[r]].aux] += 2.0 * sums[r] * row_scale[r];
    }
    if (n_restores) {
        *n_restores = count;
    }
}

}  // namespace kq

using namespace kq;

extern "C" {

int kq_energy_gradient(kq_state* state, const kq_op* ops, size_t n_ops, size_t n_strings, const uint32_t* term_offsets,
                       const uint32_t* term_qubits, const uint8_t* term_kinds, const uint8_t* phases, const double* coeffs,
                       size_t n_params, double* energy, double* grad, uint32_t checkpoint_every, double drift_tol,
                       size_t* n_restores)
{
    return guarded([&] {
        // a pending basis state is built by the forward sweep's first pass (which reads nothing)
        check_state_lazy(state);
        if (!energy || (n_params > 0 && !grad)) {
            throw InvalidArg {"null output"};
        }
        ensure_init();
        if (!state->bit_of.empty()) {
            normalise_layout(state);
        }
        last_stats() = PlanStats {};
        energy_gradient(state, ops, n_ops, n_strings, term_offsets, term_qubits, term_kinds, phases, coeffs, n_params,
                        energy, grad, checkpoint_every, drift_tol, n_restores);
        state->zero_mask = 0;
    });
}

}  // extern "C"
