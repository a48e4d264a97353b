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

// d/dtheta of kettle's angle-gate matrices (as written in gate_matrix of kq_plan.cpp)
bool derivative_of(const kq_op& op, const GateRec& g, GateRec& d)
{
    const double th = op.angle;
    d = g;
    d.kind = GK_DENSE;
    std::fill(std::begin(d.m), std::end(d.m), 0.0);
    switch (op.gate) {
        case KQ_RX: case KQ_CRX: {
            const double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
            d.m[0] = -s / 2; d.m[3] = -c / 2; d.m[5] = -c / 2; d.m[6] = -s / 2;
            return true;
        }
        case KQ_RY: case KQ_CRY: {
            const double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
            d.m[0] = -s / 2; d.m[2] = -c / 2; d.m[4] = c / 2; d.m[6] = -s / 2;
            return true;
        }
        case KQ_RZ: case KQ_CRZ: {
            const double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
            d.m[0] = -s / 2; d.m[1] = -c / 2; d.m[6] = -s / 2; d.m[7] = c / 2;
            return true;
        }
        case KQ_P: case KQ_CP: {
            const double c = std::cos(th), s = std::sin(th);
            d.m[6] = -s; d.m[7] = c;
            return true;
        }
        default: break;
    }
    return false;
}

namespace
{

// One segment of the reverse sweep through fused adjoint tile passes.  inv[i] is the inverse of the
// i-th gate in reverse order and deriv[i] its derivative (or null); values[i] receives
// Re <lambda| dG |psi> for each i with a derivative.
void run_adjoint_segment(kq_state* s, void* psi, void* lam, const std::vector<GateRec>& inv,
                         const std::vector<const GateRec*>& deriv, std::vector<double>& values)
{
    const int max_k = fusion_max_k(s->precision, s->nvec);
    const auto passes = plan_passes(inv, s->nvec, max_k, fusion_low_qubits(), fusion_max_gates(), register_qubits(s->precision),
                                    fusion_max_stages(), false);
    const auto align = [](size_t x) { return (x + 255) & ~size_t {255}; };
    std::vector<TileProgram> progs(passes.size());
    std::vector<size_t> offs(passes.size());
    size_t total = 0;
    uint64_t n_rows_total = 0;
    std::vector<uint64_t> row_off;
    std::vector<uint32_t> row_cnt;
    std::vector<size_t> part_off(passes.size());
    uint64_t part_total = 0;
    for (size_t i = 0; i < passes.size(); ++i) {
        lower_tile_pass(passes[i], s->precision, s->nvec, progs[i], &deriv);
        const TileProgram& tp = progs[i];
        offs[i] = total;
        total += align(tp.gates.size() * sizeof(TileGate)) + align(tp.steps.size() * sizeof(TileStep))
               + align(tp.layouts.size() * sizeof(TileLayout)) + align(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate))
               + align(std::max<size_t>(tp.terms.size(), 1) * sizeof(TileTerm));
        const uint64_t n_part = uint64_t {tile_pass_blocks(tp)} * tile_pass_waves(tp);
        part_off[i] = part_total;
        for (size_t k = 0; k < tp.derivs.size(); ++k) {
            row_off.push_back(part_total + k * n_part);
            row_cnt.push_back(static_cast<uint32_t>(n_part));
        }
        part_total += tp.derivs.size() * n_part;
        n_rows_total += tp.derivs.size();
    }
    const size_t tables = align(row_off.size() * sizeof(uint64_t)) + align(row_cnt.size() * sizeof(uint32_t));
    std::vector<unsigned char> blob(std::max<size_t>(total + tables, 256), 0);
    for (size_t i = 0; i < passes.size(); ++i) {
        const TileProgram& tp = progs[i];
        size_t o = offs[i];
        std::memcpy(blob.data() + o, tp.gates.data(), tp.gates.size() * sizeof(TileGate));
        o += align(tp.gates.size() * sizeof(TileGate));
        std::memcpy(blob.data() + o, tp.steps.data(), tp.steps.size() * sizeof(TileStep));
        o += align(tp.steps.size() * sizeof(TileStep));
        std::memcpy(blob.data() + o, tp.layouts.data(), tp.layouts.size() * sizeof(TileLayout));
        o += align(tp.layouts.size() * sizeof(TileLayout));
        std::memcpy(blob.data() + o, tp.derivs.data(), tp.derivs.size() * sizeof(TileGate));
        o += align(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate));
        std::memcpy(blob.data() + o, tp.terms.data(), tp.terms.size() * sizeof(TileTerm));
    }
    std::memcpy(blob.data() + total, row_off.data(), row_off.size() * sizeof(uint64_t));
    std::memcpy(blob.data() + total + align(row_off.size() * sizeof(uint64_t)), row_cnt.data(),
                row_cnt.size() * sizeof(uint32_t));
    auto* d = static_cast<unsigned char*>(grow(s->prog, s->prog_bytes, blob.size()));
    KQ_HIP_CHECK(hipMemcpyAsync(d, blob.data(), blob.size(), hipMemcpyHostToDevice, s->stream));
    // partials, then one sum per derivative slot
    const size_t part_bytes = align(std::max<uint64_t>(part_total, 1) * sizeof(double));
    auto* aux = static_cast<unsigned char*>(grow(s->aux, s->aux_bytes, part_bytes + std::max<uint64_t>(n_rows_total, 1) * sizeof(double)));
    auto* d_part = reinterpret_cast<double*>(aux);
    auto* d_sums = reinterpret_cast<double*>(aux + part_bytes);
    const double ab = static_cast<double>(amp_bytes(s->precision));
    for (size_t i = 0; i < passes.size(); ++i) {
        const TileProgram& tp = progs[i];
        const unsigned char* g = d + offs[i];
        const unsigned char* sp = g + align(tp.gates.size() * sizeof(TileGate));
        const unsigned char* ly = sp + align(tp.steps.size() * sizeof(TileStep));
        const unsigned char* dv = ly + align(tp.layouts.size() * sizeof(TileLayout));
        const unsigned char* tm = dv + align(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate));
        profiled(s, 1, 4.0 * ab * static_cast<double>(s->n_amps), tp.gates.size(),
                 [&] { launch_adjoint_tile_pass(s, psi, lam, tp, g, sp, ly, dv, tm, d_part + part_off[i]); });
    }
    std::vector<double> sums(n_rows_total);
    if (n_rows_total) {
        launch_sum_rows(s->stream, d_part, reinterpret_cast<const uint64_t*>(d + total),
                        reinterpret_cast<const uint32_t*>(d + total + align(row_off.size() * sizeof(uint64_t))),
                        static_cast<uint32_t>(n_rows_total), d_sums);
        KQ_HIP_CHECK(hipMemcpyAsync(sums.data(), d_sums, n_rows_total * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    }
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    profiler_collect();
    size_t row = 0;
    for (const TileProgram& tp : progs) {
        for (const uint32_t src : tp.deriv_src) {
            values[src] = sums[row++];
        }
    }
}

}  // namespace

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
    std::vector<PauliGroup> groups;
    std::vector<PauliZ> strings;
    build_pauli(s->n, n_strings, off, q, k, phases, coeffs, groups, strings);

    const size_t G = recs.size();
    const size_t bytes = s->n_amps * amp_bytes(s->precision);
    if (every == 0) {
        every = static_cast<uint32_t>(std::max<double>(1.0, std::ceil(std::sqrt(static_cast<double>(G)))));
    }
    // checkpoints at gate boundaries b = every, 2 every, ... (< G), capped at 32 GiB of memory; the
    // forward and the reverse sweep are planned segment by segment between them, so the state at
    // each boundary is exactly the checkpointed one
    std::vector<size_t> bounds;
    for (size_t b = every; b < G; b += every) {
        bounds.push_back(b);
    }
    const size_t cap = std::max<size_t>(1, (size_t {32} << 30) / std::max<size_t>(bytes, 1));
    if (bounds.size() > cap) {
        bounds.clear();
        const size_t step = (G + cap) / (cap + 1);
        for (size_t b = step; b < G && bounds.size() < cap; b += step) {
            bounds.push_back(b);
        }
    }
    std::vector<size_t> seg {0};
    seg.insert(seg.end(), bounds.begin(), bounds.end());
    seg.push_back(G);
    std::vector<void*> ckpt(bounds.size(), nullptr);
    void* lam = nullptr;
    void* accs = nullptr;
    void* psi = nullptr;
    const auto cleanup = [&] {
        for (auto* p : ckpt) {
            if (p) {
                (void) hipFree(p);
            }
        }
        for (void* p : {lam, accs, psi}) {
            if (p) {
                (void) hipFree(p);
            }
        }
    };
    try {
        lam = device_alloc(bytes);
        for (auto& p : ckpt) {
            p = device_alloc(bytes);
        }
        // forward sweep with checkpoints
        for (size_t j = 0; j + 1 < seg.size(); ++j) {
            // checkpoints are compared with the uncomputed state amplitude by amplitude: qubit q on bit q
            run_records(s, std::vector<GateRec>(recs.begin() + static_cast<long>(seg[j]), recs.begin() + static_cast<long>(seg[j + 1])),
                        false);
            if (j < bounds.size()) {
                launch_copy(s, ckpt[j], s->amps, bytes);
            }
        }
        // lambda = H psi, E = <psi | lambda>
        launch_pauli_apply_sum(s, s->amps, lam, groups, strings);
        double e[2];
        launch_dot(s, s->amps, lam, e);
        *energy = e[0];
        // reverse sweep on a copy of psi (the state keeps psi(theta) for the caller)
        psi = device_alloc(bytes);
        launch_copy(s, psi, s->amps, bytes);
        std::vector<GateRec> derivs(G);
        std::vector<char> has_d(G, 0);
        for (size_t gi = 0; gi < G; ++gi) {
            has_d[gi] = ops[gi].aux >= 0 && derivative_of(ops[gi], recs[gi], derivs[gi]) ? 1 : 0;
        }
        const bool fused = fusion_max_k(s->precision, s->nvec) >= register_qubits(s->precision);
        std::fill(grad, grad + n_params, 0.0);
        std::vector<double> h;
        if (!fused) {
            accs = device_alloc(std::max<size_t>(G, 1) * 2 * sizeof(double));
        }
        size_t restores = 0;
        for (size_t j = seg.size() - 1; j-- > 0;) {
            const size_t b0 = seg[j], b1 = seg[j + 1];
            if (fused) {
                std::vector<GateRec> inv;
                std::vector<const GateRec*> dptr;
                inv.reserve(b1 - b0);
                for (size_t gi = b1; gi-- > b0;) {
                    inv.push_back(inverse_of(recs[gi]));
                    dptr.push_back(has_d[gi] ? &derivs[gi] : nullptr);
                }
                std::vector<double> values(inv.size(), 0.0);
                run_adjoint_segment(s, psi, lam, inv, dptr, values);
                for (size_t i = 0; i < inv.size(); ++i) {
                    const size_t gi = b1 - 1 - i;
                    if (has_d[gi]) {
                        grad[ops[gi].aux] += 2.0 * values[i];
                    }
                }
            }
            else {
                for (size_t gi = b1; gi-- > b0;) {
                    launch_adjoint_gate(s, psi, lam, inverse_of(recs[gi]), has_d[gi] ? &derivs[gi] : nullptr,
                                        has_d[gi] ? static_cast<double*>(accs) + 2 * gi : nullptr);
                }
                h.resize(2 * (b1 - b0));
                KQ_HIP_CHECK(hipMemcpyAsync(h.data(), static_cast<double*>(accs) + 2 * b0, h.size() * sizeof(double),
                                            hipMemcpyDeviceToHost, s->stream));
                KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
                for (size_t gi = b0; gi < b1; ++gi) {
                    if (has_d[gi]) {
                        grad[ops[gi].aux] += 2.0 * h[2 * (gi - b0)];
                    }
                }
            }
            // psi should now be the forward state after gates [0, b0): compare with the checkpoint
            if (j >= 1) {
                const double drift = launch_max_abs_diff(s, psi, ckpt[j - 1]);
                if (drift > drift_tol) {
                    launch_copy(s, psi, ckpt[j - 1], bytes);
                    ++restores;
                }
            }
        }
        KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
        if (n_restores) {
            *n_restores = restores;
        }
    }
    catch (...) {
        cleanup();
        throw;
    }
    cleanup();
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
        check_state(state);
        if (!energy || (n_params > 0 && !grad)) {
            throw InvalidArg {"null output"};
        }
        ensure_init();
        normalise_layout(state);
        last_stats() = PlanStats {};
        energy_gradient(state, ops, n_ops, n_strings, term_offsets, term_qubits, term_kinds, phases, coeffs, n_params,
                        energy, grad, checkpoint_every, drift_tol, n_restores);
    });
}

}  // extern "C"
