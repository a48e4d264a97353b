// Host-side and device-side debug entry points: pair maps, plan replays, generated pass sources.
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

namespace
{

// run-time pass kernel sources of a straight-line gate list (host only)
// One pass of a straight-line gate list as run_records plans it for a statevector (qubit moves and
// the final restore passes included): a generated tile kernel with its coefficients, or a sweep gate.
struct DebugPass
{
    bool sweep = false;
    bool out_of_place = false;   // stores into the second buffer, which then becomes the state's
    std::string src;
    std::vector<double> coefs;
    uint32_t blocks = 0;
    uint32_t threads = 0;
    GateRec gate {};
    PassInput input = PassInput::load;   // what its kernel starts from (runs from a basis state)
    uint64_t zmask = 0, zval = 0;        // skip_zero: the known outside bits; basis: zval = the index
};

// basis >= 0: the run starts from that basis state, not written yet (the first pass builds it, later
// passes skip the tiles known to be zero, the restore passes after the run start from a written state)
std::vector<DebugPass> debug_passes(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                                    int64_t basis = -1)
{
    if (n_qubits < 4 || n_qubits > 40 || (precision != KQ_C128 && precision != KQ_C64) || (!ops && n_ops)) {
        throw InvalidArg {"debug jit: bad arguments"};
    }
    std::vector<GateRec> recs;
    for (size_t i = 0; i < n_ops; ++i) {
        if (ops[i].gate < 0 || ops[i].gate >= KQ_M) {
            throw InvalidArg {"debug jit: op " + std::to_string(i) + " is not a unitary gate"};
        }
        compile_gate(ops[i], n_qubits, false, recs);
    }
    const int max_k = fusion_max_k_for(precision, n_qubits, lds_bytes ? lds_bytes : device_lds_per_block());
    if (max_k < register_qubits(precision)) {
        throw InvalidArg {"debug jit: state too small for a tile"};
    }
    std::vector<uint8_t> bit_of;
    std::vector<uint8_t> sigma;
    // as a gate run: the last pass puts every qubit back on its own bit when it can (in place, or out of
    // place into the second buffer)
    std::vector<uint8_t> final_bits;
    if (fusion_relabel()) {
        recs = exchanges_as_relabels(recs, n_qubits, sigma);
        final_bits.resize(n_qubits);
        for (uint32_t q = 0; q < n_qubits; ++q) {
            final_bits[sigma[q]] = static_cast<uint8_t>(q);
        }
    }
    auto passes = plan_passes(recs, n_qubits, max_k, fusion_low_qubits(), fusion_max_gates(), register_qubits(precision),
                              fusion_max_stages(), true, fusion_relabel() ? &bit_of : nullptr, nullptr,
                              fusion_relabel() ? &final_bits : nullptr, true);
    if (!sigma.empty() && bit_of.size() == n_qubits) {
        const std::vector<uint8_t> planned = bit_of;
        for (uint32_t q = 0; q < n_qubits; ++q) {
            bit_of[q] = planned[sigma[q]];
        }
    }
    const size_t run_passes = passes.size();
    bool moved = false;
    for (size_t q = 0; q < bit_of.size(); ++q) {
        moved = moved || bit_of[q] != q;
    }
    if (moved) {
        for (auto& p : plan_restore(bit_of, n_qubits, max_k, fusion_low_qubits(), register_qubits(precision))) {
            passes.push_back(std::move(p));
        }
    }
    if (basis >= 0 && static_cast<uint64_t>(basis) >= (uint64_t {1} << n_qubits)) {
        throw InvalidArg {"debug jit: basis index out of range"};
    }
    KnownZeros kz;
    if (basis >= 0) {
        kz.mask = (uint64_t {1} << n_qubits) - 1;
        kz.val = static_cast<uint64_t>(basis);
        kz.basis = true;
    }
    std::vector<DebugPass> out;
    for (size_t i = 0; i < passes.size(); ++i) {
        const TilePass& p = passes[i];
        DebugPass d;
        if (i == run_passes) {
            kz.forget();   // a read restores the order, and writes a pending basis state first
        }
        if (p.sweep) {
            d.sweep = true;
            d.gate = p.single;
            kz.forget();
        }
        else {
            TileProgram tp;
            lower_tile_pass(p, precision, n_qubits, tp);
            d.input = kz.input_for(tp, i == 0, true, d.zmask, d.zval);
            kz.after(tp);
            d.src = jit_source(tp, precision, d.coefs, d.input);
            d.out_of_place = !tp.out_freepos.empty();
            d.blocks = tile_pass_blocks(tp);
            d.threads = 1u << tp.n_thr_bits;
        }
        out.push_back(std::move(d));
    }
    return out;
}

std::vector<std::string> debug_jit_sources(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops,
                                           size_t lds_bytes)
{
    std::vector<std::string> srcs;
    for (auto& d : debug_passes(n_qubits, precision, ops, n_ops, lds_bytes)) {
        if (!d.sweep) {
            srcs.push_back(std::move(d.src));
        }
    }
    return srcs;
}

}  // namespace

}  // namespace kq

using namespace kq;

extern "C" {

int kq_debug_pair_map_1q(uint32_t n_qubits, uint32_t target, uint64_t* s0, uint64_t* s1)
{
    return guarded([&] {
        if (n_qubits == 0 || target >= n_qubits || n_qubits > 30) {
            throw InvalidArg {"pair map: bad sizes"};
        }
        ensure_init();
        const uint64_t np = uint64_t {1} << (n_qubits - 1);
        void* d = device_alloc(2 * np * sizeof(uint64_t));
        auto* d0 = static_cast<uint64_t*>(d);
        launch_pair_map_1q(n_qubits, target, d0, d0 + np);
        KQ_HIP_CHECK(hipDeviceSynchronize());
        KQ_HIP_CHECK(hipMemcpy(s0, d0, np * sizeof(uint64_t), hipMemcpyDeviceToHost));
        KQ_HIP_CHECK(hipMemcpy(s1, d0 + np, np * sizeof(uint64_t), hipMemcpyDeviceToHost));
        KQ_HIP_CHECK(hipFree(d));
    });
}

int kq_debug_pair_map_2q(uint32_t n_qubits, uint32_t control, uint32_t target, int four_way, uint64_t* out)
{
    return guarded([&] {
        if (n_qubits < 2 || control >= n_qubits || target >= n_qubits || control == target || n_qubits > 30) {
            throw InvalidArg {"pair map: bad sizes"};
        }
        ensure_init();
        const uint64_t np = uint64_t {1} << (n_qubits - 2);
        const uint64_t words = np * (four_way ? 4 : 2);
        void* d = device_alloc(words * sizeof(uint64_t));
        launch_pair_map_2q(n_qubits, control, target, four_way, static_cast<uint64_t*>(d));
        KQ_HIP_CHECK(hipDeviceSynchronize());
        KQ_HIP_CHECK(hipMemcpy(out, d, words * sizeof(uint64_t), hipMemcpyDeviceToHost));
        KQ_HIP_CHECK(hipFree(d));
    });
}

int kq_debug_check_relabel(uint32_t n_qubits, const kq_op* ops, size_t n_ops, double* amps, size_t* n_passes,
                           size_t* n_restore_passes)
{
    return guarded([&] {
        if (n_qubits < 1 || n_qubits > 24 || (!ops && n_ops) || !amps) {
            throw InvalidArg {"debug relabel: bad arguments"};
        }
        std::vector<GateRec> recs;
        for (size_t i = 0; i < n_ops; ++i) {
            if (ops[i].gate < 0 || ops[i].gate >= KQ_M) {
                throw InvalidArg {"debug relabel: op " + std::to_string(i) + " is not a unitary gate"};
            }
            compile_gate(ops[i], n_qubits, false, recs);
        }
        const int max_k = fusion_max_k_for(KQ_C128, n_qubits, device_lds_per_block());
        std::vector<uint8_t> bit_of(n_qubits);
        for (uint32_t q = 0; q < n_qubits; ++q) {
            bit_of[q] = static_cast<uint8_t>(q);
        }
        std::vector<uint8_t> planned = bit_of;
        std::vector<uint8_t> sigma;
        recs = exchanges_as_relabels(recs, n_qubits, sigma);
        const auto passes = plan_passes(recs, n_qubits, max_k, fusion_low_qubits(), fusion_max_gates(), register_qubits(KQ_C128),
                                        fusion_max_stages(), true, &planned);
        const size_t N = size_t {1} << n_qubits;
        auto* z = reinterpret_cast<std::complex<double>*>(amps);
        // the passes' gates in execution order on the host, each on the index bits holding its qubits
        const auto apply = [&](const GateRec& g) {
            const std::complex<double> m00 {g.m[0], g.m[1]}, m01 {g.m[2], g.m[3]}, m10 {g.m[4], g.m[5]}, m11 {g.m[6], g.m[7]};
            const size_t tb = size_t {1} << g.target;
            for (size_t i = 0; i < N; ++i) {
                if ((i & tb) || (g.control >= 0 && !((i >> g.control) & 1u))) {
                    continue;
                }
                const auto a = z[i], b = z[i | tb];
                z[i] = m00 * a + m01 * b;
                z[i | tb] = m10 * a + m11 * b;
            }
        };
        const auto move = [&](const TilePass& p) {
            if (p.out_qubits.empty()) {
                return;
            }
            std::vector<std::complex<double>> out(N);
            for (size_t i = 0; i < N; ++i) {
                size_t j = i;
                for (size_t l = 0; l < p.qubits.size(); ++l) {
                    j &= ~(size_t {1} << p.out_qubits[l]);
                }
                for (size_t l = 0; l < p.qubits.size(); ++l) {
                    j |= ((i >> p.qubits[l]) & 1u) << p.out_qubits[l];
                }
                out[j] = z[i];
            }
            std::copy(out.begin(), out.end(), z);
            std::vector<uint8_t> nb = bit_of;
            for (uint32_t q = 0; q < n_qubits; ++q) {
                for (size_t l = 0; l < p.qubits.size(); ++l) {
                    if (bit_of[q] == p.qubits[l]) {
                        nb[q] = static_cast<uint8_t>(p.out_qubits[l]);
                    }
                }
            }
            bit_of = nb;
        };
        for (const auto& p : passes) {
            for (const uint32_t r : p.src) {
                GateRec g = recs[r];
                g.target = bit_of[g.target];
                if (g.control >= 0) {
                    g.control = bit_of[static_cast<size_t>(g.control)];
                }
                apply(g);
            }
            move(p);
        }
        if (bit_of != planned) {
            throw InvalidArg {"debug relabel: replayed qubit placement differs from the planner's"};
        }
        // exchanged qubits: qubit q's amplitudes are those of the relabelled qubit sigma[q]
        for (uint32_t q = 0; q < n_qubits; ++q) {
            bit_of[q] = planned[sigma[q]];
        }
        // the one-pass out-of-place restore, its index arithmetic replayed on the host, must give what
        // the in-place restore passes give, bit for bit
        std::vector<std::complex<double>> one_pass;
        if (n_qubits >= 12) {
            const PermuteTables t = permute_tables(bit_of);
            one_pass.assign(N, {});
            std::vector<int> seen(N, 0);
            for (uint64_t blk = 0; blk < (N >> 12); ++blk) {
                uint64_t sb = 0, db = 0;
                for (uint32_t i = 0; i < t.n_free; ++i) {
                    if ((blk >> i) & 1u) {
                        sb |= uint64_t {1} << t.free_src[i];
                        db |= uint64_t {1} << t.free_dst[i];
                    }
                }
                for (uint32_t l = 0; l < 4096; ++l) {
                    uint64_t so = sb, dof = db;
                    for (int j = 0; j < 12; ++j) {
                        if ((l >> j) & 1u) {
                            so |= t.src_of[j];
                            dof |= t.dst_of[j];
                        }
                    }
                    if (so >= N || dof >= N || seen[dof]++) {
                        throw InvalidArg {"debug relabel: one-pass restore is not a permutation"};
                    }
                    one_pass[dof] = z[so];
                }
            }
            for (const uint8_t* lay : {t.ld_thr, t.st_thr}) {
                uint32_t m = 0;
                for (int j = 0; j < 8; ++j) {
                    m |= 1u << lay[j];
                }
                const uint8_t* reg = lay == t.ld_thr ? t.ld_reg : t.st_reg;
                for (int q = 0; q < 4; ++q) {
                    m |= 1u << reg[q];
                }
                if (m != 4095u) {
                    throw InvalidArg {"debug relabel: one-pass restore layout does not cover the tile"};
                }
            }
        }
        const auto restore = plan_restore(bit_of, n_qubits, max_k, fusion_low_qubits(), register_qubits(KQ_C128));
        for (const auto& p : restore) {
            move(p);
        }
        for (uint32_t q = 0; q < n_qubits; ++q) {
            if (bit_of[q] != q) {
                throw InvalidArg {"debug relabel: restore left a qubit off its bit"};
            }
        }
        if (!one_pass.empty() && !std::equal(one_pass.begin(), one_pass.end(), z)) {
            throw InvalidArg {"debug relabel: one-pass restore differs from the restore passes"};
        }
        if (n_passes) {
            *n_passes = passes.size();
        }
        if (n_restore_passes) {
            *n_restore_passes = restore.size();
        }
    });
}

int kq_debug_jit_compile(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                         size_t* n_kernels, size_t* failures)
{
    return guarded([&] {
        if (!n_kernels || !failures) {
            throw InvalidArg {"debug jit: null output"};
        }
        const auto srcs = debug_jit_sources(n_qubits, precision, ops, n_ops, lds_bytes);
        std::string log;
        *failures = jit_compile_only(srcs, &log);
        *n_kernels = srcs.size();
        if (*failures) {
            throw InvalidArg {"debug jit: " + std::to_string(*failures) + " of " + std::to_string(srcs.size())
                              + " pass kernels failed to compile:\n" + log.substr(0, 4000)};
        }
    });
}

int kq_debug_adjoint_jit_compile(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                                 size_t* n_kernels, size_t* failures)
{
    return guarded([&] {
        if (!n_kernels || !failures || n_qubits < 4 || n_qubits > 40 || (precision != KQ_C128 && precision != KQ_C64)
            || (!ops && n_ops)) {
            throw InvalidArg {"debug adjoint jit: bad arguments"};
        }
        std::vector<GateRec> recs;
        for (size_t i = 0; i < n_ops; ++i) {
            if (ops[i].gate < 0 || ops[i].gate >= KQ_M) {
                throw InvalidArg {"debug adjoint jit: op " + std::to_string(i) + " is not a unitary gate"};
            }
            compile_gate(ops[i], n_qubits, false, recs);
        }
        std::vector<GateRec> derivs(recs.size()), inv;
        std::vector<const GateRec*> dptr;
        for (size_t gi = recs.size(); gi-- > 0;) {
            inv.push_back(inverse_of(recs[gi]));
            const bool d = ops[gi].aux >= 0 && derivative_of(ops[gi], recs[gi], derivs[gi]);
            dptr.push_back(d ? &derivs[gi] : nullptr);
        }
        const int max_k = fusion_max_k_for(precision, n_qubits, lds_bytes ? lds_bytes : device_lds_per_block());
        if (max_k < register_qubits(precision)) {
            throw InvalidArg {"debug adjoint jit: state too small for a tile"};
        }
        const auto passes = plan_passes(inv, n_qubits, max_k, fusion_low_qubits(), fusion_max_gates(),
                                        register_qubits(precision), fusion_max_stages(), false);
        std::vector<std::string> srcs;
        for (const auto& p : passes) {
            TileProgram tp;
            lower_tile_pass(p, precision, n_qubits, tp, &dptr);
            AdjointCoefs coefs;
            srcs.push_back(jit_adjoint_source(tp, precision, coefs));
        }
        std::string log;
        *failures = jit_compile_only(srcs, &log);
        *n_kernels = srcs.size();
        if (*failures) {
            throw InvalidArg {"debug adjoint jit: " + std::to_string(*failures) + " of " + std::to_string(srcs.size())
                              + " adjoint pass kernels failed to compile:\n" + log.substr(0, 4000)};
        }
    });
}

int kq_debug_jit_source(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                        size_t pass_index, char* buf, size_t cap, size_t* len, size_t* n_kernels)
{
    return guarded([&] {
        if (!len || !n_kernels) {
            throw InvalidArg {"debug jit: null output"};
        }
        const auto srcs = debug_jit_sources(n_qubits, precision, ops, n_ops, lds_bytes);
        *n_kernels = srcs.size();
        if (pass_index >= srcs.size()) {
            throw InvalidArg {"debug jit: pass index out of range"};
        }
        const std::string& src = srcs[pass_index];
        *len = src.size();
        if (buf && cap) {
            const size_t n = std::min(cap - 1, src.size());
            std::memcpy(buf, src.data(), n);
            buf[n] = '\0';
        }
    });
}

int kq_debug_pass(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes, size_t pass_index,
                  char* src, size_t src_cap, size_t* src_len, double* coefs, size_t coef_cap, size_t* n_coefs,
                  uint64_t info[5], size_t* n_passes)
{
    return guarded([&] {
        if (!src_len || !n_coefs || !info || !n_passes) {
            throw InvalidArg {"debug pass: null output"};
        }
        const auto passes = debug_passes(n_qubits, precision, ops, n_ops, lds_bytes);
        *n_passes = passes.size();
        if (pass_index >= passes.size()) {
            throw InvalidArg {"debug pass: pass index out of range"};
        }
        const DebugPass& d = passes[pass_index];
        *src_len = d.src.size();
        *n_coefs = d.sweep ? 8 : d.coefs.size();
        if (src && src_cap) {
            const size_t n = std::min(src_cap - 1, d.src.size());
            std::memcpy(src, d.src.data(), n);
            src[n] = '\0';
        }
        if (coefs) {
            const double* from = d.sweep ? d.gate.m : d.coefs.data();
            std::memcpy(coefs, from, std::min(coef_cap, *n_coefs) * sizeof(double));
        }
        info[0] = d.sweep ? 1 : (d.out_of_place ? 2 : 0);
        info[1] = d.blocks;
        info[2] = d.threads;
        info[3] = d.sweep ? d.gate.target : 0;
        info[4] = d.sweep ? static_cast<uint64_t>(d.gate.control + 1) : 0;
    });
}

int kq_debug_pass_from_basis(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                             size_t pass_index, uint64_t basis, char* src, size_t src_cap, size_t* src_len, double* coefs,
                             size_t coef_cap, size_t* n_coefs, uint64_t info[8], size_t* n_passes)
{
    return guarded([&] {
        if (!src_len || !n_coefs || !info || !n_passes) {
            throw InvalidArg {"debug pass: null output"};
        }
        if (basis >= (uint64_t {1} << 62)) {
            throw InvalidArg {"debug pass: basis index out of range"};
        }
        const auto passes = debug_passes(n_qubits, precision, ops, n_ops, lds_bytes, static_cast<int64_t>(basis));
        *n_passes = passes.size();
        if (pass_index >= passes.size()) {
            throw InvalidArg {"debug pass: pass index out of range"};
        }
        const DebugPass& d = passes[pass_index];
        *src_len = d.src.size();
        *n_coefs = d.sweep ? 8 : d.coefs.size();
        if (src && src_cap) {
            const size_t n = std::min(src_cap - 1, d.src.size());
            std::memcpy(src, d.src.data(), n);
            src[n] = '\0';
        }
        if (coefs) {
            const double* from = d.sweep ? d.gate.m : d.coefs.data();
            std::memcpy(coefs, from, std::min(coef_cap, *n_coefs) * sizeof(double));
        }
        info[0] = d.sweep ? 1 : (d.out_of_place ? 2 : 0);
        info[1] = d.blocks;
        info[2] = d.threads;
        info[3] = d.sweep ? d.gate.target : 0;
        info[4] = d.sweep ? static_cast<uint64_t>(d.gate.control + 1) : 0;
        info[5] = static_cast<uint64_t>(d.input);
        info[6] = d.zmask;
        info[7] = d.zval;
    });
}

int kq_debug_plan(uint32_t n_qubits, int precision, int is_density, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                  uint32_t* order, uint32_t* pass_of, int32_t* pass_kind, double* pass_cost, size_t* n_passes)
{
    return guarded([&] {
        if (n_qubits == 0 || n_qubits > 40 || (precision != KQ_C128 && precision != KQ_C64) || (!ops && n_ops)
            || !order || !pass_of || !pass_kind || !pass_cost || !n_passes) {
            throw InvalidArg {"debug plan: bad arguments"};
        }
        std::vector<GateRec> recs;
        for (size_t i = 0; i < n_ops; ++i) {
            if (ops[i].gate < 0 || ops[i].gate >= KQ_M) {
                throw InvalidArg {"debug plan: op " + std::to_string(i) + " is not a unitary gate"};
            }
            compile_gate(ops[i], n_qubits, is_density != 0, recs);
        }
        const uint32_t nvec = is_density ? 2 * n_qubits : n_qubits;
        const int max_k = fusion_max_k_for(precision, nvec, lds_bytes ? lds_bytes : device_lds_per_block());
        size_t np = 0, j = 0;
        if (max_k < register_qubits(precision)) {
            for (size_t r = 0; r < recs.size(); ++r, ++np) {
                order[j] = static_cast<uint32_t>(r);
                pass_of[j++] = static_cast<uint32_t>(np);
                pass_kind[np] = 2;
                pass_cost[np] = gate_cost(recs[r]);
            }
        }
        else {
            std::vector<uint8_t> bit_of;   // as run_records plans a statevector: qubits may move between bits
            const bool relabel = fusion_relabel() && !is_density;
            const auto passes = plan_passes(recs, nvec, max_k, fusion_low_qubits(), fusion_max_gates(), register_qubits(precision),
                                            fusion_max_stages(), true, relabel ? &bit_of : nullptr);
            for (const auto& p : passes) {
                if (!p.sweep) {
                    TileProgram tp;
                    lower_tile_pass(p, precision, nvec, tp);   // validates the layouts and variants
                }
                double cost = 0.0;
                for (const uint32_t r : p.src) {
                    order[j] = r;
                    pass_of[j++] = static_cast<uint32_t>(np);
                    cost += gate_cost(recs[r]);
                }
                pass_kind[np] = p.sweep ? 2 : (p.c >= p.k ? 0 : 1);
                pass_cost[np] = cost;
                ++np;
            }
        }
        if (j != recs.size()) {
            throw InvalidArg {"debug plan: the plan does not apply every gate exactly once"};
        }
        *n_passes = np;
    });
}

}  // extern "C"

int kq_debug_dense_block(kq_state* s, const double* u, int use_mfma, double* ms)
{
    return guarded([&] {
        check_state(s);
        if (!u) {
            throw InvalidArg {"dense block: null matrix"};
        }
        if (s->is_dm) {
            throw InvalidArg {"dense block: statevectors only"};
        }
        ensure_init();
        normalise_layout(s);
        // U in the amplitude precision, in the state's scratch (the reductions' scratch is not in use here)
        const size_t bytes = 512 * amp_bytes(s->precision) / 2;
        void* d_u = scratch(s, bytes);
        if (s->precision == KQ_C128) {
            KQ_HIP_CHECK(hipMemcpyAsync(d_u, u, bytes, hipMemcpyHostToDevice, s->stream));
        }
        else {
            std::vector<float> f(u, u + 512);
            KQ_HIP_CHECK(hipMemcpyAsync(d_u, f.data(), bytes, hipMemcpyHostToDevice, s->stream));
            KQ_HIP_CHECK(hipStreamSynchronize(s->stream));   // f leaves scope
        }
        hipEvent_t a, b;
        KQ_HIP_CHECK(hipEventCreate(&a));
        KQ_HIP_CHECK(hipEventCreate(&b));
        try {
            KQ_HIP_CHECK(hipEventRecord(a, s->stream));
            launch_dense4(s, d_u, use_mfma != 0);
            KQ_HIP_CHECK(hipEventRecord(b, s->stream));
            KQ_HIP_CHECK(hipEventSynchronize(b));
            float t = 0.0f;
            KQ_HIP_CHECK(hipEventElapsedTime(&t, a, b));
            if (ms) {
                *ms = t;
            }
        }
        catch (...) {
            (void) hipEventDestroy(a);
            (void) hipEventDestroy(b);
            throw;
        }
        (void) hipEventDestroy(a);
        (void) hipEventDestroy(b);
    });
}

int kq_debug_mfma16_layout(int precision, const double* a, const double* b, double* d)
{
    return guarded([&] {
        if (!a || !b || !d || (precision != KQ_C128 && precision != KQ_C64)) {
            throw InvalidArg {"mfma layout: bad arguments"};
        }
        ensure_init();
        const size_t es = precision == KQ_C128 ? 8 : 4;
        void* buf = nullptr;
        KQ_HIP_CHECK(hipMalloc(&buf, 384 * es));
        try {
            std::vector<unsigned char> h(384 * es);
            for (int i = 0; i < 64; ++i) {
                if (es == 8) {
                    std::memcpy(h.data() + i * 8, a + i, 8);
                    std::memcpy(h.data() + (64 + i) * 8, b + i, 8);
                }
                else {
                    const float fa = static_cast<float>(a[i]), fb = static_cast<float>(b[i]);
                    std::memcpy(h.data() + i * 4, &fa, 4);
                    std::memcpy(h.data() + (64 + i) * 4, &fb, 4);
                }
            }
            KQ_HIP_CHECK(hipMemcpy(buf, h.data(), 128 * es, hipMemcpyHostToDevice));
            auto* base = static_cast<unsigned char*>(buf);
            launch_mfma16_layout(nullptr, precision, base, base + 64 * es, base + 128 * es);
            KQ_HIP_CHECK(hipDeviceSynchronize());
            KQ_HIP_CHECK(hipMemcpy(h.data(), base + 128 * es, 256 * es, hipMemcpyDeviceToHost));
            for (int i = 0; i < 256; ++i) {
                if (es == 8) {
                    std::memcpy(d + i, h.data() + i * 8, 8);
                }
                else {
                    float f;
                    std::memcpy(&f, h.data() + i * 4, 4);
                    d[i] = f;
                }
            }
        }
        catch (...) {
            (void) hipFree(buf);
            throw;
        }
        (void) hipFree(buf);
    });
}
