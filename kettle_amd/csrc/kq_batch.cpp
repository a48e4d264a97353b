// Batched energy + adjoint gradient of many independent parameterised circuits of uneven qubit
// counts (the VQE inner loop of kettle's optimize/ drivers, run for thousands of instances at once).
//
// Every instance is planned once into fused tile passes (forward) and adjoint tile passes (reverse
// sweep with derivative partials), with angle-independent gate kinds so that new parameter values
// only re-lower coefficients.  Execution is a fixed sequence of launches over all instances:
//   reset every state to |0...0>
//   forward round r of each tile-size group: pass r of every instance that has one, in one launch
//   lambda = H psi and the energy partials of every instance, in one launch (with a complex128
//   forward sweep and a complex64 reverse sweep: energies from the complex128 states, then psi and
//   lambda narrowed to complex64)
//   reverse round r of each group, then one FP64 row sum of that round's derivative partials
//   distance of every uncomputed state from |0...0> (inverse-gate drift), in one launch
// The sequence contains no allocation, no host synchronisation and no memset, so after a first
// eager run it is captured in a hipGraph and replayed by later calls.
// With a drift tolerance set, every instance whose drift exceeds it after the run is evaluated again
// by the single-state checkpointed adjoint in complex128 (on a work state per qubit count, kept by
// the batch), and that energy and gradient replace the batched ones.
#include <algorithm>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kq_host.hpp"
#include "kq_internal.hpp"

namespace kq
{

struct BatchInstance
{
    uint32_t n = 0;                    // qubits as given
    uint32_t nv = 0;                   // qubits simulated (at least 4: one register block)
    std::vector<kq_op> ops;
    size_t n_params = 0;
    std::vector<double> params;
    bool dirty = true;                 // parameters changed since its programs were last written
    std::vector<PauliGroup> groups;
    std::vector<PauliZ> strings;
    // plan
    int k = 0;
    std::vector<TilePass> fwd, bwd;
    std::vector<uint32_t> bwd_gate;    // reversed position -> gate index
    std::vector<char> has_d;
    // lowered programs and their byte offsets in the program arena
    std::vector<TileProgram> fwd_prog, bwd_prog;
    std::vector<size_t> fwd_off, bwd_off;
    size_t fwd_end = 0, bwd_end = 0;   // end of the last forward / reverse program
    // per reverse pass, per derivative slot: |F|^2, F the pivots divided out of the gates before it
    std::vector<std::vector<double>> bwd_dscale;
    // per forward pass: the index bits known to be 0 wherever the state is non-zero when it starts (the
    // sweep starts from |0...0>; a pass's tile bits are unknown after it)
    std::vector<uint64_t> fwd_zero_mask;
    uint64_t amp_off = 0;              // offset of psi / lambda in the arenas (amplitudes)
    uint64_t n_amps = 0;
    size_t grad_off = 0;               // first gradient entry of the instance in the output
    size_t pauli_off = 0;              // byte offset of the groups (strings follow) in the Pauli arena
    // the operator as given, for a single-state recompute
    std::vector<uint32_t> h_off, h_q;
    std::vector<uint8_t> h_k, h_ph;
    std::vector<double> h_c;
};

struct BatchRound
{
    bool adjoint = false;
    int k = 0;
    uint32_t n_items = 0;
    uint32_t n_blocks = 0;
    size_t items_off = 0;              // byte offset of the BatchTileItem array in the table arena
    // adjoint rounds: one row per derivative slot
    uint32_t n_rows = 0;
    size_t row_off_off = 0;            // byte offsets of the uint64 row offsets / uint32 row counts
    size_t row_cnt_off = 0;
    size_t row_base = 0;               // first global row of the round
};

// one derivative slot of a reverse pass: its instance, gate, pass and slot
struct BatchRow
{
    uint32_t inst;
    uint32_t gate;
    uint32_t pass;
    uint32_t slot;
};

}  // namespace kq

struct kq_batch
{
    int precision = KQ_C64;             // reverse sweep and lambda
    int fwd_precision = KQ_C64;         // forward sweep and energy
    hipStream_t stream = nullptr;
    std::vector<kq::BatchInstance> inst;
    bool planned = false;
    bool coeffs_dirty = true;
    // plan-wide tables
    std::vector<kq::BatchRound> rounds;
    std::vector<kq::BatchRow> rows;     // global derivative rows
    uint64_t total_amps = 0;
    uint64_t part_doubles = 0;          // largest adjoint round's partials
    uint32_t pauli_blocks = 0;
    size_t pauli_items_off = 0;
    size_t e_row_off_off = 0, e_row_cnt_off = 0;
    size_t ones_off = 0;
    bool reset = true;                  // some instance has no forward pass: the arena is reset first
    size_t prog_bytes = 0;
    size_t table_bytes = 0;
    size_t pauli_bytes = 0;
    size_t fwd_passes = 0, rev_passes = 0;
    double algorithmic_bytes = 0.0;     // per run: one read and/or write of every amplitude each launch updates
    // device memory
    void* psi_fwd = nullptr;            // forward states when the forward precision differs
    void* psi = nullptr;
    void* lam = nullptr;
    void* prog = nullptr;
    void* tables = nullptr;
    void* pauli = nullptr;
    double* part = nullptr;             // derivative partials (largest round), then energy partials
    double* slot_values = nullptr;
    double* energies = nullptr;
    double* drift = nullptr;
    // host staging of the program arena (pinned)
    unsigned char* prog_host = nullptr;
    // captured launch sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    size_t n_launches = 0;
    // drift-triggered recompute (disabled while negative)
    double recompute_tol = -1.0;
    size_t last_recomputed = 0;
    size_t total_recomputed = 0;
    std::vector<std::pair<uint32_t, kq_state*>> work;   // complex128 work state per qubit count
};

namespace kq
{
namespace
{

size_t align256(size_t x) { return (x + 255) & ~size_t {255}; }

size_t program_bytes(const TileProgram& tp)
{
    return align256(tp.gates.size() * sizeof(TileGate)) + align256(tp.steps.size() * sizeof(TileStep))
         + align256(tp.layouts.size() * sizeof(TileLayout)) + align256(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate))
         + align256(std::max<size_t>(tp.terms.size(), 1) * sizeof(TileTerm));
}

void write_program(const TileProgram& tp, unsigned char* dst)
{
    std::memcpy(dst, tp.gates.data(), tp.gates.size() * sizeof(TileGate));
    dst += align256(tp.gates.size() * sizeof(TileGate));
    std::memcpy(dst, tp.steps.data(), tp.steps.size() * sizeof(TileStep));
    dst += align256(tp.steps.size() * sizeof(TileStep));
    std::memcpy(dst, tp.layouts.data(), tp.layouts.size() * sizeof(TileLayout));
    dst += align256(tp.layouts.size() * sizeof(TileLayout));
    std::memcpy(dst, tp.derivs.data(), tp.derivs.size() * sizeof(TileGate));
    dst += align256(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate));
    std::memcpy(dst, tp.terms.data(), tp.terms.size() * sizeof(TileTerm));
}

// records of the instance's current angles: forward records, inverse records in reverse order and
// derivative records (indexed like the reversed list)
void compile_instance(const BatchInstance& in, std::vector<GateRec>& recs, std::vector<GateRec>& inv,
                      std::vector<GateRec>& derivs, std::vector<char>& has_d)
{
    const size_t G = in.ops.size();
    recs.clear();
    recs.reserve(G);
    for (const kq_op& op0 : in.ops) {
        kq_op op = op0;
        if (op.aux >= 0) {
            op.angle = in.params[static_cast<size_t>(op.aux)];
        }
        compile_gate_stable(op, in.nv, recs);
    }
    inv.resize(G);
    derivs.assign(G, GateRec {});
    has_d.assign(G, 0);
    for (size_t i = 0; i < G; ++i) {
        const size_t gi = G - 1 - i;
        inv[i] = inverse_of(recs[gi]);
        kq_op op = in.ops[gi];
        if (op.aux >= 0) {
            op.angle = in.params[static_cast<size_t>(op.aux)];
            has_d[i] = derivative_of(op, recs[gi], derivs[i]) ? 1 : 0;
        }
    }
}

void plan_instance(BatchInstance& in, int precision, int fwd_precision)
{
    std::vector<GateRec> recs, inv, derivs;
    std::vector<char> has_d;
    compile_instance(in, recs, inv, derivs, has_d);
    in.k = std::min(fusion_max_k(precision, in.nv), fusion_max_k(fwd_precision, in.nv));
    if (in.k < register_qubits(precision)) {
        throw InvalidArg {"batch: the tile does not hold one register block"};
    }
    in.fwd = plan_passes(recs, in.nv, in.k, fusion_low_qubits(), fusion_max_gates(), register_qubits(precision), fusion_max_stages(), false);
    in.bwd = plan_passes(inv, in.nv, in.k, fusion_low_qubits(), fusion_max_gates(), register_qubits(precision), fusion_max_stages(), false);
    in.has_d = has_d;
    in.bwd_gate.resize(inv.size());
    for (size_t i = 0; i < inv.size(); ++i) {
        in.bwd_gate[i] = static_cast<uint32_t>(inv.size() - 1 - i);
    }
}

// lower the instance's passes with its current angles (the plan and the program sizes are fixed)
void lower_instance(BatchInstance& in, int precision, int fwd_precision)
{
    std::vector<GateRec> recs, inv, derivs;
    std::vector<char> has_d;
    compile_instance(in, recs, inv, derivs, has_d);
    std::vector<const GateRec*> dptr(inv.size(), nullptr);
    for (size_t i = 0; i < inv.size(); ++i) {
        if (has_d[i]) {
            dptr[i] = &derivs[i];
        }
    }
    const auto patch = [](TilePass& p, const std::vector<GateRec>& src) {
        for (size_t j = 0; j < p.gates.size(); ++j) {
            std::memcpy(p.gates[j].m, src[p.src[j]].m, sizeof p.gates[j].m);
        }
    };
    in.fwd_prog.resize(in.fwd.size());
    for (size_t p = 0; p < in.fwd.size(); ++p) {
        patch(in.fwd[p], recs);
        lower_tile_pass(in.fwd[p], fwd_precision, in.nv, in.fwd_prog[p]);
        normalise_program(in.fwd_prog[p], fwd_precision, nullptr);
    }
    in.bwd_prog.resize(in.bwd.size());
    in.bwd_dscale.resize(in.bwd.size());
    for (size_t p = 0; p < in.bwd.size(); ++p) {
        patch(in.bwd[p], inv);
        lower_tile_pass(in.bwd[p], precision, in.nv, in.bwd_prog[p], &dptr);
        normalise_program(in.bwd_prog[p], precision, &in.bwd_dscale[p]);
    }
}

void* balloc(size_t bytes) { return device_alloc(std::max<size_t>(bytes, 256)); }

void release(kq_batch* b)
{
    if (b->graph_exec) {
        (void) hipGraphExecDestroy(b->graph_exec);
        b->graph_exec = nullptr;
    }
    if (b->graph) {
        (void) hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    for (void** p : {&b->psi_fwd, &b->psi, &b->lam, &b->prog, &b->tables, &b->pauli}) {
        if (*p) {
            (void) hipFree(*p);
            *p = nullptr;
        }
    }
    if (b->prog_host) {
        (void) hipHostFree(b->prog_host);
        b->prog_host = nullptr;
    }
    for (double** p : {&b->part, &b->slot_values, &b->energies, &b->drift}) {
        if (*p) {
            (void) hipFree(*p);
            *p = nullptr;
        }
    }
}

uint64_t free_mask_of(const TileProgram& tp)
{
    uint64_t m = 0;
    for (const uint8_t q : tp.freepos) {
        m |= uint64_t {1} << q;
    }
    return m;
}

// plan every instance, place states, programs and tables, allocate device memory
void build(kq_batch* b)
{
    release(b);
    b->rounds.clear();
    b->rows.clear();
    const size_t ab = amp_bytes(b->precision);
    // instances: plan, place states and programs
    uint64_t amps = 0;
    size_t prog = 0, pauli = 0, grad = 0;
    for (BatchInstance& in : b->inst) {
        plan_instance(in, b->precision, b->fwd_precision);
        lower_instance(in, b->precision, b->fwd_precision);
        in.amp_off = amps;
        in.n_amps = uint64_t {1} << in.nv;
        amps += in.n_amps;
        in.fwd_off.clear();
        in.bwd_off.clear();
        for (const TileProgram& tp : in.fwd_prog) {
            in.fwd_off.push_back(prog);
            prog += program_bytes(tp);
        }
        in.fwd_end = prog;
        for (const TileProgram& tp : in.bwd_prog) {
            in.bwd_off.push_back(prog);
            prog += program_bytes(tp);
        }
        in.bwd_end = prog;
        in.pauli_off = pauli;
        pauli += align256(in.groups.size() * sizeof(PauliGroup)) + align256(in.strings.size() * sizeof(PauliZ));
        in.grad_off = grad;
        grad += in.n_params;
    }
    b->total_amps = amps;
    {
        const double af = static_cast<double>(amp_bytes(b->fwd_precision));
        const double ar = static_cast<double>(ab);
        b->fwd_passes = b->rev_passes = 0;
        b->algorithmic_bytes = 0.0;
        b->reset = false;
        for (BatchInstance& in : b->inst) {
            b->reset = b->reset || in.fwd_prog.empty();
        }
        for (BatchInstance& in : b->inst) {
            const double A = static_cast<double>(in.n_amps);
            b->fwd_passes += in.fwd.size();
            b->rev_passes += in.bwd.size();
            // Forward passes from |0...0>: the first writes the state without reading it (its tile 0 built,
            // the other tiles stored as zeros); a later one reads and writes the tiles whose outside bits
            // under the known-zero mask are all 0 and skips the others.
            in.fwd_zero_mask.assign(in.fwd_prog.size(), 0);
            uint64_t known = in.n_amps - 1;
            double fwd_bytes = 0.0;
            for (size_t p = 0; p < in.fwd_prog.size(); ++p) {
                uint64_t free_mask = 0;
                for (const uint8_t q : in.fwd_prog[p].freepos) {
                    free_mask |= uint64_t {1} << q;
                }
                in.fwd_zero_mask[p] = known & free_mask;
                const double live = A / static_cast<double>(uint64_t {1} << std::popcount(in.fwd_zero_mask[p]));
                fwd_bytes += p == 0 ? af * A : 2.0 * af * live;
                known &= free_mask;
            }
            // (reset), forward passes, lambda = H psi, narrowing, reverse passes (two vectors), drift read
            b->algorithmic_bytes += A * (af + ar + (b->fwd_precision != b->precision ? af + ar : 0.0)
                                         + 4.0 * ar * static_cast<double>(in.bwd.size()) + ar)
                                  + fwd_bytes;
        }
        if (b->reset) {
            b->algorithmic_bytes += af * static_cast<double>(b->total_amps);
        }
    }
    b->prog_bytes = std::max<size_t>(prog, 256);
    b->pauli_bytes = std::max<size_t>(pauli, 256);
    // rounds, per tile size (which fixes the workgroup size)
    std::vector<int> ks;
    for (const BatchInstance& in : b->inst) {
        ks.push_back(in.k);
    }
    std::sort(ks.begin(), ks.end());
    ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
    std::vector<unsigned char> tables;
    const auto put = [&](const void* src, size_t bytes) {
        const size_t off = tables.size();
        tables.resize(off + align256(std::max<size_t>(bytes, 1)), 0);
        std::memcpy(tables.data() + off, src, bytes);
        return off;
    };
    std::vector<std::vector<BatchTileItem>> round_items;
    std::vector<std::vector<uint64_t>> round_row_off;
    std::vector<std::vector<uint32_t>> round_row_cnt;
    b->part_doubles = 1;
    const auto add_rounds = [&](bool adjoint) {
        for (const int k : ks) {
            size_t n_rounds = 0;
            for (const BatchInstance& in : b->inst) {
                if (in.k == k) {
                    n_rounds = std::max(n_rounds, adjoint ? in.bwd_prog.size() : in.fwd_prog.size());
                }
            }
            for (size_t r = 0; r < n_rounds; ++r) {
                BatchRound rd;
                rd.adjoint = adjoint;
                rd.k = k;
                std::vector<BatchTileItem> items;
                std::vector<uint64_t> roff;
                std::vector<uint32_t> rcnt;
                uint32_t blocks = 0;
                uint64_t part = 0;
                for (uint32_t ii = 0; ii < b->inst.size(); ++ii) {
                    const BatchInstance& in = b->inst[ii];
                    const auto& progs = adjoint ? in.bwd_prog : in.fwd_prog;
                    if (in.k != k || r >= progs.size()) {
                        continue;
                    }
                    const TileProgram& tp = progs[r];
                    const size_t off = adjoint ? in.bwd_off[r] : in.fwd_off[r];
                    BatchTileItem it {};
                    it.a_off = in.amp_off;
                    it.b_off = in.amp_off;
                    // device addresses are filled in once the program arena exists (offsets here)
                    it.gates = off;
                    it.steps = off + align256(tp.gates.size() * sizeof(TileGate));
                    it.layouts = it.steps + align256(tp.steps.size() * sizeof(TileStep));
                    it.derivs = it.layouts + align256(tp.layouts.size() * sizeof(TileLayout));
                    it.terms = it.derivs + align256(std::max<size_t>(tp.derivs.size(), 1) * sizeof(TileGate));
                    it.free_mask = free_mask_of(tp);
                    it.n_thr_bits = tp.n_thr_bits;
                    it.n_steps = static_cast<int32_t>(tp.steps.size());
                    it.out_layout = tp.out_layout;
                    it.block_begin = blocks;
                    if (!adjoint && !b->reset) {
                        it.input = r == 0 ? 1u : (in.fwd_zero_mask[r] != 0 ? 2u : 0u);
                        it.zmask = in.fwd_zero_mask[r];
                    }
                    const uint32_t nb = tile_pass_blocks(tp);
                    it.n_part = nb * tile_pass_waves(tp);
                    it.partials = part;   // offset in doubles, made an address later
                    for (size_t sl = 0; sl < tp.derivs.size(); ++sl) {
                        roff.push_back(part + sl * it.n_part);
                        rcnt.push_back(it.n_part);
                        b->rows.push_back({ii, in.bwd_gate[tp.deriv_src[sl]], static_cast<uint32_t>(r), static_cast<uint32_t>(sl)});
                    }
                    part += tp.derivs.size() * it.n_part;
                    blocks += nb;
                    items.push_back(it);
                }
                rd.n_items = static_cast<uint32_t>(items.size());
                rd.n_blocks = blocks;
                rd.n_rows = static_cast<uint32_t>(roff.size());
                rd.row_base = b->rows.size() - roff.size();
                b->part_doubles = std::max<uint64_t>(b->part_doubles, part);
                b->rounds.push_back(rd);
                round_items.push_back(std::move(items));
                round_row_off.push_back(std::move(roff));
                round_row_cnt.push_back(std::move(rcnt));
            }
        }
    };
    add_rounds(false);
    add_rounds(true);
    // Pauli / read-out items and energy rows
    std::vector<BatchPauliItem> pitems;
    std::vector<uint64_t> e_off;
    std::vector<uint32_t> e_cnt;
    std::vector<uint64_t> ones;
    uint32_t pblocks = 0;
    for (const BatchInstance& in : b->inst) {
        BatchPauliItem it {};
        it.a_off = in.amp_off;
        it.b_off = in.amp_off;
        it.n_amps = in.n_amps;
        it.groups = in.pauli_off;
        it.zs = in.pauli_off + align256(in.groups.size() * sizeof(PauliGroup));
        it.n_groups = static_cast<uint32_t>(in.groups.size());
        it.block_begin = pblocks;
        it.part_off = pblocks;
        const uint32_t nb = batch_blocks(in.n_amps);
        e_off.push_back(pblocks);
        e_cnt.push_back(nb);
        pblocks += nb;
        pitems.push_back(it);
        ones.push_back(in.amp_off);
    }
    b->pauli_blocks = pblocks;
    const uint64_t part_total = b->part_doubles + pblocks;
    // device memory
    b->psi = balloc(b->total_amps * ab);
    if (b->fwd_precision != b->precision) {
        b->psi_fwd = balloc(b->total_amps * amp_bytes(b->fwd_precision));
    }
    b->lam = balloc(b->total_amps * ab);
    b->prog = balloc(b->prog_bytes);
    b->pauli = balloc(b->pauli_bytes);
    b->part = static_cast<double*>(balloc(part_total * sizeof(double)));
    b->slot_values = static_cast<double*>(balloc(std::max<size_t>(b->rows.size(), 1) * sizeof(double)));
    b->energies = static_cast<double*>(balloc(b->inst.size() * sizeof(double)));
    b->drift = static_cast<double*>(balloc(b->inst.size() * sizeof(double)));
    // fix up device addresses and write the tables
    const auto prog_base = reinterpret_cast<uint64_t>(b->prog);
    const auto part_base = reinterpret_cast<uint64_t>(b->part);
    for (size_t ri = 0; ri < b->rounds.size(); ++ri) {
        for (BatchTileItem& it : round_items[ri]) {
            it.gates += prog_base;
            it.steps += prog_base;
            it.layouts += prog_base;
            it.derivs += prog_base;
            it.terms += prog_base;
            it.partials = part_base + it.partials * sizeof(double);
        }
        b->rounds[ri].items_off = put(round_items[ri].data(), round_items[ri].size() * sizeof(BatchTileItem));
        b->rounds[ri].row_off_off = put(round_row_off[ri].data(), round_row_off[ri].size() * sizeof(uint64_t));
        b->rounds[ri].row_cnt_off = put(round_row_cnt[ri].data(), round_row_cnt[ri].size() * sizeof(uint32_t));
    }
    const auto pauli_base = reinterpret_cast<uint64_t>(b->pauli);
    for (BatchPauliItem& it : pitems) {
        it.groups += pauli_base;
        it.zs += pauli_base;
        it.part_off += b->part_doubles;   // energy partials follow the derivative partials
    }
    for (uint64_t& o : e_off) {
        o += b->part_doubles;
    }
    b->pauli_items_off = put(pitems.data(), pitems.size() * sizeof(BatchPauliItem));
    b->e_row_off_off = put(e_off.data(), e_off.size() * sizeof(uint64_t));
    b->e_row_cnt_off = put(e_cnt.data(), e_cnt.size() * sizeof(uint32_t));
    b->ones_off = put(ones.data(), ones.size() * sizeof(uint64_t));
    b->table_bytes = tables.size();
    b->tables = balloc(b->table_bytes);
    KQ_HIP_CHECK(hipMemcpyAsync(b->tables, tables.data(), tables.size(), hipMemcpyHostToDevice, b->stream));
    // Pauli tables
    std::vector<unsigned char> ph(b->pauli_bytes, 0);
    for (const BatchInstance& in : b->inst) {
        std::memcpy(ph.data() + in.pauli_off, in.groups.data(), in.groups.size() * sizeof(PauliGroup));
        std::memcpy(ph.data() + in.pauli_off + align256(in.groups.size() * sizeof(PauliGroup)), in.strings.data(),
                    in.strings.size() * sizeof(PauliZ));
    }
    KQ_HIP_CHECK(hipMemcpyAsync(b->pauli, ph.data(), ph.size(), hipMemcpyHostToDevice, b->stream));
    KQ_HIP_CHECK(hipStreamSynchronize(b->stream));
    KQ_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b->prog_host), std::max<size_t>(b->prog_bytes, 256), hipHostMallocDefault));
    std::memset(b->prog_host, 0, b->prog_bytes);
    b->planned = true;
    b->coeffs_dirty = true;
}

// Writes the pass programs of the instances whose parameters changed (every instance after a build)
// into the pinned staging arena and copies their byte ranges to the device.  Re-lowering runs on up to
// 16 host threads, one instance at a time each.
void upload_programs(kq_batch* b, bool relower)
{
    std::vector<uint32_t> todo;
    for (uint32_t i = 0; i < b->inst.size(); ++i) {
        if (!relower || b->inst[i].dirty) {
            todo.push_back(i);
        }
    }
    const auto one = [&](BatchInstance& in) {
        if (relower) {
            lower_instance(in, b->precision, b->fwd_precision);
        }
        const auto put = [&](const TileProgram& tp, size_t off, size_t end) {
            // new angles change a program's values and variants, never its size
            if (program_bytes(tp) != end - off) {
                throw InvalidArg {"batch: a re-lowered pass program changed size"};
            }
            check_tile_records(tp);
            write_program(tp, b->prog_host + off);
        };
        for (size_t p = 0; p < in.fwd_prog.size(); ++p) {
            put(in.fwd_prog[p], in.fwd_off[p], p + 1 < in.fwd_off.size() ? in.fwd_off[p + 1] : in.fwd_end);
        }
        for (size_t p = 0; p < in.bwd_prog.size(); ++p) {
            put(in.bwd_prog[p], in.bwd_off[p], p + 1 < in.bwd_off.size() ? in.bwd_off[p + 1] : in.bwd_end);
        }
    };
    std::atomic<size_t> next {0};
    std::exception_ptr err;
    std::mutex err_mu;
    const auto work = [&] {
        for (size_t j = next++; j < todo.size(); j = next++) {
            try {
                one(b->inst[todo[j]]);
            }
            catch (...) {
                std::lock_guard<std::mutex> lk(err_mu);
                if (!err) {
                    err = std::current_exception();
                }
                next = todo.size();
            }
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t workers = std::min<size_t>({todo.size() / 8 + 1, hw, 16});
    std::vector<std::thread> pool;
    for (size_t w = 1; w < workers; ++w) {
        pool.emplace_back(work);
    }
    work();
    for (auto& t : pool) {
        t.join();
    }
    if (err) {
        std::rethrow_exception(err);
    }
    // an instance's programs are contiguous: one copy per changed instance, or one for the arena
    if (todo.size() == b->inst.size()) {
        KQ_HIP_CHECK(hipMemcpyAsync(b->prog, b->prog_host, b->prog_bytes, hipMemcpyHostToDevice, b->stream));
    }
    else {
        for (const uint32_t i : todo) {
            const BatchInstance& in = b->inst[i];
            const size_t begin = !in.fwd_off.empty() ? in.fwd_off[0] : (!in.bwd_off.empty() ? in.bwd_off[0] : in.bwd_end);
            if (in.bwd_end > begin) {
                KQ_HIP_CHECK(hipMemcpyAsync(static_cast<unsigned char*>(b->prog) + begin, b->prog_host + begin,
                                            in.bwd_end - begin, hipMemcpyHostToDevice, b->stream));
            }
        }
    }
    for (const uint32_t i : todo) {
        b->inst[i].dirty = false;
    }
    b->coeffs_dirty = false;
}

// the launch sequence (no allocation, no synchronisation: capturable)
void enqueue(kq_batch* b)
{
    auto* tab = static_cast<unsigned char*>(b->tables);
    const uint32_t n = static_cast<uint32_t>(b->inst.size());
    size_t launches = 0;
    const bool mixed = b->fwd_precision != b->precision;
    void* fwd = mixed ? b->psi_fwd : b->psi;
    if (b->reset) {
        // (otherwise every state's first forward pass writes its |0...0> itself)
        launch_batch_reset(b->stream, b->fwd_precision, fwd, b->total_amps, reinterpret_cast<const uint64_t*>(tab + b->ones_off), n);
        launches += 2;
    }
    for (const BatchRound& rd : b->rounds) {
        if (rd.adjoint) {
            continue;
        }
        launch_tile_batch(b->stream, b->fwd_precision, false, fwd, nullptr,
                          reinterpret_cast<const BatchTileItem*>(tab + rd.items_off), rd.n_items, rd.n_blocks, rd.k);
        ++launches;
    }
    launch_pauli_batch(b->stream, b->fwd_precision, b->precision, fwd, b->lam,
                       reinterpret_cast<const BatchPauliItem*>(tab + b->pauli_items_off), n, b->pauli_blocks, b->part);
    launch_sum_rows(b->stream, b->part, reinterpret_cast<const uint64_t*>(tab + b->e_row_off_off),
                    reinterpret_cast<const uint32_t*>(tab + b->e_row_cnt_off), n, b->energies);
    launches += 2;
    if (mixed) {
        launch_narrow(b->stream, b->psi_fwd, b->psi, b->total_amps);
        ++launches;
    }
    for (const BatchRound& rd : b->rounds) {
        if (!rd.adjoint) {
            continue;
        }
        launch_tile_batch(b->stream, b->precision, true, b->psi, b->lam,
                          reinterpret_cast<const BatchTileItem*>(tab + rd.items_off), rd.n_items, rd.n_blocks, rd.k);
        launch_sum_rows(b->stream, b->part, reinterpret_cast<const uint64_t*>(tab + rd.row_off_off),
                        reinterpret_cast<const uint32_t*>(tab + rd.row_cnt_off), rd.n_rows, b->slot_values + rd.row_base);
        launches += rd.n_rows ? 2 : 1;
    }
    launch_basis_distance_batch(b->stream, b->precision, b->psi, reinterpret_cast<const BatchPauliItem*>(tab + b->pauli_items_off),
                                n, b->pauli_blocks, b->part);
    launch_max_rows(b->stream, b->part, reinterpret_cast<const uint64_t*>(tab + b->e_row_off_off),
                    reinterpret_cast<const uint32_t*>(tab + b->e_row_cnt_off), n, b->drift);
    launches += 2;
    b->n_launches = launches;
}

}  // namespace

void batch_add(kq_batch* b, uint32_t n_qubits, const kq_op* ops, size_t n_ops, size_t n_params, size_t n_strings,
               const uint32_t* off, const uint32_t* q, const uint8_t* k, const uint8_t* phases, const double* coeffs,
               size_t* index)
{
    if (n_qubits == 0 || n_qubits > 34) {
        throw InvalidArg {"batch: n_qubits must be 1..34"};
    }
    if (n_ops > 0 && ops == nullptr) {
        throw InvalidArg {"batch: null op array"};
    }
    BatchInstance in;
    in.n = n_qubits;
    in.nv = std::max<uint32_t>(n_qubits, 4);
    in.ops.assign(ops, ops + n_ops);
    in.n_params = n_params;
    in.params.assign(n_params, 0.0);
    std::vector<char> seen(n_params, 0);
    for (size_t i = 0; i < n_ops; ++i) {
        const kq_op& op = ops[i];
        if (op.gate < 0 || op.gate >= KQ_M) {
            throw InvalidArg {"batch: op " + std::to_string(i) + " is not a unitary gate"};
        }
        const bool ctrl = (op.gate >= KQ_CH && op.gate <= KQ_CP) || op.gate == KQ_CU;
        const int32_t t = ctrl ? op.q1 : op.q0;
        if (t < 0 || static_cast<uint32_t>(t) >= n_qubits || (ctrl && (op.q0 < 0 || static_cast<uint32_t>(op.q0) >= n_qubits))) {
            throw InvalidArg {"batch: gate qubit index out of range"};
        }
        if (op.aux >= 0) {
            if (static_cast<size_t>(op.aux) >= n_params) {
                throw InvalidArg {"batch: parameter index out of range"};
            }
            if (!seen[static_cast<size_t>(op.aux)]) {
                in.params[static_cast<size_t>(op.aux)] = op.angle;
                seen[static_cast<size_t>(op.aux)] = 1;
            }
        }
    }
    build_pauli(n_qubits, n_strings, off, q, k, phases, coeffs, in.groups, in.strings);
    if (n_strings > 0) {   // build_pauli has checked the arrays
        in.h_off.assign(off, off + n_strings + 1);
        in.h_q.assign(q, q + off[n_strings]);
        in.h_k.assign(k, k + off[n_strings]);
        in.h_ph.assign(phases, phases + n_strings);
        in.h_c.assign(coeffs, coeffs + 2 * n_strings);
    }
    b->inst.push_back(std::move(in));
    b->planned = false;
    if (index) {
        *index = b->inst.size() - 1;
    }
}

void batch_set_parameters(kq_batch* b, size_t index, const double* values)
{
    if (index >= b->inst.size()) {
        throw InvalidArg {"batch: instance index out of range"};
    }
    BatchInstance& in = b->inst[index];
    if (in.n_params > 0 && values == nullptr) {
        throw InvalidArg {"batch: null parameter values"};
    }
    in.params.assign(values, values + in.n_params);
    in.dirty = true;
    b->coeffs_dirty = true;
}

void batch_set_all_parameters(kq_batch* b, const double* values)
{
    size_t total = 0;
    for (const BatchInstance& in : b->inst) {
        total += in.n_params;
    }
    if (total > 0 && values == nullptr) {
        throw InvalidArg {"batch: null parameter values"};
    }
    size_t o = 0;
    for (BatchInstance& in : b->inst) {
        in.params.assign(values + o, values + o + in.n_params);
        in.dirty = true;
        o += in.n_params;
    }
    b->coeffs_dirty = true;
}

// energy and gradient of instance i by the single-state checkpointed adjoint in complex128, with a
// restore at every checkpoint (drift tolerance 0)
void recompute_instance(kq_batch* b, size_t i, double* energy, double* grad)
{
    const BatchInstance& in = b->inst[i];
    kq_state* st = nullptr;
    for (auto& w : b->work) {
        if (w.first == in.n) {
            st = w.second;
        }
    }
    if (st == nullptr) {
        st = create_state(in.n, KQ_C128, false);
        b->work.emplace_back(in.n, st);
    }
    st->bit_of.clear();   // the adjoint runs from |0...0> with qubit q on bit q
    launch_set_basis(st, 0);
    std::vector<kq_op> ops = in.ops;
    for (kq_op& op : ops) {
        if (op.aux >= 0) {
            op.angle = in.params[static_cast<size_t>(op.aux)];
        }
    }
    const size_t n_strings = in.h_ph.size();
    double e = 0.0;
    std::vector<double> g(std::max<size_t>(in.n_params, 1), 0.0);
    energy_gradient(st, ops.data(), ops.size(), n_strings, in.h_off.data(), in.h_q.data(), in.h_k.data(), in.h_ph.data(),
                    in.h_c.data(), in.n_params, &e, g.data(), 0, 0.0, nullptr);
    if (energy) {
        *energy = e;
    }
    if (grad) {
        std::copy(g.begin(), g.begin() + static_cast<std::ptrdiff_t>(in.n_params), grad);
    }
}

void batch_run(kq_batch* b, bool use_graph, double* energies, double* grads, double* drift)
{
    if (b->inst.empty()) {
        throw InvalidArg {"batch: no instances"};
    }
    const bool fresh = !b->planned;
    if (fresh) {
        build(b);
    }
    if (b->coeffs_dirty) {
        upload_programs(b, !fresh);
    }
    if (use_graph && b->graph_exec == nullptr && !fresh) {
        KQ_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue(b);
        }
        catch (...) {
            hipGraph_t g = nullptr;
            (void) hipStreamEndCapture(b->stream, &g);
            if (g) {
                (void) hipGraphDestroy(g);
            }
            throw;
        }
        KQ_HIP_CHECK(hipStreamEndCapture(b->stream, &b->graph));
        KQ_HIP_CHECK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    }
    if (use_graph && b->graph_exec != nullptr) {
        KQ_HIP_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
    }
    else {
        enqueue(b);   // first run eager: kernel attributes are set before any capture
    }
    const size_t n = b->inst.size();
    std::vector<double> e(n), dr(n), sv(b->rows.size());
    KQ_HIP_CHECK(hipMemcpyAsync(e.data(), b->energies, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    KQ_HIP_CHECK(hipMemcpyAsync(dr.data(), b->drift, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (!sv.empty()) {
        KQ_HIP_CHECK(hipMemcpyAsync(sv.data(), b->slot_values, sv.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    }
    KQ_HIP_CHECK(hipStreamSynchronize(b->stream));
    if (energies) {
        std::copy(e.begin(), e.end(), energies);
    }
    if (drift) {
        std::copy(dr.begin(), dr.end(), drift);
    }
    if (grads) {
        size_t total = 0;
        for (const BatchInstance& in : b->inst) {
            total += in.n_params;
        }
        std::fill(grads, grads + total, 0.0);
        // rows are in a fixed order, so the sums over gates sharing a parameter are deterministic
        for (size_t r = 0; r < b->rows.size(); ++r) {
            const BatchRow& row = b->rows[r];
            const BatchInstance& in = b->inst[row.inst];
            const kq_op& op = in.ops[row.gate];
            // the pass carries psi / F and lambda / F: the slot holds Re <lambda| Q |psi> / |F|^2
            grads[in.grad_off + static_cast<size_t>(op.aux)] += 2.0 * sv[r] * in.bwd_dscale[row.pass][row.slot];
        }
    }
    b->last_recomputed = 0;
    if (b->recompute_tol >= 0.0) {
        for (size_t i = 0; i < n; ++i) {
            if (!(dr[i] <= b->recompute_tol)) {   // NaN drift is recomputed too
                recompute_instance(b, i, energies ? energies + i : nullptr, grads ? grads + b->inst[i].grad_off : nullptr);
                ++b->last_recomputed;
            }
        }
        b->total_recomputed += b->last_recomputed;
    }
}

void batch_plan_info(const kq_batch* b, size_t* fwd_passes, size_t* rev_passes, double* algorithmic_bytes)
{
    if (!b->planned) {
        throw InvalidArg {"batch: not planned yet (run it once)"};
    }
    if (fwd_passes) {
        *fwd_passes = b->fwd_passes;
    }
    if (rev_passes) {
        *rev_passes = b->rev_passes;
    }
    if (algorithmic_bytes) {
        *algorithmic_bytes = b->algorithmic_bytes;
    }
}

void batch_set_drift_recompute(kq_batch* b, double drift_tol)
{
    if (std::isnan(drift_tol)) {
        throw InvalidArg {"batch: drift tolerance is NaN"};
    }
    b->recompute_tol = drift_tol;
}

void batch_recomputed(const kq_batch* b, size_t* last_run, size_t* total)
{
    if (last_run) {
        *last_run = b->last_recomputed;
    }
    if (total) {
        *total = b->total_recomputed;
    }
}

kq_batch* batch_create(int precision, int forward_precision, hipStream_t stream)
{
    if (forward_precision != precision && !(precision == KQ_C64 && forward_precision == KQ_C128)) {
        throw InvalidArg {"batch: the forward precision must equal the precision, or be complex128 with complex64"};
    }
    auto* b = new kq_batch;
    b->precision = precision;
    b->fwd_precision = forward_precision;
    b->stream = stream;
    return b;
}

void batch_destroy(kq_batch* b)
{
    if (b->stream) {
        (void) hipStreamSynchronize(b->stream);
    }
    release(b);
    for (auto& w : b->work) {
        (void) kq_state_destroy(w.second);
    }
    b->work.clear();
    if (b->stream) {
        (void) hipStreamDestroy(b->stream);
    }
    delete b;
}

void batch_info(const kq_batch* b, size_t* n_instances, size_t* n_params, size_t* device_bytes, size_t* n_launches)
{
    size_t np = 0;
    for (const BatchInstance& in : b->inst) {
        np += in.n_params;
    }
    if (n_instances) {
        *n_instances = b->inst.size();
    }
    if (n_params) {
        *n_params = np;
    }
    if (device_bytes) {
        const size_t fwd_bytes = b->fwd_precision != b->precision ? b->total_amps * amp_bytes(b->fwd_precision) : 0;
        *device_bytes = b->planned ? 2 * b->total_amps * amp_bytes(b->precision) + fwd_bytes + b->prog_bytes + b->table_bytes + b->pauli_bytes
                                         + (b->part_doubles + b->pauli_blocks + b->rows.size() + 2 * b->inst.size()) * sizeof(double)
                                   : 0;
    }
    if (n_launches) {
        *n_launches = b->n_launches;
    }
}

}  // namespace kq
