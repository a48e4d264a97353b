// Gate runs on a state: fusion planning per call (the plan tuner), qubit placement between passes and
// its restore, SWAP gates as relabels, and the streaming of planned passes to the device.
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

// ---- running straight-line gate batches -------------------------------------------------------

uint64_t sweep_amps(const GateRec& g, uint64_t n_amps)
{
    const bool ctrl = g.control >= 0;
    if (g.kind == GK_DIAG || g.kind == GK_DIAG1) {
        const bool skip0 = g.m[0] == 1.0 && g.m[1] == 0.0;
        return n_amps >> ((ctrl ? 1 : 0) + (skip0 ? 1 : 0));
    }
    return n_amps >> (ctrl ? 1 : 0);
}

void execute_passes(kq_state* s, const std::vector<TilePass>& passes);
// final_bits: the placement the last pass should store (see plan_passes); out_of_place: it may store
// into the state's spare buffer
void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of,
                  const std::vector<uint8_t>* final_bits = nullptr, bool out_of_place = false);
void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of, int low_c,
                  int max_stages, const std::vector<uint8_t>* final_bits = nullptr, bool out_of_place = false);
// tile settings of one gate run, chosen by the plan tuner
struct FusionChoice
{
    int low_c = 5;
    int max_stages = 12;
};
bool tuned(const kq_state* s, int max_k);
FusionChoice choose_fusion(const kq_state* s, const std::vector<GateRec>& recs, int max_k, const std::vector<uint8_t>& bit_of);

// Tables of the one-pass permutation that moves the qubit on index bit bit_of[q] to bit q.
PermuteTables permute_tables(const std::vector<uint8_t>& bit_of)
{
    const uint32_t nv = static_cast<uint32_t>(bit_of.size());
    if (nv < 12 || nv > 63) {
        throw InvalidArg {"index-bit permutation: 12 to 63 index bits"};
    }
    std::vector<uint8_t> qubit_at(nv);
    for (uint32_t q = 0; q < nv; ++q) {
        qubit_at[bit_of[q]] = static_cast<uint8_t>(q);
    }
    // tile: the source's six lowest bits and the six bits that land on the destination's lowest
    std::vector<int> tile;
    const auto add = [&](int b) {
        if (static_cast<int>(tile.size()) < 12 && std::find(tile.begin(), tile.end(), b) == tile.end()) {
            tile.push_back(b);
        }
    };
    for (int b = 0; b < 6; ++b) {
        add(b);
    }
    for (int d = 0; d < 6; ++d) {
        add(bit_of[static_cast<size_t>(d)]);
    }
    for (int b = 0; b < static_cast<int>(nv); ++b) {
        add(b);
    }
    PermuteTables t {};
    std::vector<int> local_of(nv, -1);
    for (int j = 0; j < 12; ++j) {
        local_of[static_cast<size_t>(tile[static_cast<size_t>(j)])] = j;
        t.src_of[j] = uint64_t {1} << tile[static_cast<size_t>(j)];
        t.dst_of[j] = uint64_t {1} << qubit_at[static_cast<size_t>(tile[static_cast<size_t>(j)])];
    }
    for (int b = 0; b < static_cast<int>(nv); ++b) {
        if (local_of[static_cast<size_t>(b)] < 0) {
            t.free_src[t.n_free] = static_cast<uint8_t>(b);
            t.free_dst[t.n_free] = qubit_at[static_cast<size_t>(b)];
            ++t.n_free;
        }
    }
    // layouts: six lanes on the given local bits, the other six local bits as two wave bits and four
    // registers
    const auto layout = [&](const int* lanes, uint8_t* thr, uint8_t* reg) {
        uint32_t used = 0;
        for (int j = 0; j < 6; ++j) {
            thr[j] = static_cast<uint8_t>(lanes[j]);
            used |= 1u << lanes[j];
        }
        int nt = 6, nr = 0;
        for (int l = 0; l < 12; ++l) {
            if (!((used >> l) & 1u)) {
                if (nt < 8) {
                    thr[nt++] = static_cast<uint8_t>(l);
                }
                else {
                    reg[nr++] = static_cast<uint8_t>(l);
                }
            }
        }
    };
    int ld_lanes[6], st_lanes[6];
    for (int j = 0; j < 6; ++j) {
        ld_lanes[j] = local_of[static_cast<size_t>(j)];
        st_lanes[j] = local_of[bit_of[static_cast<size_t>(j)]];
    }
    layout(ld_lanes, t.ld_thr, t.ld_reg);
    layout(st_lanes, t.st_thr, t.st_reg);
    return t;
}

// The state's spare buffer (the destination of out-of-place passes), allocated on first use; false if
// it cannot be allocated.
bool ensure_spare(kq_state* s)
{
    if (s->spare) {
        return true;
    }
    try {
        s->spare = device_alloc(s->n_amps * amp_bytes(s->precision));
    }
    catch (const AllocError&) {
        return false;
    }
    return true;
}

// Moves every qubit q of the state back on index bit q in one out-of-place pass (k_permute_bits) into
// the spare buffer, or a new buffer, which replaces the state's.  False (nothing done) if the state has
// fewer than 12 index bits or no second buffer can be had.
bool permute_out_of_place(kq_state* s)
{
    if (s->nvec < 12 || s->nvec > 63) {
        return false;
    }
    const PermuteTables t = permute_tables(s->bit_of);
    if (s->spare) {
        launch_permute_bits(s, s->amps, s->spare, t);
        std::swap(s->amps, s->spare);
        return true;
    }
    void* dst = nullptr;
    try {
        dst = device_alloc(s->n_amps * amp_bytes(s->precision));
    }
    catch (const AllocError&) {
        return false;
    }
    launch_permute_bits(s, s->amps, dst, t);
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    KQ_HIP_CHECK(hipFree(s->amps));
    s->amps = dst;
    return true;
}

// Stores every qubit q of the state back on index bit q, if gate passes left them permuted: one
// out-of-place pass when a second buffer fits, else gate-free in-place tile passes.
void materialise_basis(kq_state* s)
{
    if (s->lazy_basis >= 0) {
        const uint64_t index = static_cast<uint64_t>(s->lazy_basis);
        ensure_init();
        launch_set_basis(s, index);   // clears lazy_basis
        s->zero_mask = s->n_amps - 1;   // e_index: every other amplitude is 0
        s->zero_val = index;
    }
}

void normalise_layout(kq_state* s)
{
    materialise_basis(s);
    if (s->bit_of.empty()) {
        return;
    }
    s->zero_mask = 0;
    if (permute_out_of_place(s)) {
        s->bit_of.clear();
        PlanStats& st = last_stats();
        ++st.n_passes;
        return;
    }
    const int max_k = fusion_max_k(s->precision, s->nvec);
    const auto passes = plan_restore(s->bit_of, s->nvec, max_k, fusion_low_qubits(), register_qubits(s->precision));
    s->bit_of.clear();
    execute_passes(s, passes);
}

void normalise_layout(const kq_state* s) { normalise_layout(const_cast<kq_state*>(s)); }

PassInput KnownZeros::input_for(const TileProgram& tp, bool first, bool compiled, uint64_t& zmask, uint64_t& zval) const
{
    zmask = 0;
    zval = 0;
    if (!compiled) {
        return PassInput::load;
    }
    if (first && basis) {
        zval = val;   // the basis index
        return PassInput::basis;
    }
    uint64_t free_mask = 0;
    for (const uint8_t q : tp.freepos) {
        free_mask |= uint64_t {1} << q;
    }
    if ((mask & free_mask) == 0) {
        return PassInput::load;
    }
    zmask = mask & free_mask;
    zval = val & zmask;
    return PassInput::skip_zero;
}

void KnownZeros::after(const TileProgram& tp)
{
    uint64_t free_mask = 0;
    for (const uint8_t q : tp.freepos) {
        free_mask |= uint64_t {1} << q;
    }
    if (tp.out_freepos.empty()) {
        mask &= free_mask;
    }
    else {
        uint64_t m = 0, v = 0;
        for (size_t i = 0; i < tp.freepos.size(); ++i) {
            const uint64_t from = uint64_t {1} << tp.freepos[i], to = uint64_t {1} << tp.out_freepos[i];
            if (mask & from) {
                m |= to;
                v |= (val & from) ? to : 0;
            }
        }
        mask = m;
        val = v;
    }
    val &= mask;
    basis = false;
}

// kettle's SWAP gate is three controlled X gates, CX(a, b) CX(b, a) CX(a, b) (circuit.cpp:801-810).
// Where qubits may sit on any index bit, the exchange moves no amplitude: the triple is dropped and
// every later gate on a acts on b and vice versa.  Returns the other gates, relabelled; sigma[q] is
// the label that qubit q's amplitudes carry at the end (a permutation).
std::vector<GateRec> exchanges_as_relabels(const std::vector<GateRec>& recs, uint32_t nvec, std::vector<uint8_t>& sigma)
{
    sigma.resize(nvec);
    for (uint32_t q = 0; q < nvec; ++q) {
        sigma[q] = static_cast<uint8_t>(q);
    }
    const auto cx = [](const GateRec& g, int32_t c, uint32_t t) {
        return g.kind == GK_SWAP && g.control == c && g.target == t;
    };
    std::vector<GateRec> out;
    out.reserve(recs.size());
    std::vector<uint8_t> label = sigma;   // label[q]: the qubit that true qubit q's gates act on now
    for (size_t i = 0; i < recs.size(); ++i) {
        const GateRec& g = recs[i];
        if (g.kind == GK_SWAP && g.control >= 0 && i + 2 < recs.size()) {
            const uint32_t a = static_cast<uint32_t>(g.control), b = g.target;
            if (cx(recs[i + 1], static_cast<int32_t>(b), a) && cx(recs[i + 2], static_cast<int32_t>(a), b)) {
                std::swap(label[a], label[b]);
                i += 2;
                continue;
            }
        }
        GateRec r = g;
        r.target = label[g.target];
        if (g.control >= 0) {
            r.control = label[static_cast<uint32_t>(g.control)];
        }
        out.push_back(r);
    }
    sigma = label;
    return out;
}

// Whether the last pass of a statevector gate run may store out of place: only run-time compiled passes
// do, into the state's spare buffer (allocated here when it can be).
bool fold_out_of_place(kq_state* s, int max_k)
{
    return jit_wanted(s) && jit_available() && max_k < static_cast<int>(s->nvec) && ensure_spare(s);
}

void run_records(kq_state* s, const std::vector<GateRec>& recs, bool allow_relabel)
{
    PlanStats& st = last_stats();
    st.n_gates += recs.size();
    if (recs.empty()) {
        return;
    }
    const double ab = static_cast<double>(amp_bytes(s->precision));
    const int max_k = fusion_max_k(s->precision, s->nvec);
    // tiles smaller than one thread's register block (states of a few qubits) run gate by gate
    if (max_k < register_qubits(s->precision)) {
        normalise_layout(s);
        s->zero_mask = 0;   // (known zeros are followed through tile passes only)
        for (const auto& g : recs) {
            profiled(s, 2, 2.0 * ab * static_cast<double>(sweep_amps(g, s->n_amps)), 1, [&] { launch_gate_sweep(s, g); });
            ++st.n_passes;
        }
        KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
        profiler_collect();
        return;
    }
    // statevector passes may leave qubits on other index bits (the state's bit_of); everything
    // else, and density matrices, run with qubit q on bit q
    const bool relabel = allow_relabel && fusion_relabel() && !s->is_dm;
    if (!relabel && !s->bit_of.empty()) {
        normalise_layout(s);   // (a pending basis state keeps qubit q on bit q and stays pending)
    }
    std::vector<uint8_t> bit_of = s->bit_of;
    // a gate run leaves every qubit on its own bit when its last pass can store them there
    std::vector<uint8_t> identity_bits(s->nvec);
    for (uint32_t q = 0; q < s->nvec; ++q) {
        identity_bits[q] = static_cast<uint8_t>(q);
    }
    if (relabel) {
        std::vector<uint8_t> sigma;
        const std::vector<GateRec> rest = exchanges_as_relabels(recs, s->nvec, sigma);
        if (rest.size() != recs.size()) {
            // the planner works on the relabelled gates; qubit q ends on the bit of qubit sigma[q]
            if (bit_of.empty()) {
                bit_of.resize(s->nvec);
                for (uint32_t q = 0; q < s->nvec; ++q) {
                    bit_of[q] = static_cast<uint8_t>(q);
                }
            }
            if (!rest.empty()) {
                const FusionChoice fc = tuned(s, max_k) ? choose_fusion(s, rest, max_k, bit_of)
                                                        : FusionChoice {fusion_low_qubits(), fusion_max_stages()};
                // qubit q ends on the bit of the relabelled qubit sigma[q]: for q on bit q, that one's
                // final bit is q
                std::vector<uint8_t> final_bits(s->nvec);
                for (uint32_t q = 0; q < s->nvec; ++q) {
                    final_bits[sigma[q]] = static_cast<uint8_t>(q);
                }
                plan_and_run(s, rest, max_k, &bit_of, fc.low_c, fc.max_stages, &final_bits, fold_out_of_place(s, max_k));
            }
            std::vector<uint8_t> out(s->nvec);
            bool identity = true;
            for (uint32_t q = 0; q < s->nvec; ++q) {
                out[q] = bit_of[sigma[q]];
                identity = identity && out[q] == q;
            }
            s->bit_of = identity ? std::vector<uint8_t> {} : out;
            return;
        }
    }
    if (relabel && tuned(s, max_k)) {
        if (bit_of.empty()) {
            bit_of.resize(s->nvec);
            for (uint32_t q = 0; q < s->nvec; ++q) {
                bit_of[q] = static_cast<uint8_t>(q);
            }
        }
        const FusionChoice fc = choose_fusion(s, recs, max_k, bit_of);
        plan_and_run(s, recs, max_k, &bit_of, fc.low_c, fc.max_stages, &identity_bits, fold_out_of_place(s, max_k));
    }
    else if (relabel) {
        plan_and_run(s, recs, max_k, &bit_of, &identity_bits, fold_out_of_place(s, max_k));
    }
    else {
        plan_and_run(s, recs, max_k, nullptr);
    }
    if (relabel) {
        bool identity = true;
        for (size_t q = 0; q < bit_of.size(); ++q) {
            identity = identity && bit_of[q] == q;
        }
        s->bit_of = identity ? std::vector<uint8_t> {} : bit_of;
    }
}

size_t align256(size_t x) { return (x + 255) & ~size_t {255}; }

// Streams passes to the device as the planner produces them: each pass is lowered, its compiled
// kernel looked up, its program (coefficients, or the interpreted kernel's tables) staged in pinned
// host memory, copied and the pass launched, while the GPU runs the passes before it and the host
// plans the next one.  A pass whose kernel is not compiled yet is held back, with every later pass
// (launch order is execution order), and the held kernels compile together, in parallel, once the
// plan is complete.  Programs take consecutive regions of the state's device and pinned buffers; a
// region is reused only after the stream has drained.
class PassStream
{
public:
    explicit PassStream(kq_state* s) : s_(s), use_jit_(jit_wanted(s) && jit_available())
    {
        // the known zeros the passes start from, followed as they are pushed (launch order)
        if (s->lazy_basis >= 0) {
            kz_.mask = s->n_amps - 1;
            kz_.val = static_cast<uint64_t>(s->lazy_basis);
            kz_.basis = true;
        }
        else {
            kz_.mask = s->zero_mask;
            kz_.val = s->zero_val & s->zero_mask;
        }
    }

    void push(TilePass&& p)
    {
        Item it;
        it.pass = std::move(p);
        const bool first = first_;
        first_ = false;
        if (!it.pass.sweep) {
            lower_tile_pass(it.pass, s_->precision, s_->nvec, it.prog);
            // the run's first pass on a pending basis state builds its tiles from the basis index; a pass
            // whose outside bits include known ones skips (or zero-fills) the zero tiles
            it.input = kz_.input_for(it.prog, first, use_jit_ && jit_tuning().blocks_per_cu == 0, it.zmask, it.zval);
            if (use_jit_) {
                it.src = jit_source(it.prog, s_->precision, it.coefs, it.input);
                it.fn = jit_lookup(it.src);
            }
            kz_.after(it.prog);
        }
        else {
            kz_.forget();   // (sweeps: not followed)
        }
        it.mask_after = kz_.mask;
        it.val_after = kz_.val;
        if (held_.empty() && (it.pass.sweep || !use_jit_ || it.fn)) {
            launch(it);
        }
        else {
            held_.push_back(std::move(it));
        }
    }

    // device work (copies, small kernels) that must run after every pass pushed so far and before the
    // passes pushed later, whether or not those passes wait for their kernels to compile
    void push_action(std::function<void()> action)
    {
        first_ = false;
        kz_.forget();   // an action may change the state in any way
        Item it;
        it.action = std::move(action);
        it.mask_after = 0;
        it.val_after = 0;
        if (held_.empty()) {
            launch(it);   // (writes a pending basis state first)
        }
        else {
            held_.push_back(std::move(it));
        }
    }

    void finish(bool sync = true)
    {
        if (!held_.empty()) {
            std::vector<std::string> srcs;
            std::vector<size_t> idx;
            for (size_t i = 0; i < held_.size(); ++i) {
                if (!held_[i].action && !held_[i].pass.sweep && !held_[i].fn) {
                    srcs.push_back(held_[i].src);
                    idx.push_back(i);
                }
            }
            const auto fns = jit_kernels(srcs);
            for (size_t j = 0; j < idx.size(); ++j) {
                held_[idx[j]].fn = fns[j];   // null (compile failure): the interpreted tile kernel runs it
            }
            for (Item& it : held_) {
                launch(it);
            }
            held_.clear();
        }
        if (sync) {
            KQ_HIP_CHECK(hipStreamSynchronize(s_->stream));
            profiler_collect();
        }
    }

private:
    struct Item
    {
        TilePass pass;
        TileProgram prog;
        std::string src;
        std::vector<double> coefs;
        hipFunction_t fn = nullptr;
        PassInput input = PassInput::load;   // what its compiled kernel starts from
        uint64_t zmask = 0, zval = 0;        // skip_zero: the known outside bits
        uint64_t mask_after = 0, val_after = 0;   // the known zeros once it has run
        std::function<void()> action;        // set: not a pass, run this instead
    };

    // device and pinned staging regions of `bytes` for one program
    std::pair<unsigned char*, unsigned char*> reserve(size_t bytes)
    {
        bytes = align256(std::max<size_t>(bytes, 1));
        if (used_ + bytes > s_->prog_bytes || used_ + bytes > s_->stage_bytes) {
            if (used_ > 0) {
                KQ_HIP_CHECK(hipStreamSynchronize(s_->stream));   // every earlier program has been read
                used_ = 0;
            }
            const size_t want = std::max<size_t>({bytes, size_t {1} << 20, 2 * s_->prog_bytes});
            if (bytes > s_->prog_bytes) {
                grow(s_->prog, s_->prog_bytes, want);
            }
            if (bytes > s_->stage_bytes) {
                if (s_->stage) {
                    KQ_HIP_CHECK(hipHostFree(s_->stage));
                    s_->stage = nullptr;
                    s_->stage_bytes = 0;
                }
                KQ_HIP_CHECK(hipHostMalloc(&s_->stage, want, hipHostMallocDefault));
                s_->stage_bytes = want;
            }
        }
        auto* d = static_cast<unsigned char*>(s_->prog) + used_;
        auto* h = static_cast<unsigned char*>(s_->stage) + used_;
        used_ += bytes;
        return {d, h};
    }

    void launch(Item& it)
    {
        if (s_->lazy_basis >= 0 && !(it.input == PassInput::basis && it.fn)) {
            materialise_basis(s_);   // anything but a pass built from the basis index reads the buffer
        }
        s_->zero_mask = 0;   // (nothing is known if the launch fails)
        launch_item(it);
        s_->zero_mask = it.mask_after;
        s_->zero_val = it.val_after;
    }

    void launch_item(Item& it)
    {
        if (it.action) {
            it.action();
            return;
        }
        PlanStats& st = last_stats();
        const double ab = static_cast<double>(amp_bytes(s_->precision));
        const TilePass& p = it.pass;
        ++st.n_passes;
        if (p.sweep) {
            profiled(s_, 2, 2.0 * ab * static_cast<double>(sweep_amps(p.single, s_->n_amps)), 1,
                     [&] { launch_gate_sweep(s_, p.single); });
            return;
        }
        ++st.n_tile_passes;
        const TileProgram& tp = it.prog;
        const int cls = p.c >= p.k ? 0 : 1;
        if (it.fn) {
            const size_t bytes = std::max<size_t>(it.coefs.size(), 1) * sizeof(double);
            auto [d, h] = reserve(bytes);
            std::memcpy(h, it.coefs.data(), it.coefs.size() * sizeof(double));
            KQ_HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s_->stream));
            if (it.input == PassInput::basis) {
                // no read: only the store moves bytes
                const uint64_t basis = static_cast<uint64_t>(s_->lazy_basis);
                profiled(s_, cls, ab * static_cast<double>(s_->n_amps), p.gates.size(),
                         [&] { launch_jit_pass(s_, it.fn, tp, reinterpret_cast<const double*>(d), &basis, 1); });
                s_->lazy_basis = -1;   // written by the pass just launched
            }
            else if (it.input == PassInput::skip_zero) {
                // the tiles that may be non-zero are read and written; out of place, the others are written
                const uint64_t tiles = tile_pass_blocks(tp);
                const uint64_t live = tiles >> std::popcount(it.zmask);
                const double tile_bytes = ab * static_cast<double>(uint64_t {1} << tp.k);
                const double moved = tile_bytes * (2.0 * static_cast<double>(live)
                                                   + (tp.out_freepos.empty() ? 0.0 : static_cast<double>(tiles - live)));
                const uint64_t zx[2] = {it.zmask, it.zval};
                profiled(s_, cls, moved, p.gates.size(),
                         [&] { launch_jit_pass(s_, it.fn, tp, reinterpret_cast<const double*>(d), zx, 2); });
            }
            else {
                profiled(s_, cls, 2.0 * ab * static_cast<double>(s_->n_amps), p.gates.size(),
                         [&] { launch_jit_pass(s_, it.fn, tp, reinterpret_cast<const double*>(d)); });
            }
            if (!tp.out_freepos.empty()) {
                std::swap(s_->amps, s_->spare);   // the pass stored the state into the spare buffer
            }
            return;
        }
        if (!tp.out_freepos.empty()) {
            // no compiled kernel for an out-of-place pass: the interpreted kernel runs it in place without
            // moving qubits, then one permutation pass stores every index bit b at its destination
            TilePass in_place = p;
            std::vector<uint8_t> src_of(s_->nvec);   // destination bit -> source bit
            for (size_t j = 0; j < p.qubits.size(); ++j) {
                src_of[static_cast<size_t>(p.out_qubits[j])] = static_cast<uint8_t>(p.qubits[j]);
            }
            for (size_t i = 0; i < tp.freepos.size(); ++i) {
                src_of[tp.out_freepos[i]] = tp.freepos[i];
            }
            in_place.out_qubits.clear();
            in_place.out_free.clear();
            Item plain;
            plain.pass = std::move(in_place);
            lower_tile_pass(plain.pass, s_->precision, s_->nvec, plain.prog);
            --st.n_passes;
            --st.n_tile_passes;
            launch(plain);
            ++st.n_passes;
            launch_permute_bits(s_, s_->amps, s_->spare, permute_tables(src_of));
            std::swap(s_->amps, s_->spare);
            return;
        }
        const size_t g_b = align256(tp.gates.size() * sizeof(TileGate)), s_b = align256(tp.steps.size() * sizeof(TileStep)),
                     l_b = align256(tp.layouts.size() * sizeof(TileLayout));
        const size_t bytes = g_b + s_b + l_b + std::max<size_t>(tp.terms.size(), 1) * sizeof(TileTerm);
        auto [d, h] = reserve(bytes);
        std::memcpy(h, tp.gates.data(), tp.gates.size() * sizeof(TileGate));
        std::memcpy(h + g_b, tp.steps.data(), tp.steps.size() * sizeof(TileStep));
        std::memcpy(h + g_b + s_b, tp.layouts.data(), tp.layouts.size() * sizeof(TileLayout));
        std::memcpy(h + g_b + s_b + l_b, tp.terms.data(), tp.terms.size() * sizeof(TileTerm));
        KQ_HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s_->stream));
        profiled(s_, cls, 2.0 * ab * static_cast<double>(s_->n_amps), p.gates.size(),
                 [&] { launch_tile_pass(s_, tp, d, d + g_b, d + g_b + s_b, d + g_b + s_b + l_b); });
    }

    kq_state* s_;
    bool use_jit_;
    bool first_ = true;
    KnownZeros kz_;   // known zeros after the passes pushed so far
    std::vector<Item> held_;
    size_t used_ = 0;
};

void execute_passes(kq_state* s, const std::vector<TilePass>& passes)
{
    PassStream stream(s);
    for (const TilePass& p : passes) {
        stream.push(TilePass {p});
    }
    stream.finish();
}

namespace
{

// Plans of plan_statevector_passes by gate structure (kinds, bits; not the coefficients) and the
// settings they were planned with: a repeated structure (an optimiser's next angles) takes the cached
// passes with its own coefficients.
struct PlanCache
{
    struct Entry
    {
        std::vector<uint32_t> key;
        std::vector<TilePass> passes;
        std::vector<uint8_t> bit_of;
    };
    std::mutex mu;
    std::unordered_map<uint64_t, Entry> entries;
};

PlanCache& plan_cache()
{
    static PlanCache c;
    return c;
}

}  // namespace

std::vector<TilePass> plan_statevector_passes(kq_state* s, const std::vector<GateRec>& recs, bool allow_sweep,
                                              std::vector<uint8_t>& bit_of)
{
    const int max_k = fusion_max_k(s->precision, s->nvec);
    bit_of.resize(s->nvec);
    for (uint32_t q = 0; q < s->nvec; ++q) {
        bit_of[q] = static_cast<uint8_t>(q);
    }
    const bool relabel = fusion_relabel() && !s->is_dm && max_k < static_cast<int>(s->nvec);
    const bool tuner = relabel && tuned(s, max_k);
    std::vector<uint32_t> key {s->nvec, static_cast<uint32_t>(s->precision), static_cast<uint32_t>(max_k),
                               static_cast<uint32_t>(fusion_low_qubits()), static_cast<uint32_t>(fusion_max_stages()),
                               static_cast<uint32_t>(fusion_max_gates()), relabel ? 1u : 0u, tuner ? 1u : 0u,
                               allow_sweep ? 1u : 0u};
    key.reserve(key.size() + 3 * recs.size());
    for (const GateRec& g : recs) {
        key.push_back(g.kind);
        key.push_back(g.target);
        key.push_back(static_cast<uint32_t>(g.control));
    }
    uint64_t h = 1469598103934665603ull;
    for (const uint32_t v : key) {
        h = (h ^ v) * 1099511628211ull;
    }
    PlanCache& c = plan_cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        const auto it = c.entries.find(h);
        if (it != c.entries.end() && it->second.key == key) {
            std::vector<TilePass> passes = it->second.passes;
            for (TilePass& p : passes) {
                for (size_t i = 0; i < p.gates.size(); ++i) {
                    std::memcpy(p.gates[i].m, recs[p.src[i]].m, sizeof p.gates[i].m);
                }
                if (p.sweep) {
                    std::memcpy(p.single.m, recs[p.src[0]].m, sizeof p.single.m);
                }
            }
            bit_of = it->second.bit_of;
            return passes;
        }
    }
    const FusionChoice fc = tuner ? choose_fusion(s, recs, max_k, bit_of) : FusionChoice {fusion_low_qubits(), fusion_max_stages()};
    std::vector<TilePass> passes = plan_passes(recs, s->nvec, max_k, fc.low_c, fusion_max_gates(), register_qubits(s->precision),
                                               fc.max_stages, allow_sweep, relabel ? &bit_of : nullptr);
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.entries.size() >= 64) {
        c.entries.clear();
    }
    c.entries[h] = PlanCache::Entry {std::move(key), passes, bit_of};
    return passes;
}

void execute_passes_with(kq_state* s, const std::vector<TilePass>& passes,
                         const std::vector<std::pair<size_t, std::function<void()>>>& after, bool sync)
{
    PassStream stream(s);
    size_t a = 0;
    for (size_t i = 0; i < passes.size(); ++i) {
        stream.push(TilePass {passes[i]});
        for (; a < after.size() && after[a].first == i; ++a) {
            stream.push_action(after[a].second);
        }
    }
    stream.finish(sync);
}

// Plans a gate list and runs its passes as they are planned.
void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of, int low_c,
                  int max_stages, const std::vector<uint8_t>* final_bits, bool out_of_place)
{
    PassStream stream(s);
    const std::function<void(TilePass&&)> emit = [&](TilePass&& p) { stream.push(std::move(p)); };
    plan_passes(recs, s->nvec, max_k, low_c, fusion_max_gates(), register_qubits(s->precision), max_stages, true, bit_of, &emit,
                final_bits, out_of_place);
    stream.finish();
}

void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of,
                  const std::vector<uint8_t>* final_bits, bool out_of_place)
{
    plan_and_run(s, recs, max_k, bit_of, fusion_low_qubits(), fusion_max_stages(), final_bits, out_of_place);
}

// ---- plan tuner ------------------------------------------------------------------------------
// The greedy planner's pass count moves a lot with the stage cap and the number of low qubits per
// tile (30-qubit QFT + layers: 13 to 23 passes over the grid below) while the total number of LDS
// remaps stays near 150-175, so the tuner plans a statevector gate list of a compiled-pass state
// once per candidate (host only) and keeps the cheapest by a model fitted on MI355X: a pass costs
// its HBM time, plus 0.126 of it for each LDS remap above six (the remaps and the arithmetic beyond
// that no longer overlap the HBM traffic; measured within 2% over six plans).  The choice is cached
// per gate structure (kinds, bits, controls; not the angles), starting placement and tile size.
struct TunerState
{
    std::mutex mu;
    std::unordered_map<uint64_t, FusionChoice> cache;
    FusionChoice last {};
    size_t plans_tried = 0;
};

TunerState& tuner()
{
    static TunerState t;
    return t;
}

double plan_cost(const std::vector<TilePass>& passes)
{
    double c = 0.0;
    for (const TilePass& p : passes) {
        c += p.sweep ? 0.5 : 1.0 + 0.126 * std::max(0, pass_remaps(p) - 6);
    }
    return c;
}

FusionChoice choose_fusion(const kq_state* s, const std::vector<GateRec>& recs, int max_k, const std::vector<uint8_t>& bit_of)
{
    uint64_t h = 1469598103934665603ull;
    const auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    };
    mix(s->nvec);
    mix(static_cast<uint64_t>(s->precision));
    mix(static_cast<uint64_t>(max_k));
    mix(static_cast<uint64_t>(fusion_max_gates()));
    for (const uint8_t b : bit_of) {
        mix(b);
    }
    for (const GateRec& g : recs) {
        mix((uint64_t {g.kind} << 40) ^ (uint64_t {g.target} << 20) ^ static_cast<uint64_t>(static_cast<int64_t>(g.control) & 0xfffff));
    }
    TunerState& t = tuner();
    {
        std::lock_guard<std::mutex> lk(t.mu);
        auto it = t.cache.find(h);
        if (it != t.cache.end()) {
            t.last = it->second;
            return it->second;
        }
    }
    FusionChoice best {fusion_low_qubits(), fusion_max_stages()};
    double best_cost = 1.0e300;
    size_t tried = 0;
    for (const int low : {4, 5, 6, 7}) {
        if (low >= max_k || max_k - low > 10) {
            continue;
        }
        for (const int st : {12, 14, 16, 18, 20, 24, 28}) {
            std::vector<uint8_t> b = bit_of;
            try {
                const double c = plan_cost(plan_passes(recs, s->nvec, max_k, low, fusion_max_gates(), register_qubits(s->precision),
                                                       st, true, &b));
                ++tried;
                if (c < best_cost - 1e-9) {
                    best_cost = c;
                    best = {low, st};
                }
            }
            catch (const InvalidArg&) {
                // a candidate the planner cannot build (tile shape) is skipped
            }
        }
    }
    std::lock_guard<std::mutex> lk(t.mu);
    if (t.cache.size() >= 256) {
        t.cache.clear();
    }
    t.cache.emplace(h, best);
    t.last = best;
    t.plans_tried += tried;
    return best;
}

// the tuner applies to statevector runs with qubit moves through the compiled pass kernels
bool tuned(const kq_state* s, int max_k)
{
    return fusion_autotune() && !s->is_dm && jit_wanted(s) && jit_available() && max_k >= 8
        && max_k < static_cast<int>(s->nvec);
}

void run_ops(kq_state* s, const kq_op* ops, size_t n)
{
    std::vector<GateRec> recs;
    recs.reserve(s->is_dm ? 2 * n : n);
    for (size_t i = 0; i < n; ++i) {
        if (ops[i].gate < 0 || ops[i].gate >= KQ_M) {
            throw InvalidArg {"kq_apply_gates: op " + std::to_string(i) + " is not a unitary gate"};
        }
        compile_gate(ops[i], s->n, s->is_dm != 0, recs);
    }
    run_records(s, recs);
}

}  // namespace kq

using namespace kq;

extern "C" {

int kq_apply_gates(kq_state* s, const kq_op* ops, size_t n_ops)
{
    return guarded([&] {
        check_state_lazy(s);
        if (n_ops > 0 && ops == nullptr) {
            throw InvalidArg {"null op array"};
        }
        ensure_init();
        last_stats() = PlanStats {};
        run_ops(s, ops, n_ops);
    });
}

int kq_set_fusion(int max_tile_qubits, int max_gates_per_pass, int max_stages_per_pass, double max_cost_per_pass)
{
    return guarded([&] {
        if (max_tile_qubits > 13) {
            throw InvalidArg {"tile size above 2^13 amplitudes does not fit LDS"};
        }
        set_fusion(max_tile_qubits, max_gates_per_pass, max_stages_per_pass, max_cost_per_pass);
    });
}

int kq_get_fusion(int* max_tile_qubits, int* max_gates_per_pass, int* max_stages_per_pass, double* max_cost_per_pass)
{
    return guarded([&] {
        if (max_tile_qubits) { *max_tile_qubits = fusion_setting_k(); }
        if (max_gates_per_pass) { *max_gates_per_pass = fusion_max_gates(); }
        if (max_stages_per_pass) { *max_stages_per_pass = fusion_max_stages(); }
        if (max_cost_per_pass) { *max_cost_per_pass = fusion_max_cost(); }
    });
}

int kq_last_plan_stats(size_t* n_gates, size_t* n_passes, size_t* n_tile_passes)
{
    return guarded([&] {
        const auto& st = last_stats();
        if (n_gates) { *n_gates = st.n_gates; }
        if (n_passes) { *n_passes = st.n_passes; }
        if (n_tile_passes) { *n_tile_passes = st.n_tile_passes; }
    });
}

int kq_set_fusion_relabel(int on)
{
    return guarded([&] { set_fusion_relabel(on != 0); });
}

int kq_set_fusion_autotune(int on)
{
    return guarded([&] { set_fusion_autotune(on != 0); });
}

int kq_fusion_autotune_info(int* on, int* low_qubits, int* max_stages, size_t* plans_tried, size_t* cached)
{
    return guarded([&] {
        TunerState& t = tuner();
        std::lock_guard<std::mutex> lk(t.mu);
        if (on) {
            *on = fusion_autotune() ? 1 : 0;
        }
        if (low_qubits) {
            *low_qubits = t.last.low_c;
        }
        if (max_stages) {
            *max_stages = t.last.max_stages;
        }
        if (plans_tried) {
            *plans_tried = t.plans_tried;
        }
        if (cached) {
            *cached = t.cache.size();
        }
    });
}

int kq_set_fusion_low_qubits(int low_qubits)
{
    return guarded([&] { set_fusion_low_qubits(low_qubits); });
}

}  // extern "C"
