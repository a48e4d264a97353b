// The C-ABI entry points of libkq (include/kq.h) and their host-side logic: state handles, the
// simulate() interpreter, measurement draws, sampling, Pauli canonicalisation and the adjoint
// energy gradient.  Device work goes through the launchers in kq_kernels.hip.
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

namespace kq
{
namespace
{

thread_local std::string g_last_error;
std::mutex g_init_mutex;
bool g_initialised = false;
int g_device = 0;

template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return KQ_OK;
    }
    catch (const InvalidArg& e) {
        g_last_error = e.what();
        return KQ_ERR_INVALID;
    }
    catch (const HipError& e) {
        g_last_error = e.what();
        return KQ_ERR_HIP;
    }
    catch (const AllocError& e) {
        g_last_error = e.what();
        return KQ_ERR_ALLOC;
    }
    catch (const std::out_of_range& e) {
        g_last_error = e.what();
        return KQ_ERR_OUT_OF_RANGE;
    }
    catch (const std::exception& e) {
        g_last_error = e.what();
        return KQ_ERR_RUNTIME;
    }
    catch (...) {
        g_last_error = "unknown exception";
        return KQ_ERR_RUNTIME;
    }
}

void ensure_init()
{
    std::lock_guard<std::mutex> lock {g_init_mutex};
    if (!g_initialised) {
        int count = 0;
        KQ_HIP_CHECK(hipGetDeviceCount(&count));
        if (count <= 0) {
            throw HipError {"no HIP device available"};
        }
        KQ_HIP_CHECK(hipSetDevice(g_device));
        hipDeviceProp_t prop;
        KQ_HIP_CHECK(hipGetDeviceProperties(&prop, g_device));
        device_lds_per_block() = prop.sharedMemPerBlock;
        device_cus() = static_cast<uint32_t>(prop.multiProcessorCount);
        g_initialised = true;
    }
    else {
        KQ_HIP_CHECK(hipSetDevice(g_device));
    }
}

void check_state(const kq_state* s)
{
    if (s == nullptr || s->amps == nullptr) {
        throw InvalidArg {"null or destroyed state handle"};
    }
}

void* device_alloc(size_t bytes)
{
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess || p == nullptr) {
        (void) hipGetLastError();
        throw AllocError {"hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e)};
    }
    return p;
}

void* grow(void*& buf, size_t& have, size_t need)
{
    if (need > have) {
        if (buf != nullptr) {
            KQ_HIP_CHECK(hipFree(buf));
            buf = nullptr;
            have = 0;
        }
        const size_t bytes = std::max<size_t>(need, 1 << 16);
        buf = device_alloc(bytes);
        have = bytes;
    }
    return buf;
}

kq_state* create_state(uint32_t n, int precision, bool dm)
{
    if (precision != KQ_C128 && precision != KQ_C64) {
        throw InvalidArg {"precision must be KQ_C128 or KQ_C64"};
    }
    if (n == 0) {
        throw std::runtime_error {dm ? "ERROR: cannot perform a DensityMatrix simulation with 0 qubits.\n"
                                     : "There must be at least 1 qubit in the Statevector.\n"};
    }
    const uint32_t nvec = dm ? 2 * n : n;
    if (nvec > 36) {
        throw InvalidArg {"state too large for one device"};
    }
    ensure_init();
    auto* s = new kq_state {};
    s->n = n;
    s->nvec = nvec;
    s->precision = precision;
    s->is_dm = dm ? 1 : 0;
    s->n_amps = uint64_t {1} << nvec;
    try {
        s->amps = device_alloc(s->n_amps * amp_bytes(precision));
        KQ_HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        launch_set_basis(s, 0);
        KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    }
    catch (...) {
        if (s->amps) {
            (void) hipFree(s->amps);
        }
        if (s->stream) {
            (void) hipStreamDestroy(s->stream);
        }
        delete s;
        throw;
    }
    return s;
}

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
void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of);
void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of, int low_c,
                  int max_stages);
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

// Moves every qubit q of the state back on index bit q in one out-of-place pass (k_permute_bits) into
// a new buffer that replaces the state's.  False (nothing done) if the state has fewer than 12 index
// bits or the second buffer cannot be allocated.
bool permute_out_of_place(kq_state* s)
{
    if (s->nvec < 12 || s->nvec > 63) {
        return false;
    }
    const PermuteTables t = permute_tables(s->bit_of);
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
void normalise_layout(kq_state* s)
{
    if (s->bit_of.empty()) {
        return;
    }
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

void run_records(kq_state* s, const std::vector<GateRec>& recs, bool allow_relabel = true)
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
    if (!relabel) {
        normalise_layout(s);
    }
    std::vector<uint8_t> bit_of = s->bit_of;
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
                plan_and_run(s, rest, max_k, &bit_of, fc.low_c, fc.max_stages);
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
        plan_and_run(s, recs, max_k, &bit_of, fc.low_c, fc.max_stages);
    }
    else {
        plan_and_run(s, recs, max_k, relabel ? &bit_of : nullptr);
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
    explicit PassStream(kq_state* s) : s_(s), use_jit_(jit_wanted(s) && jit_available()) {}

    void push(TilePass&& p)
    {
        Item it;
        it.pass = std::move(p);
        if (!it.pass.sweep) {
            lower_tile_pass(it.pass, s_->precision, s_->nvec, it.prog);
            if (use_jit_) {
                it.src = jit_source(it.prog, s_->precision, it.coefs);
                it.fn = jit_lookup(it.src);
            }
        }
        if (held_.empty() && (it.pass.sweep || !use_jit_ || it.fn)) {
            launch(it);
        }
        else {
            held_.push_back(std::move(it));
        }
    }

    void finish()
    {
        if (!held_.empty()) {
            std::vector<std::string> srcs;
            std::vector<size_t> idx;
            for (size_t i = 0; i < held_.size(); ++i) {
                if (!held_[i].pass.sweep && !held_[i].fn) {
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
        KQ_HIP_CHECK(hipStreamSynchronize(s_->stream));
        profiler_collect();
    }

private:
    struct Item
    {
        TilePass pass;
        TileProgram prog;
        std::string src;
        std::vector<double> coefs;
        hipFunction_t fn = nullptr;
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
            profiled(s_, cls, 2.0 * ab * static_cast<double>(s_->n_amps), p.gates.size(),
                     [&] { launch_jit_pass(s_, it.fn, tp, reinterpret_cast<const double*>(d)); });
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

// Plans a gate list and runs its passes as they are planned.
void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of, int low_c,
                  int max_stages)
{
    PassStream stream(s);
    const std::function<void(TilePass&&)> emit = [&](TilePass&& p) { stream.push(std::move(p)); };
    plan_passes(recs, s->nvec, max_k, low_c, fusion_max_gates(), register_qubits(s->precision), max_stages, true, bit_of, &emit);
    stream.finish();
}

void plan_and_run(kq_state* s, const std::vector<GateRec>& recs, int max_k, std::vector<uint8_t>* bit_of)
{
    plan_and_run(s, recs, max_k, bit_of, fusion_low_qubits(), fusion_max_stages());
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
void sample_from_device_probs(hipStream_t st, double* d_p, uint64_t n, uint64_t shots, bool has_seed, int32_t seed,
                              uint64_t* out)
{
    const size_t sb = cdf_scratch_bytes(n);
    void* sc = device_alloc(sb + shots * (sizeof(double) + sizeof(uint64_t)) + 256);
    try {
        double last = 0.0, second_last = 0.0;
        launch_cdf(st, d_p, n, sc, sb, &last, &second_last);
        const double offset = (last - second_last) * 1.0e-4;
        auto prng = get_prng(has_seed, seed);
        auto dist = std::uniform_real_distribution<double> {0.0, last - offset};
        std::vector<double> u(shots);
        for (auto& x : u) {
            x = dist(prng);
        }
        auto* d_u = reinterpret_cast<double*>(static_cast<char*>(sc) + ((sb + 127) / 128) * 128);
        auto* d_out = reinterpret_cast<uint64_t*>(d_u + shots);
        if (shots) {
            KQ_HIP_CHECK(hipMemcpyAsync(d_u, u.data(), shots * sizeof(double), hipMemcpyHostToDevice, st));
            launch_lower_bound(st, d_p, n, d_u, shots, d_out);
            KQ_HIP_CHECK(hipMemcpyAsync(out, d_out, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        }
        KQ_HIP_CHECK(hipStreamSynchronize(st));
        for (uint64_t i = 0; i < shots; ++i) {
            if (out[i] >= n) {
                throw std::runtime_error {"LOGIC BUG: Ended up with measurement of state past end of cumulative\nprobability distribution, which shouldn't happen?"};
            }
        }
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

}  // namespace

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

}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

PassProfiler& profiler()
{
    static PassProfiler p;
    return p;
}

void profiler_collect()
{
    PassProfiler& p = profiler();
    for (auto& e : p.pending) {
        float ms = 0.0f;
        KQ_HIP_CHECK(hipEventSynchronize(e.b));
        KQ_HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
        p.cls[e.cls].ms += ms;
        p.cls[e.cls].launches += 1;
        p.cls[e.cls].bytes += e.bytes;
        p.cls[e.cls].gates += e.gates;
        (void) hipEventDestroy(e.a);
        (void) hipEventDestroy(e.b);
    }
    p.pending.clear();
}

void* scratch(kq_state* s, size_t bytes) { return grow(s->scratch, s->scratch_bytes, bytes); }

}  // namespace kq

using namespace kq;

extern "C" {

const char* kq_last_error(void) { return g_last_error.c_str(); }

int kq_init(int device)
{
    return guarded([&] {
        {
            std::lock_guard<std::mutex> lock {g_init_mutex};
            g_device = device;
            g_initialised = false;
        }
        ensure_init();
    });
}

int kq_device_info(char* name, size_t name_len, int* n_cu, size_t* total_mem)
{
    return guarded([&] {
        ensure_init();
        hipDeviceProp_t prop;
        KQ_HIP_CHECK(hipGetDeviceProperties(&prop, g_device));
        if (name && name_len) {
            std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
        }
        if (n_cu) {
            *n_cu = prop.multiProcessorCount;
        }
        if (total_mem) {
            *total_mem = prop.totalGlobalMem;
        }
    });
}

int kq_synchronize(void)
{
    return guarded([&] { KQ_HIP_CHECK(hipDeviceSynchronize()); });
}

int kq_sv_create(kq_state** out, uint32_t n_qubits, int precision)
{
    return guarded([&] {
        if (!out) {
            throw InvalidArg {"null output pointer"};
        }
        *out = create_state(n_qubits, precision, false);
    });
}

int kq_dm_create(kq_state** out, uint32_t n_qubits, int precision)
{
    return guarded([&] {
        if (!out) {
            throw InvalidArg {"null output pointer"};
        }
        *out = create_state(n_qubits, precision, true);
    });
}

int kq_state_destroy(kq_state* s)
{
    return guarded([&] {
        if (!s) {
            return;
        }
        (void) hipSetDevice(g_device);
        if (s->stream) {
            (void) hipStreamSynchronize(s->stream);
            (void) hipStreamDestroy(s->stream);
        }
        if (s->amps) {
            (void) hipFree(s->amps);
        }
        if (s->scratch) {
            (void) hipFree(s->scratch);
        }
        if (s->prog) {
            (void) hipFree(s->prog);
        }
        if (s->stage) {
            (void) hipHostFree(s->stage);
        }
        if (s->aux) {
            (void) hipFree(s->aux);
        }
        delete s;
    });
}

int kq_state_info(const kq_state* s, uint32_t* n_qubits, int* precision, int* is_density)
{
    return guarded([&] {
        check_state(s);
        if (n_qubits) { *n_qubits = s->n; }
        if (precision) { *precision = s->precision; }
        if (is_density) { *is_density = s->is_dm; }
    });
}

int kq_state_device_ptr(kq_state* s, void** ptr, size_t* bytes)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (ptr) { *ptr = s->amps; }
        if (bytes) { *bytes = s->n_amps * amp_bytes(s->precision); }
    });
}

int kq_state_upload(kq_state* s, const double* host)
{
    return guarded([&] {
        check_state(s);
        if (!host) {
            throw InvalidArg {"null host buffer"};
        }
        ensure_init();
        s->bit_of.clear();   // overwritten: qubit q on bit q
        if (s->precision == KQ_C128) {
            KQ_HIP_CHECK(hipMemcpy(s->amps, host, s->n_amps * 16, hipMemcpyHostToDevice));
        }
        else {
            std::vector<float> f(2 * s->n_amps);
            for (size_t i = 0; i < f.size(); ++i) {
                f[i] = static_cast<float>(host[i]);
            }
            KQ_HIP_CHECK(hipMemcpy(s->amps, f.data(), s->n_amps * 8, hipMemcpyHostToDevice));
        }
    });
}

int kq_state_download(const kq_state* s, double* host)
{
    return guarded([&] {
        check_state(s);
        ensure_init();
        normalise_layout(s);
        if (!host) {
            throw InvalidArg {"null host buffer"};
        }
        ensure_init();
        KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
        const auto h = download_host(s);
        std::memcpy(host, h.data(), h.size() * sizeof(double));
    });
}

int kq_state_set_basis(kq_state* s, uint64_t index)
{
    return guarded([&] {
        check_state(s);
        const uint64_t N = uint64_t {1} << s->n;
        if (index >= N) {
            throw InvalidArg {"basis index out of range"};
        }
        ensure_init();
        s->bit_of.clear();   // overwritten: qubit q on bit q
        launch_set_basis(s, s->is_dm ? index + index * N : index);
        KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    });
}

int kq_state_copy(kq_state* dst, const kq_state* src)
{
    return guarded([&] {
        check_state(dst);
        check_state(src);
        if (dst->n_amps != src->n_amps || dst->precision != src->precision || dst->is_dm != src->is_dm) {
            throw InvalidArg {"state copy: shapes differ"};
        }
        ensure_init();
        normalise_layout(src);
        dst->bit_of.clear();
        KQ_HIP_CHECK(hipStreamSynchronize(src->stream));
        launch_copy_kernel(dst, dst->amps, src->amps, dst->n_amps);
        KQ_HIP_CHECK(hipStreamSynchronize(dst->stream));
    });
}

int kq_simulate(kq_state* s, uint32_t circuit_n_qubits, const kq_op* ops, size_t n_ops, const int32_t* preds,
                size_t n_preds, int has_seed, int32_t seed, int8_t* creg, size_t n_bits, kq_logger_fn logger,
                void* logger_user)
{
    return guarded([&] {
        check_state(s);
        if (n_ops > 0 && ops == nullptr) {
            throw InvalidArg {"null op array"};
        }
        ensure_init();
        last_stats() = PlanStats {};
        simulate(s, circuit_n_qubits, ops, n_ops, preds, n_preds, has_seed != 0, seed, creg, n_bits, logger, logger_user);
    });
}

int kq_apply_gates(kq_state* s, const kq_op* ops, size_t n_ops)
{
    return guarded([&] {
        check_state(s);
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

int kq_profile_begin(void)
{
    return guarded([&] {
        PassProfiler& p = profiler();
        profiler_collect();
        p = PassProfiler {};
        p.enabled = true;
    });
}

int kq_profile_end(kq_pass_stats out[3])
{
    return guarded([&] {
        PassProfiler& p = profiler();
        profiler_collect();
        if (out) {
            for (int i = 0; i < 3; ++i) {
                out[i] = p.cls[i];
            }
        }
        p.enabled = false;
    });
}

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

int kq_stream_probe(uint32_t n_qubits, double gbps[4])
{
    return guarded([&] {
        if (n_qubits < 16 || n_qubits > 32 || !gbps) {
            throw InvalidArg {"stream probe: 16 <= n_qubits <= 32 and a 4-value output"};
        }
        ensure_init();
        stream_probe(n_qubits, gbps);
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

int kq_set_jit(int mode)
{
    return guarded([&] { set_jit_mode(mode); });
}

int kq_set_jit_tuning(int min_waves_per_simd, int blocks_per_cu, int normalise)
{
    return guarded([&] {
        JitTuning t = jit_tuning();
        t.min_waves_per_simd = min_waves_per_simd;
        t.blocks_per_cu = blocks_per_cu;
        t.normalise = normalise != 0;
        set_jit_tuning(t);
    });
}

int kq_set_jit_priority(int mode)
{
    return guarded([&] {
        JitTuning t = jit_tuning();
        t.priority = mode;
        set_jit_tuning(t);
    });
}

int kq_jit_info(int* mode, size_t* compiled, size_t* cache_hits, size_t* failures, double* compile_seconds,
                size_t* uncapped)
{
    return guarded([&] {
        const JitStats& st = jit_stats();
        if (mode) {
            *mode = jit_mode();
        }
        if (compiled) {
            *compiled = st.compiled;
        }
        if (cache_hits) {
            *cache_hits = st.hits;
        }
        if (failures) {
            *failures = st.failures;
        }
        if (compile_seconds) {
            *compile_seconds = st.compile_seconds;
        }
        if (uncapped) {
            *uncapped = st.uncapped;
        }
    });
}

namespace
{

// run-time pass kernel sources of a straight-line gate list (host only)
// One pass of a straight-line gate list as run_records plans it for a statevector (qubit moves and
// the final restore passes included): a generated tile kernel with its coefficients, or a sweep gate.
struct DebugPass
{
    bool sweep = false;
    std::string src;
    std::vector<double> coefs;
    uint32_t blocks = 0;
    uint32_t threads = 0;
    GateRec gate {};
};

std::vector<DebugPass> debug_passes(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes)
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
    if (fusion_relabel()) {
        recs = exchanges_as_relabels(recs, n_qubits, sigma);
    }
    auto passes = plan_passes(recs, n_qubits, max_k, fusion_low_qubits(), fusion_max_gates(), register_qubits(precision),
                              fusion_max_stages(), true, fusion_relabel() ? &bit_of : nullptr);
    if (!sigma.empty() && bit_of.size() == n_qubits) {
        const std::vector<uint8_t> planned = bit_of;
        for (uint32_t q = 0; q < n_qubits; ++q) {
            bit_of[q] = planned[sigma[q]];
        }
    }
    bool moved = false;
    for (size_t q = 0; q < bit_of.size(); ++q) {
        moved = moved || bit_of[q] != q;
    }
    if (moved) {
        for (auto& p : plan_restore(bit_of, n_qubits, max_k, fusion_low_qubits(), register_qubits(precision))) {
            passes.push_back(std::move(p));
        }
    }
    std::vector<DebugPass> out;
    for (const auto& p : passes) {
        DebugPass d;
        if (p.sweep) {
            d.sweep = true;
            d.gate = p.single;
        }
        else {
            TileProgram tp;
            lower_tile_pass(p, precision, n_qubits, tp);
            d.src = jit_source(tp, precision, d.coefs);
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
        info[0] = d.sweep ? 1 : 0;
        info[1] = d.blocks;
        info[2] = d.threads;
        info[3] = d.sweep ? d.gate.target : 0;
        info[4] = d.sweep ? static_cast<uint64_t>(d.gate.control + 1) : 0;
    });
}

int kq_batch_create(kq_batch** out, int precision, int forward_precision)
{
    return guarded([&] {
        if (!out || (precision != KQ_C128 && precision != KQ_C64)
            || (forward_precision != KQ_C128 && forward_precision != KQ_C64)) {
            throw InvalidArg {"batch: bad arguments"};
        }
        ensure_init();
        hipStream_t st = nullptr;
        KQ_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        try {
            *out = batch_create(precision, forward_precision, st);
        }
        catch (...) {
            (void) hipStreamDestroy(st);
            throw;
        }
    });
}

int kq_batch_add(kq_batch* b, uint32_t n_qubits, const kq_op* ops, size_t n_ops, size_t n_params, size_t n_strings,
                 const uint32_t* term_offsets, const uint32_t* term_qubits, const uint8_t* term_kinds, const uint8_t* phases,
                 const double* coeffs, size_t* index)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        batch_add(b, n_qubits, ops, n_ops, n_params, n_strings, term_offsets, term_qubits, term_kinds, phases, coeffs, index);
    });
}

int kq_batch_set_parameters(kq_batch* b, size_t index, const double* values)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        batch_set_parameters(b, index, values);
    });
}

int kq_batch_run(kq_batch* b, int use_graph, double* energies, double* grads, double* drift)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        ensure_init();
        batch_run(b, use_graph != 0, energies, grads, drift);
    });
}

int kq_batch_info(const kq_batch* b, size_t* n_instances, size_t* n_params_total, size_t* device_bytes, size_t* n_launches)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        batch_info(b, n_instances, n_params_total, device_bytes, n_launches);
    });
}

int kq_batch_destroy(kq_batch* b)
{
    return guarded([&] {
        if (b) {
            batch_destroy(b);
        }
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
