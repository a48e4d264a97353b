#include <cstdint>
// Gate compilation (kettle gate -> device record) and the fusion planner.
//
// Coefficients are evaluated here, on the host, with kettle's own expressions: cos/sin of
// theta / 2.0 for RX/RY/RZ and of theta for P (operations.cpp:111-173), M_SQRT1_2 and 0.5
// literals for H, T, SX (operations.cpp:13-109).  The device never evaluates a transcendental.
#include <algorithm>
#include <array>
#include <complex>
#include <bit>
#include <cmath>
#include <cstring>

#include "kq_internal.hpp"

namespace kq
{
namespace
{

struct M
{
    double m[8];
};

M dense(double a, double b, double c, double d, double e, double f, double g, double h)
{
    return M {{a, b, c, d, e, f, g, h}};
}

// kettle's per-pair formula of each gate, written as the 2x2 matrix it applies
GateKind gate_matrix(int gate, double theta, const double* u, M& out)
{
    const double r = M_SQRT1_2;
    switch (gate) {
        case KQ_H: case KQ_CH: out = dense(r, 0, r, 0, r, 0, -r, 0); return GK_DENSE;
        case KQ_X: case KQ_CX: out = dense(0, 0, 1, 0, 1, 0, 0, 0); return GK_SWAP;
        case KQ_Y: case KQ_CY: out = dense(0, 0, 0, -1, 0, 1, 0, 0); return GK_DENSE;
        case KQ_Z: case KQ_CZ: out = dense(1, 0, 0, 0, 0, 0, -1, 0); return GK_DIAG;
        case KQ_S: case KQ_CS: out = dense(1, 0, 0, 0, 0, 0, 0, 1); return GK_DIAG;
        case KQ_SDAG: case KQ_CSDAG: out = dense(1, 0, 0, 0, 0, 0, 0, -1); return GK_DIAG;
        case KQ_T: case KQ_CT: out = dense(1, 0, 0, 0, 0, 0, r, r); return GK_DIAG;
        case KQ_TDAG: case KQ_CTDAG: out = dense(1, 0, 0, 0, 0, 0, r, -r); return GK_DIAG;
        case KQ_SX: case KQ_CSX: out = dense(0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5); return GK_DENSE;
        case KQ_SXDAG: case KQ_CSXDAG: out = dense(0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5, -0.5); return GK_DENSE;
        case KQ_RX: case KQ_CRX: {
            const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
            out = dense(c, 0, 0, -s, 0, -s, c, 0);
            return GK_DENSE;
        }
        case KQ_RY: case KQ_CRY: {
            const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
            out = dense(c, 0, -s, 0, s, 0, c, 0);
            return GK_DENSE;
        }
        case KQ_RZ: case KQ_CRZ: {
            const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
            out = dense(c, -s, 0, 0, 0, 0, c, s);
            return GK_DIAG;
        }
        case KQ_P: case KQ_CP: {
            const double c = std::cos(theta), s = std::sin(theta);
            out = dense(1, 0, 0, 0, 0, 0, c, s);
            return GK_DIAG;
        }
        case KQ_U: case KQ_CU: {
            std::memcpy(out.m, u, sizeof out.m);
            const bool diag = u[2] == 0.0 && u[3] == 0.0 && u[4] == 0.0 && u[5] == 0.0;
            return diag ? GK_DIAG : GK_DENSE;
        }
        default: break;
    }
    throw InvalidArg {"unknown gate code " + std::to_string(gate)};
}

bool is_controlled(int gate) { return (gate >= KQ_CH && gate <= KQ_CP) || gate == KQ_CU; }

// the cheapest exact arithmetic form of a matrix (kernels skip the terms that are zero)
GateKind classify(const M& m)
{
    const double* e = m.m;
    if (e[2] == 0.0 && e[3] == 0.0 && e[4] == 0.0 && e[5] == 0.0) {
        return (e[0] == 1.0 && e[1] == 0.0) ? GK_DIAG1 : GK_DIAG;
    }
    if (e[0] == 0.0 && e[1] == 0.0 && e[6] == 0.0 && e[7] == 0.0 && e[2] == 1.0 && e[3] == 0.0 && e[4] == 1.0
        && e[5] == 0.0) {
        return GK_SWAP;
    }
    if (e[1] == 0.0 && e[3] == 0.0 && e[5] == 0.0 && e[7] == 0.0) {
        return GK_REAL;
    }
    if (e[1] == 0.0 && e[7] == 0.0 && e[2] == 0.0 && e[4] == 0.0) {
        return GK_AXIS;
    }
    return GK_DENSE;
}

GateRec make_rec(GateKind kind, uint32_t target, int32_t control, const M& m)
{
    GateRec g {};
    g.kind = kind;
    g.target = target;
    g.control = control;
    std::memcpy(g.m, m.m, sizeof g.m);
    return g;
}

PlanStats g_stats;
int g_max_k = -1;          // -1: automatic
int g_max_gates = 4096;
int g_max_stages = 12;
bool g_autotune = true;
double g_max_cost = 1.0e9;

}  // namespace

PlanStats& last_stats() { return g_stats; }

size_t& device_lds_per_block()
{
    static size_t bytes = 65536;
    return bytes;
}

uint32_t& device_cus()
{
    static uint32_t n = 256;
    return n;
}

uint32_t device_cu_count() { return device_cus(); }

double fusion_max_cost() { return g_max_cost; }

void set_fusion(int max_k, int max_gates, int max_stages, double max_cost)
{
    if (max_cost > 0.0) {
        g_max_cost = max_cost;
    }
    g_max_k = max_k;
    if (max_gates > 0) {
        g_max_gates = max_gates;
    }
    if (max_stages > 0) {
        g_max_stages = max_stages;
        g_autotune = false;   // an explicit setting wins over the tuner
    }
}

int fusion_setting_k() { return g_max_k; }
int fusion_max_stages() { return g_max_stages; }

namespace
{
// 5: 512-byte contiguous chunks per access.  With qubits moving onto the low bits between passes
// the 30-qubit QFT + layers circuit needs 28 passes at 5 and 29 at 4 (35 and 39 without moves), and
// 5 streams faster.
int g_low_qubits = 5;
}

int fusion_low_qubits() { return g_low_qubits; }

namespace
{
bool g_relabel = true;
}

bool fusion_relabel() { return g_relabel; }
void set_fusion_relabel(bool on) { g_relabel = on; }

bool fusion_autotune() { return g_autotune; }
void set_fusion_autotune(bool on) { g_autotune = on; }

void set_fusion_low_qubits(int c)
{
    if (c < 0 || c > 12) {
        throw InvalidArg {"low qubits per tile must be in [0, 12]"};
    }
    g_low_qubits = c;
    g_autotune = false;   // an explicit setting wins over the tuner
}

// LDS remaps a tile pass runs: one per change of register bits between its layouts, counting the
// load layout (when the first stage does not load directly) and the store layout.
int pass_remaps(const TilePass& p)
{
    if (p.sweep || p.stages.empty()) {
        return 0;
    }
    const auto regs = [&](const TileStage& s) {
        uint32_t m = 0;
        for (int q = 0; q < p.rq; ++q) {
            m |= 1u << s.reg[q];
        }
        return m;
    };
    int n = 0;
    uint32_t cur = regs(p.load_direct ? p.stages.front() : p.io_layout);
    n += p.load_direct ? 0 : 1;
    for (const TileStage& s : p.stages) {
        if (regs(s) != cur) {
            ++n;
            cur = regs(s);
        }
    }
    if (!p.out_qubits.empty() || !(p.store_direct || regs(p.io_layout) == cur)) {
        ++n;
    }
    return n;
}
int register_qubits(int) { return 4; }
int fusion_max_gates() { return g_max_gates; }

// Tile size: 2^k amplitudes.  k = 12 gives 256 threads of 16 register amplitudes each and a
// 64 KiB (complex128) / 32 KiB (complex64) LDS tile for the stage remaps.
int fusion_max_k(int precision, uint32_t nvec) { return fusion_max_k_for(precision, nvec, device_lds_per_block()); }

int fusion_max_k_for(int precision, uint32_t nvec, size_t lds_bytes)
{
    int k = g_max_k;
    if (k < 0) {
        k = 12;
    }
    const size_t ab = precision == KQ_C64 ? 8 : 16;
    while (k > 1 && (size_t {1} << k) * ab + TILE_STATIC_LDS > lds_bytes) {
        --k;
    }
    return std::min<int>(k, static_cast<int>(nvec));
}

bool is_diag(const GateRec& g) { return g.kind == GK_DIAG || g.kind == GK_DIAG1; }

double gate_cost(const GateRec& g)
{
    switch (g.kind) {
        case GK_DENSE: return 1.0;
        case GK_REAL: case GK_AXIS: return 0.5;
        case GK_DIAG: return 0.5;
        case GK_DIAG1: return 0.25;
        case GK_SWAP: return 0.1;
        default: break;
    }
    return 1.0;
}

void compile_gate(const kq_op& op, uint32_t n, bool is_dm, std::vector<GateRec>& out)
{
    M m;
    (void) gate_matrix(op.gate, op.angle, op.u, m);
    const GateKind kind = classify(m);
    const bool ctrl = is_controlled(op.gate);
    const uint32_t target = static_cast<uint32_t>(ctrl ? op.q1 : op.q0);
    const int32_t control = ctrl ? op.q0 : -1;
    if (target >= n || (ctrl && static_cast<uint32_t>(control) >= n)) {
        throw InvalidArg {"gate qubit index out of range"};
    }
    if (ctrl && static_cast<uint32_t>(control) == target) {
        throw InvalidArg {"control and target qubits coincide"};
    }
    out.push_back(make_rec(kind, target, control, m));
    if (is_dm) {
        // rho <- U rho U^dagger: U on the row bit t, conj(U) on the column bit n + t
        M cm = m;
        for (int i = 1; i < 8; i += 2) {
            cm.m[i] = -cm.m[i];
        }
        out.push_back(make_rec(classify(cm), target + n, ctrl ? control + static_cast<int32_t>(n) : -1, cm));
    }
}

void compile_gate_stable(const kq_op& op, uint32_t n, std::vector<GateRec>& out)
{
    const size_t first = out.size();
    compile_gate(op, n, false, out);
    if (op.aux < 0) {
        return;
    }
    // a parameterised gate keeps the arithmetic form of its gate type for every angle, so a plan made
    // for one set of angles stays valid for any other
    GateRec& g = out[first];
    switch (op.gate) {
        case KQ_RX: case KQ_CRX: g.kind = GK_AXIS; break;
        case KQ_RY: case KQ_CRY: g.kind = GK_REAL; break;
        case KQ_RZ: case KQ_CRZ: g.kind = GK_DIAG; break;
        case KQ_P: case KQ_CP: g.kind = GK_DIAG1; break;
        default: break;
    }
}

namespace
{

uint32_t enc(uint32_t src, uint32_t idx) { return (idx << 2) | src; }

// choose the register bits of each stage: the stage's target bits, completed from the previous
// stage's register bits (equal layouts need no LDS remap), then from local bits >= 3 (keeps the
// lowest local bits on lanes for coalesced HBM access), then any
void complete_stage(TileStage& st, uint32_t need, uint32_t prev, int k, int rq)
{
    uint32_t regs = need;
    const auto fill = [&](uint32_t from) {
        for (int b = 0; b < k && std::popcount(regs) < rq; ++b) {
            if (((from >> b) & 1u) && !((regs >> b) & 1u)) {
                regs |= 1u << b;
            }
        }
    };
    fill(prev);
    fill(((1u << k) - 1) & ~7u);
    fill((1u << k) - 1);
    const uint32_t all = (k >= 32) ? ~0u : ((1u << k) - 1);
    int p = 0;
    for (int b = 0; b < k; ++b) {
        if ((regs >> b) & 1u) {
            st.reg[p++] = static_cast<uint8_t>(b);
        }
    }
    st.thr_mask = all & ~regs;
}

uint32_t stage_regs(const TileStage& st, int rq)
{
    uint32_t m = 0;
    for (int p = 0; p < rq; ++p) {
        m |= 1u << st.reg[p];
    }
    return m;
}

}  // namespace

namespace
{

// One greedy pass over `remaining` (records in index bits): which gates join, their stages, the
// deferred rest and the tile's index bits.
struct GreedyPass
{
    std::vector<size_t> taken;
    std::vector<int> taken_stage;
    std::vector<uint64_t> stage_targets {0};
    std::vector<uint64_t> stage_all {0};   // every bit a gate of the stage touches (targets, controls)
    std::vector<size_t> deferred;
    uint64_t q_mask = 0;
};

GreedyPass greedy_pass(const std::vector<GateRec>& gates, const std::vector<size_t>& remaining, uint64_t low_mask, int k,
                       int rq, int max_gates, int max_stages, size_t max_scan = SIZE_MAX)
{
    GreedyPass gp;
    gp.q_mask = low_mask;
    int q_count = std::popcount(low_mask);
    uint64_t d_any = 0, d_nd = 0;
    double cost = 0.0;
    const double max_cost = fusion_max_cost();
    size_t scanned = 0;
    for (const size_t r : remaining) {
        if (++scanned > max_scan) {
            break;   // scoring only: the rest would be deferred or join rarely
        }
        const GateRec& g = gates[r];
        const uint64_t tbit = uint64_t {1} << g.target;
        const uint64_t nd = is_diag(g) ? 0 : tbit;
        const uint64_t all = tbit | (g.control >= 0 ? (uint64_t {1} << g.control) : 0);
        bool ok = static_cast<int>(gp.taken.size()) < max_gates && !(nd & d_any) && !(all & d_nd)
               && (gp.taken.empty() || cost + gate_cost(g) <= max_cost);
        uint64_t need = 0;
        bool new_stage = false;
        int place = -1;
        if (ok && nd) {
            need = nd & ~gp.q_mask;
            if (q_count + std::popcount(need) > k) {
                ok = false;
            }
            else {
                // Backfill: a non-diagonal gate may join an earlier stage that holds its target in
                // registers (or has a free register) when it commutes with every gate of the later
                // stages (bit-mask test, as for deferral); it runs after that stage's gates.  A stage
                // holding the target wins over one with room; among those the latest, which leaves
                // earlier stages open to later gates.
                uint64_t later_nd = 0, later_all = 0;
                int room = -1;
                for (int st = static_cast<int>(gp.stage_targets.size()) - 1; st >= 0; --st) {
                    if ((nd & later_all) || (all & later_nd)) {
                        break;
                    }
                    if (gp.stage_targets[static_cast<size_t>(st)] & nd) {
                        place = st;
                        break;
                    }
                    if (room < 0 && std::popcount(gp.stage_targets[static_cast<size_t>(st)]) < rq) {
                        room = st;
                    }
                    later_nd |= gp.stage_targets[static_cast<size_t>(st)];
                    later_all |= gp.stage_all[static_cast<size_t>(st)];
                }
                if (place < 0) {
                    place = room;
                }
                if (place < 0) {
                    if (static_cast<int>(gp.stage_targets.size()) >= max_stages) {
                        ok = false;
                    }
                    else {
                        new_stage = true;
                    }
                }
            }
        }
        if (ok && !nd) {
            // a diagonal gate needs no register: it joins the earliest stage after which every gate
            // commutes with it, so it does not stand in the way of later gates joining earlier stages
            uint64_t later_nd = 0;
            place = static_cast<int>(gp.stage_targets.size()) - 1;
            for (int st = place; st > 0; --st) {
                later_nd |= gp.stage_targets[static_cast<size_t>(st)];
                if (all & later_nd) {
                    break;
                }
                place = st - 1;
            }
        }
        if (ok) {
            cost += gate_cost(g);
            gp.q_mask |= need;
            q_count += std::popcount(need);
            if (new_stage) {
                gp.stage_targets.push_back(0);
                gp.stage_all.push_back(0);
            }
            if (place < 0) {
                place = static_cast<int>(gp.stage_targets.size()) - 1;
            }
            gp.stage_targets[static_cast<size_t>(place)] |= nd;
            gp.stage_all[static_cast<size_t>(place)] |= all;
            gp.taken.push_back(r);
            gp.taken_stage.push_back(place);
        }
        else {
            gp.deferred.push_back(r);
            d_any |= all;
            d_nd |= nd;
        }
    }
    return gp;
}

// the record with its qubits replaced by the index bits that hold them
GateRec at_bits(const GateRec& g, const std::vector<uint8_t>& bit_of)
{
    GateRec t = g;
    t.target = bit_of[g.target];
    if (g.control >= 0) {
        t.control = bit_of[static_cast<size_t>(g.control)];
    }
    return t;
}

// After a pass with tile index bits q_mask: which qubits of the tile should sit on index bits
// 0..c-1 for the next pass.  Candidates are the tile's qubits in the order the deferred gates
// first touch them; every c-subset of the first c + 3 is scored by how many gates a greedy next
// pass would take, and the current placement wins ties.  Returns the new qubit -> bit map.
std::vector<uint8_t> choose_low_qubits(const std::vector<GateRec>& gates, const std::vector<size_t>& deferred,
                                       const std::vector<uint8_t>& bit_of, uint64_t q_mask, int c, int k, int rq,
                                       int max_gates, int max_stages)
{
    const size_t nvec = bit_of.size();
    std::vector<int> qubit_at(nvec);
    for (size_t q = 0; q < nvec; ++q) {
        qubit_at[bit_of[q]] = static_cast<int>(q);
    }
    const auto in_tile = [&](int q) { return ((q_mask >> bit_of[static_cast<size_t>(q)]) & 1ull) != 0; };
    std::vector<int> order;
    const auto add = [&](int q) {
        if (q >= 0 && in_tile(q) && std::find(order.begin(), order.end(), q) == order.end()) {
            order.push_back(q);
        }
    };
    for (const size_t r : deferred) {
        add(static_cast<int>(gates[r].target));
        add(gates[r].control);
        if (static_cast<int>(order.size()) >= c + 3) {
            break;
        }
    }
    for (uint32_t b = 0; b < nvec && static_cast<int>(order.size()) < c + 3; ++b) {
        if ((q_mask >> b) & 1ull) {
            add(qubit_at[b]);
        }
    }
    const int m = static_cast<int>(order.size());
    const uint64_t low_mask = (uint64_t {1} << c) - 1;
    // a placement: the chosen qubits take bits 0..c-1; the qubits they displace take the bits
    // the chosen ones leave (all within the tile)
    const auto place = [&](const std::vector<int>& chosen) {
        std::vector<uint8_t> nb = bit_of;
        std::vector<int> incoming, outgoing;
        for (const int q : chosen) {
            if (bit_of[static_cast<size_t>(q)] >= c) {
                incoming.push_back(q);
            }
        }
        for (int b = 0; b < c; ++b) {
            const int q = qubit_at[static_cast<size_t>(b)];
            if (std::find(chosen.begin(), chosen.end(), q) == chosen.end()) {
                outgoing.push_back(q);
            }
        }
        for (size_t i = 0; i < incoming.size(); ++i) {
            std::swap(nb[static_cast<size_t>(incoming[i])], nb[static_cast<size_t>(outgoing[i])]);
        }
        return nb;
    };
    // a candidate is scored on the first deferred gates only (a pass rarely reaches further)
    constexpr size_t SCAN = 400;
    std::vector<GateRec> trial(gates.size());
    const auto score = [&](const std::vector<uint8_t>& nb) {
        size_t i = 0;
        for (const size_t r : deferred) {
            if (++i > SCAN) {
                break;
            }
            trial[r] = at_bits(gates[r], nb);
        }
        return greedy_pass(trial, deferred, low_mask, k, rq, max_gates, max_stages, SCAN).taken.size();
    };
    std::vector<uint8_t> best = bit_of;
    size_t best_score = score(bit_of);
    if (m < c) {
        return best;
    }
    std::vector<int> idx(static_cast<size_t>(c));
    for (int i = 0; i < c; ++i) {
        idx[static_cast<size_t>(i)] = i;
    }
    while (true) {
        std::vector<int> chosen;
        for (const int i : idx) {
            chosen.push_back(order[static_cast<size_t>(i)]);
        }
        const auto nb = place(chosen);
        const size_t sc = score(nb);
        if (sc > best_score) {
            best_score = sc;
            best = nb;
        }
        int i = c - 1;
        while (i >= 0 && idx[static_cast<size_t>(i)] == m - c + i) {
            --i;
        }
        if (i < 0) {
            break;
        }
        ++idx[static_cast<size_t>(i)];
        for (int j = i + 1; j < c; ++j) {
            idx[static_cast<size_t>(j)] = idx[static_cast<size_t>(j - 1)] + 1;
        }
    }
    return best;
}

}  // namespace

// Greedy pass builder.  A pass owns a tile set Q of index bits (|Q| <= k, always containing the
// low bits 0..c-1).  Gates are taken in circuit order; a gate joins the current pass when
//   - its non-diagonal target is in Q or Q still has room for it,
//   - it commutes with every gate deferred so far (it shares no qubit with a deferred
//     non-diagonal target, and its own non-diagonal target touches no deferred gate), and
//   - its target fits the current register stage (at most rq distinct non-diagonal targets per
//     stage) or a new stage may still be opened (at most max_stages per pass).
// Diagonal gates and controls never need to be in Q: their effect depends only on index bits.
// With `bit_of` (qubit -> index bit holding it, updated on return) a pass may also move qubits of
// its tile onto bits 0..c-1 as it stores the tile, chosen for the next pass (`out_qubits`); without
// it qubit q is index bit q throughout.
std::vector<TilePass> plan_passes(const std::vector<GateRec>& gates_in, uint32_t nvec, int max_k, int low_c, int max_gates,
                                  int rq_max, int max_stages, bool allow_sweep, std::vector<uint8_t>* bit_of_io,
                                  const std::function<void(TilePass&&)>* emit, const std::vector<uint8_t>* final_bits,
                                  bool out_of_place)
{
    std::vector<TilePass> passes;
    const auto hand_over = [&](TilePass&& p) {
        if (emit) {
            (*emit)(std::move(p));
        }
        else {
            passes.push_back(std::move(p));
        }
    };
    const int k = std::min<int>(max_k, static_cast<int>(nvec));
    const int rq = std::min(rq_max, k);
    // a tile covering every qubit needs no free slots; otherwise keep at least one slot free
    const int c = k == static_cast<int>(nvec) ? k : std::max(0, std::min(low_c, k - 1));
    std::vector<uint8_t> bit_of(nvec);
    for (uint32_t q = 0; q < nvec; ++q) {
        bit_of[q] = static_cast<uint8_t>(q);
    }
    if (bit_of_io && bit_of_io->size() == nvec) {
        bit_of = *bit_of_io;
    }
    const bool relabel = bit_of_io != nullptr && k < static_cast<int>(nvec) && c > 0;
    std::vector<GateRec> gates(gates_in.size());
    std::vector<size_t> remaining(gates_in.size());
    for (size_t i = 0; i < gates_in.size(); ++i) {
        remaining[i] = i;
    }
    while (!remaining.empty()) {
        for (const size_t r : remaining) {
            gates[r] = at_bits(gates_in[r], bit_of);
        }
        const uint64_t low_mask = (c == 64) ? ~0ull : ((uint64_t {1} << c) - 1);
        GreedyPass gp = greedy_pass(gates, remaining, low_mask, k, rq, max_gates, max_stages);
        {
            // backfilled gates run with their stage: group by stage, circuit order within a stage
            std::vector<size_t> ord(gp.taken.size());
            for (size_t i = 0; i < ord.size(); ++i) {
                ord[i] = i;
            }
            std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return gp.taken_stage[a] < gp.taken_stage[b]; });
            std::vector<size_t> t2(ord.size());
            std::vector<int> s2(ord.size());
            for (size_t i = 0; i < ord.size(); ++i) {
                t2[i] = gp.taken[ord[i]];
                s2[i] = gp.taken_stage[ord[i]];
            }
            gp.taken = std::move(t2);
            gp.taken_stage = std::move(s2);
        }
        uint64_t q_mask = gp.q_mask;
        int q_count = std::popcount(q_mask);
        std::vector<size_t>& taken = gp.taken;
        std::vector<int>& taken_stage = gp.taken_stage;
        std::vector<uint64_t>& stage_targets = gp.stage_targets;
        std::vector<size_t>& deferred = gp.deferred;
        if (taken.empty()) {
            throw InvalidArg {"fusion plan: no gate fits a tile"};
        }
        TilePass p;
        if (allow_sweep && taken.size() == 1 && is_diag(gates[taken[0]])) {
            // a lone diagonal gate touches only a quarter / half of the state: plain sweep
            p.sweep = true;
            p.single = gates[taken[0]];
            p.src.push_back(static_cast<uint32_t>(taken[0]));
            hand_over(std::move(p));
            remaining = std::move(deferred);
            continue;
        }
        // The last pass stores every qubit on its final bit when the bits that takes fit the tile.  Index
        // bit b holds a qubit whose final bit is dest[b] (a permutation).  In place, the tile must hold
        // every bit that moves (the moved bits are closed under dest); out of place it must hold the bits
        // whose contents land on the destination's low bits 0..cs-1, so that each wavefront still stores
        // runs of 2^cs amplitudes: cs = c when those bits fit, else fewer, down to 3 (runs of 64 or 128
        // bytes).
        std::vector<int> dest;
        bool fold_oop = false;
        if (relabel && deferred.empty() && final_bits && final_bits->size() == nvec && *final_bits != bit_of) {
            std::vector<uint32_t> qubit_on(nvec);
            for (uint32_t q = 0; q < nvec; ++q) {
                qubit_on[bit_of[q]] = q;
            }
            std::vector<int> d(nvec);
            uint64_t moved = 0;
            for (uint32_t b = 0; b < nvec; ++b) {
                d[b] = (*final_bits)[qubit_on[b]];
                moved |= static_cast<uint64_t>(d[b] != static_cast<int>(b)) << b;
            }
            if (std::popcount(q_mask | moved) <= k) {
                q_mask |= moved;
                dest = std::move(d);
            }
            else if (out_of_place) {
                for (int cs = c; cs >= 3 && dest.empty(); --cs) {
                    uint64_t low_dest = 0;
                    for (uint32_t b = 0; b < nvec; ++b) {
                        low_dest |= static_cast<uint64_t>(d[b] < cs) << b;
                    }
                    if (std::popcount(q_mask | low_dest) <= k) {
                        q_mask |= low_dest;
                        dest = d;
                        fold_oop = true;
                    }
                }
            }
            q_count = std::popcount(q_mask);
        }
        // fill Q with the lowest free qubits (longer contiguous runs coalesce better)
        for (uint32_t q = 0; q < nvec && q_count < k; ++q) {
            if (!((q_mask >> q) & 1ull)) {
                q_mask |= uint64_t {1} << q;
                ++q_count;
            }
        }
        p.k = k;
        p.rq = rq;
        for (uint32_t q = 0; q < nvec; ++q) {
            if ((q_mask >> q) & 1ull) {
                p.qubits.push_back(static_cast<int>(q));
            }
        }
        p.c = std::min(std::countr_one(q_mask), k);
        // keep the high-offset table within 2^10 entries
        if (p.k - p.c > 10) {
            throw InvalidArg {"fusion plan: tile has more than 10 non-contiguous qubits"};
        }
        const auto local = [&](uint64_t q) { return static_cast<uint32_t>(std::popcount(q_mask & ((uint64_t {1} << q) - 1))); };
        // stages: register bits and layouts
        const int n_st = static_cast<int>(stage_targets.size());
        p.stages.resize(static_cast<size_t>(n_st));
        uint32_t prev = 0;
        for (int si = 0; si < n_st; ++si) {
            uint32_t need_local = 0;
            for (uint32_t q = 0; q < nvec; ++q) {
                if ((stage_targets[si] >> q) & 1ull) {
                    need_local |= 1u << local(q);
                }
            }
            complete_stage(p.stages[si], need_local, prev, k, rq);
            prev = stage_regs(p.stages[si], rq);
        }
        // gates, grouped by stage (taken order is stage order), with encoded bit sources
        const auto encode = [&](const TileStage& st, uint64_t q) -> uint32_t {
            if (!((q_mask >> q) & 1ull)) {
                return enc(2, static_cast<uint32_t>(q));
            }
            const uint32_t lb = local(q);
            for (int pp = 0; pp < rq; ++pp) {
                if (st.reg[pp] == lb) {
                    return enc(0, static_cast<uint32_t>(pp));
                }
            }
            return enc(1, static_cast<uint32_t>(std::popcount(st.thr_mask & ((1u << lb) - 1))));
        };
        size_t gi = 0;
        for (int si = 0; si < n_st; ++si) {
            TileStage& st = p.stages[si];
            st.g_begin = static_cast<uint32_t>(p.gates.size());
            while (gi < taken.size() && taken_stage[gi] == si) {
                GateRec g = gates[taken[gi]];
                const uint32_t t_enc = encode(st, g.target);
                if (!is_diag(g) && (t_enc & 3u) != 0) {
                    throw InvalidArg {"fusion plan: non-diagonal target outside the stage registers"};
                }
                if (g.control >= 0) {
                    g.control = static_cast<int32_t>(encode(st, static_cast<uint64_t>(g.control)));
                }
                g.target = t_enc;
                g.pad = 0;
                p.gates.push_back(g);
                p.src.push_back(static_cast<uint32_t>(taken[gi]));
                ++gi;
            }
            st.g_count = static_cast<uint32_t>(p.gates.size()) - st.g_begin;
        }
        // HBM I/O layout: the stage's own when its register bits leave local bits 0..2 on lanes
        const uint32_t low3 = k > rq + 2 ? 7u : 0u;
        p.load_direct = (stage_regs(p.stages.front(), rq) & low3) == 0;
        p.store_direct = (stage_regs(p.stages.back(), rq) & low3) == 0;
        complete_stage(p.io_layout, 0, 0, k, rq);
        {
            // default I/O layout: the top rq local bits in registers
            uint32_t top = 0;
            for (int b = k - 1; b >= 0 && std::popcount(top) < rq; --b) {
                top |= 1u << b;
            }
            complete_stage(p.io_layout, top, 0, k, rq);
        }
        if (!dest.empty()) {
            p.out_qubits.resize(p.qubits.size());
            for (size_t j = 0; j < p.qubits.size(); ++j) {
                p.out_qubits[j] = dest[static_cast<size_t>(p.qubits[j])];
            }
            if (fold_oop) {
                for (uint32_t b = 0; b < nvec; ++b) {
                    if (!((q_mask >> b) & 1ull)) {
                        p.out_free.push_back(dest[b]);
                    }
                }
            }
            bit_of = *final_bits;
        }
        else if (relabel && !deferred.empty()) {
            const auto nb = choose_low_qubits(gates_in, deferred, bit_of, q_mask, c, k, rq, max_gates, max_stages);
            if (nb != bit_of) {
                // local bit j holds the qubit on index bit p.qubits[j]; after the store it is on nb[qubit]
                std::vector<int> qubit_at(nvec);
                for (uint32_t q = 0; q < nvec; ++q) {
                    qubit_at[bit_of[q]] = static_cast<int>(q);
                }
                p.out_qubits.resize(p.qubits.size());
                for (size_t j = 0; j < p.qubits.size(); ++j) {
                    p.out_qubits[j] = nb[static_cast<size_t>(qubit_at[static_cast<size_t>(p.qubits[j])])];
                }
                bit_of = nb;
            }
        }
        hand_over(std::move(p));
        remaining = std::move(deferred);
    }
    if (bit_of_io) {
        *bit_of_io = bit_of;
    }
    return passes;
}

// Gate-free tile passes that bring every qubit q back to index bit q.  A pass may permute the
// qubits its tile's bits hold freely, and every tile holds the low bits 0..c-1, which act as a hub:
// a pass takes in the home bits of the displaced qubits sitting on low bits and the bits holding the
// low bits' own qubits (both then go home), follows the displacement cycles while it has room, and
// sends home every qubit whose bit and home bit it holds.  When nothing can go home (every displaced
// qubit sits on a high bit with a high home) it moves one displaced qubit onto a low bit, from where
// the next pass sends it home.
std::vector<TilePass> plan_restore(std::vector<uint8_t> bit_of, uint32_t nvec, int max_k, int low_c, int rq_max)
{
    std::vector<TilePass> passes;
    const int k = std::min<int>(max_k, static_cast<int>(nvec));
    const int rq = std::min(rq_max, k);
    const int c = k == static_cast<int>(nvec) ? k : std::max(1, std::min(low_c, k - 1));
    const auto is_home = [&](uint32_t q) { return bit_of[q] == q; };
    for (int guard = 0;; ++guard) {
        bool displaced = false;
        for (uint32_t q = 0; q < nvec; ++q) {
            displaced = displaced || !is_home(q);
        }
        if (!displaced) {
            break;
        }
        if (guard > 8 * static_cast<int>(nvec)) {
            throw InvalidArg {"restore plan: no progress"};
        }
        std::vector<uint32_t> qubit_at(nvec);
        for (uint32_t q = 0; q < nvec; ++q) {
            qubit_at[bit_of[q]] = q;
        }
        const uint64_t low = (uint64_t {1} << c) - 1;
        uint64_t t_mask = low;
        int room = k - c;
        const auto in_t = [&](uint32_t b) { return ((t_mask >> b) & 1ull) != 0; };
        const auto add = [&](uint32_t b) {
            if (!in_t(b) && room > 0) {
                t_mask |= uint64_t {1} << b;
                --room;
            }
        };
        for (bool grew = true; grew && room > 0;) {
            const uint64_t before = t_mask;
            for (uint32_t b = 0; b < nvec; ++b) {
                if (in_t(b) && qubit_at[b] != b) {
                    add(qubit_at[b]);   // the home of the qubit on b
                    add(bit_of[b]);     // the bit holding the qubit whose home is b
                }
            }
            grew = t_mask != before;
        }
        // then whole displacement cycles away from the low bits, while they fit
        for (uint32_t b0 = 0; b0 < nvec && room > 0; ++b0) {
            if (in_t(b0) || qubit_at[b0] == b0) {
                continue;
            }
            std::vector<uint32_t> cycle {b0};
            for (uint32_t b = qubit_at[b0]; b != b0; b = qubit_at[b]) {
                cycle.push_back(b);
            }
            int need = 0;
            for (const uint32_t b : cycle) {
                need += in_t(b) ? 0 : 1;
            }
            if (need <= room) {
                for (const uint32_t b : cycle) {
                    add(b);
                }
            }
        }
        bool fixable = false;
        for (uint32_t b = 0; b < nvec; ++b) {
            fixable = fixable || (in_t(b) && qubit_at[b] != b && in_t(qubit_at[b]));
        }
        uint32_t hub_qubit = nvec;   // a displaced qubit to move onto a low bit when none can go home
        if (!fixable) {
            for (uint32_t q = 0; q < nvec && hub_qubit == nvec; ++q) {
                if (!is_home(q) && room > 0) {
                    add(bit_of[q]);
                    hub_qubit = q;
                }
            }
        }
        for (uint32_t b = 0; b < nvec && room > 0; ++b) {
            add(b);   // fill the tile (these qubits stay where they are)
        }
        // the permutation within the tile
        std::vector<uint8_t> nb = bit_of;
        if (hub_qubit != nvec) {
            // nothing can go home: the hub qubit takes low bit 0, whose qubit takes the hub qubit's bit
            const uint32_t ql = qubit_at[0];
            nb[hub_qubit] = 0;
            nb[ql] = bit_of[hub_qubit];
        }
        uint64_t taken = 0;
        std::vector<uint32_t> rest;
        for (uint32_t b = 0; b < nvec && hub_qubit == nvec; ++b) {
            if (!in_t(b)) {
                continue;
            }
            const uint32_t q = qubit_at[b];
            if (in_t(q)) {
                nb[q] = static_cast<uint8_t>(q);
                taken |= uint64_t {1} << q;
            }
            else {
                rest.push_back(q);
            }
        }
        // the rest take the remaining bits in order
        std::vector<uint32_t> moving = rest;
        std::vector<uint32_t> free_bits;
        for (uint32_t b = 0; b < nvec; ++b) {
            if (in_t(b) && !((taken >> b) & 1ull)) {
                free_bits.push_back(b);
            }
        }
        if (hub_qubit == nvec) {
            if (free_bits.size() != moving.size()) {
                throw InvalidArg {"restore plan: inconsistent tile"};
            }
            for (size_t i = 0; i < moving.size(); ++i) {
                nb[moving[i]] = static_cast<uint8_t>(free_bits[i]);
            }
        }
        if (nb == bit_of) {
            throw InvalidArg {"restore plan: no progress"};
        }
        TilePass p;
        p.k = k;
        p.rq = rq;
        for (uint32_t b = 0; b < nvec; ++b) {
            if (in_t(b)) {
                p.qubits.push_back(static_cast<int>(b));
                p.out_qubits.push_back(nb[qubit_at[b]]);
            }
        }
        p.c = std::min(std::countr_one(t_mask), k);
        if (p.k - p.c > 10) {
            throw InvalidArg {"restore plan: tile has more than 10 non-contiguous qubits"};
        }
        p.stages.resize(1);
        complete_stage(p.stages[0], 0, 0, k, rq);
        p.stages[0].g_begin = 0;
        p.stages[0].g_count = 0;
        p.load_direct = true;
        p.store_direct = true;
        p.io_layout = p.stages[0];
        passes.push_back(std::move(p));
        bit_of = nb;
    }
    return passes;
}

namespace
{

TileLayout make_layout(const TileStage& st, const TilePass& p, size_t ab, const std::vector<int>* bits = nullptr,
                       const std::vector<int>* lane_order = nullptr)
{
    TileLayout L {};
    const std::vector<int>& q_of = bits ? *bits : p.qubits;
    const auto hbm = [&](uint32_t local) {
        uint64_t off = 0;
        for (int j = 0; j < p.k; ++j) {
            if ((local >> j) & 1u) {
                off |= uint64_t {1} << q_of[static_cast<size_t>(j)];
            }
        }
        return off;
    };
    for (uint32_t r = 0; r < 16; ++r) {
        uint32_t s_r = 0;
        for (int q = 0; q < p.rq; ++q) {
            if ((r >> q) & 1u) {
                s_r |= 1u << st.reg[q];
            }
        }
        L.lds_slot[r] = static_cast<uint32_t>(lds_swizzle(s_r) * ab);
        L.hbm_slot[r] = hbm(s_r);
    }
    // thread bit j is the j-th local bit of thr_mask, or lane_order[j]
    std::vector<int> order;
    if (lane_order) {
        order = *lane_order;
    }
    else {
        for (int b = 0; b < p.k; ++b) {
            if ((st.thr_mask >> b) & 1u) {
                order.push_back(b);
            }
        }
    }
    for (size_t j = 0; j < order.size(); ++j) {
        const uint32_t b = static_cast<uint32_t>(order[j]);
        L.lds_thr[j] = static_cast<uint32_t>(lds_swizzle(1u << b) * ab);
        L.hbm_thr[j] = hbm(1u << b);
    }
    return L;
}

// A store layout for a pass that moves qubits: its lanes take the local bits in the order of the
// index bits they are stored to, so each wavefront still writes contiguous runs from index bit 0
// up, and the registers take the local bits stored highest.
TileLayout moved_store_layout(const TilePass& p, size_t ab)
{
    std::vector<int> by_bit(static_cast<size_t>(p.k));
    for (int j = 0; j < p.k; ++j) {
        by_bit[static_cast<size_t>(j)] = j;
    }
    std::sort(by_bit.begin(), by_bit.end(), [&](int a, int b) {
        return p.out_qubits[static_cast<size_t>(a)] < p.out_qubits[static_cast<size_t>(b)];
    });
    TileStage st {};
    uint32_t regs = 0;
    for (int i = 0; i < p.rq; ++i) {
        regs |= 1u << by_bit[static_cast<size_t>(p.k - 1 - i)];
    }
    int r = 0;
    for (int b = 0; b < p.k; ++b) {
        if ((regs >> b) & 1u) {
            st.reg[r++] = static_cast<uint8_t>(b);
        }
    }
    st.thr_mask = ((1u << p.k) - 1) & ~regs;
    std::vector<int> lanes(by_bit.begin(), by_bit.end() - p.rq);
    return make_layout(st, p, ab, &p.out_qubits, &lanes);
}

bool same_regs(const TileStage& a, const TileStage& b, int rq)
{
    for (int q = 0; q < rq; ++q) {
        if (a.reg[q] != b.reg[q]) {
            return false;
        }
    }
    return true;
}

uint32_t tile_kind(uint32_t kind)
{
    switch (kind) {
        case GK_DENSE: return TK_DENSE;
        case GK_REAL: return TK_REAL;
        case GK_AXIS: return TK_AXIS;
        case GK_DIAG: return TK_DIAG;
        case GK_DIAG1: return TK_DIAG1;
        case GK_SWAP: return TK_SWAP;
        default: break;
    }
    throw InvalidArg {"tile lowering: unknown gate kind"};
}

int derivative_generator(const double m[8])
{
    const auto is = [&](std::initializer_list<double> want) { return std::equal(want.begin(), want.end(), m); };
    if (is({0, 0, 0, -0.5, 0, -0.5, 0, 0})) {
        return 0;
    }
    if (is({0, 0, -0.5, 0, 0.5, 0, 0, 0})) {
        return 1;
    }
    if (is({0, -0.5, 0, 0, 0, 0, 0, 0.5})) {
        return 2;
    }
    if (is({0, 0, 0, 0, 0, 0, 0, 1.0})) {
        return 3;
    }
    return -1;
}

// One encoded gate record -> kernel variant.  derivative: the record is a derivative dG (never
// given the d0 = 1 treatment, its diagonal may have d0 = 0); on a register target, an angle gate's
// fixed generator gets its own variant (TV_DGEN_BASE), which needs no coefficients.
TileGate lower_gate(const GateRec& g, int precision, bool derivative)
{
    TileGate t {};
    uint32_t creg = 0;
    if (g.control >= 0) {
        const uint32_t src = static_cast<uint32_t>(g.control) & 3u, idx = static_cast<uint32_t>(g.control) >> 2;
        if (src == 0) {
            creg = idx + 1;
        }
        else if (src == 1) {
            t.lane_ctrl |= 1u << idx;
        }
        else {
            t.glb_ctrl |= uint64_t {1} << idx;
        }
    }
    const uint32_t tsrc = g.target & 3u, tidx = g.target >> 2;
    double m[8];
    std::memcpy(m, g.m, sizeof m);
    const bool diag = g.kind == GK_DIAG || g.kind == GK_DIAG1;
    if (tsrc == 0) {
        if (creg == tidx + 1) {
            throw InvalidArg {"tile lowering: control equals target"};
        }
        t.code = tile_code(tile_kind(g.kind), tidx, creg);
        const int gen = derivative ? derivative_generator(g.m) : -1;
        if (gen >= 0) {
            t.code = TV_DGEN_BASE + static_cast<uint32_t>(gen) * 20 + tidx * 5 + creg;
        }
    }
    else {
        if (!diag) {
            throw InvalidArg {"tile lowering: non-diagonal gate outside the registers"};
        }
        t.code = TV_UNIFORM_BASE + creg;
        if (g.kind == GK_DIAG1 && !derivative) {
            // d0 = 1: the gate acts only where the target bit is set, with the factor d1
            if (tsrc == 1) {
                t.lane_ctrl |= 1u << tidx;
            }
            else {
                t.glb_ctrl |= uint64_t {1} << tidx;
            }
            m[0] = m[6];
            m[1] = m[7];
        }
        else if (tsrc == 1) {
            t.lane_sel = 1u << tidx;
        }
        else {
            t.glb_sel = uint64_t {1} << tidx;
        }
    }
    if (precision == KQ_C64) {
        float f[8];
        for (int i = 0; i < 8; ++i) {
            f[i] = static_cast<float>(m[i]);
        }
        std::memset(t.m, 0, sizeof t.m);
        std::memcpy(t.m, f, sizeof f);
    }
    else {
        std::memcpy(t.m, m, sizeof m);
    }
    return t;
}

// the coefficients of a lowered record as doubles
void coef_values(const TileGate& t, int precision, double out[8])
{
    if (precision == KQ_C64) {
        float f[8];
        std::memcpy(f, t.m, sizeof f);
        for (int i = 0; i < 8; ++i) {
            out[i] = f[i];
        }
    }
    else {
        std::memcpy(out, t.m, 8 * sizeof(double));
    }
}

void set_term_coefs(TileTerm& term, int precision, const double d[4])
{
    std::memset(term.d, 0, sizeof term.d);
    if (precision == KQ_C64) {
        const float f[4] = {static_cast<float>(d[0]), static_cast<float>(d[1]), static_cast<float>(d[2]),
                            static_cast<float>(d[3])};
        std::memcpy(term.d, f, sizeof f);
    }
    else {
        std::memcpy(term.d, d, 4 * sizeof(double));
    }
}

// a record's coefficients from doubles (complex64 records hold 8 packed floats)
void store_values(TileGate& g, int precision, const std::array<double, 8>& v)
{
    std::memset(g.m, 0, sizeof g.m);
    if (precision == KQ_C64) {
        float f[8];
        for (int i = 0; i < 8; ++i) {
            f[i] = static_cast<float>(v[static_cast<size_t>(i)]);
        }
        std::memcpy(g.m, f, sizeof f);
    }
    else {
        std::memcpy(g.m, v.data(), sizeof g.m);
    }
}
}  // namespace

void normalise_program(TileProgram& tp, int precision, std::vector<double>* dscale)
{
    if (dscale) {
        dscale->assign(tp.derivs.size(), 1.0);
    }
    if (tp.steps.empty()) {
        throw InvalidArg {"normalise: program without steps"};
    }
    // the vectors grow by 1 / |F| inside the pass: keep F far from underflow (products of two
    // complex64 vectors by 1 / |F|^2 stay inside the FP32 range)
    const double floor = precision == KQ_C64 ? 1.0e-12 : 1.0e-200;
    const auto diag_pivot_of = [](double diag, double off) { return std::abs(diag) >= 1.0e-6 * std::abs(off); };
    std::complex<double> factor {1.0, 0.0};
    const auto fits = [&](std::complex<double> pivot) { return std::abs(pivot) > 0.0 && std::abs(factor * pivot) >= floor; };
    for (const TileStep& st : tp.steps) {
        for (uint32_t gi = st.g_begin; gi < st.g_begin + st.g_count; ++gi) {
            TileGate& g = tp.gates[gi];
            if (g.pad != 0 && dscale) {
                (*dscale)[g.pad - 1] = std::norm(factor);   // |F|^2 where the derivative is taken
            }
            if (g.code >= TV_UNIFORM_BASE || g.lane_ctrl != 0 || g.glb_ctrl != 0 || g.code % 5 != 0) {
                continue;
            }
            const uint32_t kind = g.code / 20, P = (g.code / 5) % 4;
            std::array<double, 8> m;
            coef_values(g, precision, m.data());
            std::array<double, 8> out {};
            uint32_t code = 0;
            std::complex<double> pivot;
            if (kind == TK_REAL) {
                const bool rotation = m[6] == m[0] && m[4] == -m[2];
                const bool reflection = m[6] == -m[0] && m[4] == m[2];
                if (!rotation && !reflection) {
                    continue;
                }
                const bool dp = diag_pivot_of(m[0], m[2]);
                pivot = dp ? m[0] : m[2];
                out[0] = (dp ? m[2] : m[0]) / pivot.real();
                code = TV_NORM_BASE + ((dp ? 0u : 2u) + (rotation ? 0u : 1u)) * 4 + P;
            }
            else if (kind == TK_AXIS) {
                if (m[6] != m[0] || m[5] != m[3]) {
                    continue;
                }
                const bool dp = diag_pivot_of(m[0], m[3]);
                pivot = dp ? m[0] : m[3];
                out[0] = (dp ? m[3] : m[0]) / pivot.real();
                code = TV_NORM_BASE + (dp ? 4u : 5u) * 4 + P;
            }
            else if (kind == TK_DIAG) {
                pivot = {m[0], m[1]};
                if (std::abs(pivot) < 0.5) {
                    continue;
                }
                const std::complex<double> d = std::complex<double> {m[6], m[7]} / pivot;
                out[0] = 1.0;
                out[6] = d.real();
                out[7] = d.imag();
                code = tile_code(TK_DIAG1, P, 0);
            }
            else if (kind == TK_DENSE) {
                const std::complex<double> m00 {m[0], m[1]}, m01 {m[2], m[3]}, m10 {m[4], m[5]}, m11 {m[6], m[7]};
                const bool dp = diag_pivot_of(std::abs(m00), std::abs(m01));
                pivot = dp ? m00 : m01;
                if (std::abs(pivot) == 0.0) {
                    continue;
                }
                const std::complex<double> x = (dp ? m01 : m00) / pivot, y = m10 / pivot, z = m11 / pivot;
                out = {x.real(), x.imag(), y.real(), y.imag(), z.real(), z.imag(), 0.0, 0.0};
                code = TV_NORM_BASE + (dp ? 6u : 7u) * 4 + P;
            }
            else {
                continue;   // swaps stay renames, d0 = 1 diagonals are normalised already
            }
            if (!fits(pivot)) {
                continue;
            }
            factor *= pivot;
            g.code = code;
            store_values(g, precision, out);
        }
    }
    // the pivots' product back into every slot: a uniform diagonal record at the end of the last step
    TileStep& last = tp.steps.back();
    if (last.g_begin + last.g_count != tp.gates.size()) {
        throw InvalidArg {"normalise: the last step does not end the record list"};
    }
    TileGate f {};
    f.code = TV_UNIFORM_BASE;   // every slot, selection bits 0: the factor in the first entry
    store_values(f, precision, {factor.real(), factor.imag(), 0.0, 0.0, 0.0, 0.0, factor.real(), factor.imag()});
    tp.gates.push_back(f);
    ++last.g_count;
}

// Merge diagonal gates of one step.  Diagonal gates commute with each other, and with a non-diagonal
// gate unless it targets their target or register-control position, so every diagonal gate without
// a derivative joins the open run of its key (target position or "uniform", register control);
// a non-diagonal gate on position P first emits the runs that touch P.  A run of one gate is emitted
// unchanged; a longer run becomes one merged record whose terms the kernel multiplies per lane.
void merge_diagonal_runs(const TileGate* gates, uint32_t count, int precision, std::vector<TileGate>& out,
                         std::vector<TileTerm>& terms)
{
    struct Run
    {
        uint32_t key;
        std::vector<const TileGate*> gates;
    };
    std::vector<Run> open;
    const auto emit = [&](const Run& r) {
        if (r.gates.size() == 1) {
            out.push_back(*r.gates[0]);
            return;
        }
        TileGate m {};
        m.code = TV_MERGE_BASE + r.key;
        m.lane_sel = static_cast<uint32_t>(terms.size());
        m.glb_sel = r.gates.size();
        const bool uniform = r.key / 5 == 4;
        for (const TileGate* g : r.gates) {
            TileTerm term {};
            term.lane_ctrl = g->lane_ctrl;
            term.glb_ctrl = g->glb_ctrl;
            double c[8];
            coef_values(*g, precision, c);
            double d[4] = {c[0], c[1], c[6], c[7]};
            if (uniform) {
                term.lane_sel = g->lane_sel;
                term.glb_sel = g->glb_sel;
            }
            else if (g->code / 20 == TK_DIAG1) {
                d[0] = 1.0;
                d[1] = 0.0;
            }
            set_term_coefs(term, precision, d);
            terms.push_back(term);
        }
        out.push_back(m);
    };
    const auto flush_if = [&](auto pred) {
        std::vector<Run> keep;
        for (Run& r : open) {
            if (pred(r.key)) {
                emit(r);
            }
            else {
                keep.push_back(std::move(r));
            }
        }
        open = std::move(keep);
    };
    for (uint32_t i = 0; i < count; ++i) {
        const TileGate& g = gates[i];
        int key = -1;
        if (g.pad == 0) {
            if (g.code < TV_UNIFORM_BASE) {
                const uint32_t kind = g.code / 20;
                if (kind == TK_DIAG || kind == TK_DIAG1) {
                    key = static_cast<int>(((g.code / 5) % 4) * 5 + g.code % 5);
                }
            }
            else if (g.code < TV_MERGE_BASE) {
                key = static_cast<int>(4 * 5 + (g.code - TV_UNIFORM_BASE));
            }
        }
        if (key >= 0) {
            auto it = std::find_if(open.begin(), open.end(), [&](const Run& r) { return r.key == static_cast<uint32_t>(key); });
            if (it == open.end()) {
                open.push_back(Run {static_cast<uint32_t>(key), {&g}});
            }
            else {
                it->gates.push_back(&g);
            }
            continue;
        }
        if (g.code < TV_UNIFORM_BASE) {
            const uint32_t kind = g.code / 20;
            if (kind != TK_DIAG && kind != TK_DIAG1) {
                const uint32_t pos = (g.code / 5) % 4;
                flush_if([&](uint32_t k) { return k / 5 == pos || (k % 5 != 0 && k % 5 - 1 == pos); });
            }
        }
        out.push_back(g);
    }
    flush_if([](uint32_t) { return true; });
}

namespace
{

}  // namespace

// Turn the planner's bit encoding (enc = idx << 2 | src; src 0 register position, 1 thread bit,
// 2 global qubit) into kernel variants, lane / workgroup conditions and per-layout offset tables.
void lower_tile_pass(const TilePass& p, int precision, uint32_t nvec, TileProgram& out,
                     const std::vector<const GateRec*>* deriv_of_src)
{
    if (p.rq != 4 || p.k < 4 || p.k > 12 || p.stages.empty()) {
        throw InvalidArg {"tile lowering: unsupported tile shape"};
    }
    const size_t ab = amp_bytes(precision);
    out = TileProgram {};
    out.k = p.k;
    out.n_thr_bits = p.k - p.rq;
    uint64_t qmask = 0;
    for (const int q : p.qubits) {
        qmask |= uint64_t {1} << q;
    }
    for (uint32_t q = 0; q < nvec; ++q) {
        if (!((qmask >> q) & 1ull)) {
            out.freepos.push_back(static_cast<uint8_t>(q));
        }
    }
    // layouts: [load][one per distinct consecutive stage layout][store]
    const TileStage& first = p.load_direct ? p.stages.front() : p.io_layout;
    out.layouts.push_back(make_layout(first, p, ab));
    const TileStage* cur = &first;
    uint32_t cur_idx = 0;
    for (const TileStage& st : p.stages) {
        if (!same_regs(st, *cur, p.rq)) {
            out.layouts.push_back(make_layout(st, p, ab));
            cur = &st;
            cur_idx = static_cast<uint32_t>(out.layouts.size() - 1);
        }
        TileStep step {};
        step.g_begin = st.g_begin;
        step.g_count = st.g_count;
        step.layout = cur_idx;
        out.steps.push_back(step);
    }
    const bool store_cur = p.store_direct || same_regs(p.io_layout, *cur, p.rq);
    const bool oop = !p.out_free.empty();
    uint64_t dest_free = 0;   // out of place: the destination bits of the outside bits
    if (oop) {
        // every index bit goes to one destination bit: the tile's bits to out_qubits, the others to out_free
        if (p.out_free.size() != out.freepos.size() || p.out_qubits.size() != p.qubits.size()) {
            throw InvalidArg {"tile lowering: out-of-place store bits do not match the tile"};
        }
        uint64_t seen = 0;
        for (const std::vector<int>* v : {&p.out_qubits, &p.out_free}) {
            for (const int b : *v) {
                if (b < 0 || b >= static_cast<int>(nvec) || ((seen >> b) & 1ull)) {
                    throw InvalidArg {"tile lowering: out-of-place store bits are not a permutation"};
                }
                seen |= uint64_t {1} << b;
            }
        }
        for (const int b : p.out_free) {
            dest_free |= uint64_t {1} << b;
        }
        out.out_freepos.assign(p.out_free.begin(), p.out_free.end());
    }
    if (!p.out_qubits.empty() && p.out_qubits != p.qubits) {
        // the store moves qubits to other index bits (of the tile, or out of place of the state): the
        // store layout's HBM offsets follow out_qubits (its LDS offsets are the current layout's when no
        // remap is needed)
        if (!oop && (p.out_qubits.size() != p.qubits.size()
                     || !std::is_permutation(p.out_qubits.begin(), p.out_qubits.end(), p.qubits.begin()))) {
            throw InvalidArg {"tile lowering: store bits are not a permutation of the tile bits"};
        }
        out.layouts.push_back(moved_store_layout(p, ab));
        out.out_layout = static_cast<uint32_t>(out.layouts.size() - 1);
    }
    else if (store_cur) {
        out.out_layout = cur_idx;
    }
    else {
        out.layouts.push_back(make_layout(p.io_layout, p, ab));
        out.out_layout = static_cast<uint32_t>(out.layouts.size() - 1);
    }
    // every layout must address each amplitude of the state exactly once: the slot bits, the
    // thread bits and the free qubits are disjoint and together cover all nvec index bits
    const uint64_t all = nvec >= 64 ? ~0ull : ((uint64_t {1} << nvec) - 1);
    const uint64_t free_bits_src = all & ~qmask;
    const uint32_t tile_bytes = static_cast<uint32_t>((size_t {1} << p.k) * ab);
    for (size_t li = 0; li < out.layouts.size(); ++li) {
        const TileLayout& L = out.layouts[li];
        // an out-of-place store layout addresses the destination, whose outside bits are out_free's
        const uint64_t free_bits = (oop && li == out.out_layout) ? dest_free : free_bits_src;
        uint64_t slot_bits = 0, thr_bits = 0;
        for (int q = 0; q < p.rq; ++q) {
            slot_bits |= L.hbm_slot[1u << q];
        }
        for (int j = 0; j < out.n_thr_bits; ++j) {
            if (std::popcount(L.hbm_thr[j]) != 1 || (thr_bits & L.hbm_thr[j])) {
                throw InvalidArg {"tile lowering: thread bits overlap"};
            }
            thr_bits |= L.hbm_thr[j];
        }
        if ((slot_bits & thr_bits) || ((slot_bits | thr_bits) & free_bits) || ((slot_bits | thr_bits | free_bits) != all)
            || std::popcount(slot_bits) != p.rq) {
            throw InvalidArg {"tile lowering: layout does not cover the state exactly once"};
        }
        for (int r = 0; r < 16; ++r) {
            if (L.lds_slot[r] >= tile_bytes) {
                throw InvalidArg {"tile lowering: LDS offset outside the tile"};
            }
        }
        for (int j = 0; j < out.n_thr_bits; ++j) {
            if (L.lds_thr[j] >= tile_bytes) {
                throw InvalidArg {"tile lowering: LDS offset outside the tile"};
            }
        }
    }
    // gates (and derivative records)
    std::vector<TileGate> lowered;
    lowered.reserve(p.gates.size());
    for (size_t i = 0; i < p.gates.size(); ++i) {
        TileGate t = lower_gate(p.gates[i], precision, false);
        const GateRec* dr = deriv_of_src ? (*deriv_of_src)[p.src[i]] : nullptr;
        if (dr) {
            // the derivative acts on the same target / control bits as the gate
            GateRec dg = *dr;
            dg.target = p.gates[i].target;
            dg.control = p.gates[i].control;
            const bool diag = dg.m[2] == 0.0 && dg.m[3] == 0.0 && dg.m[4] == 0.0 && dg.m[5] == 0.0;
            dg.kind = diag ? GK_DIAG : GK_DENSE;
            TileGate td = lower_gate(dg, precision, true);
            td.pad = static_cast<uint32_t>(out.derivs.size());
            out.derivs.push_back(td);
            out.deriv_src.push_back(p.src[i]);
            t.pad = static_cast<uint32_t>(out.derivs.size());
        }
        lowered.push_back(t);
    }
    // per step: merge runs of commuting diagonal gates
    out.gates.reserve(lowered.size());
    for (size_t si = 0; si < p.stages.size(); ++si) {
        const uint32_t begin = static_cast<uint32_t>(out.gates.size());
        merge_diagonal_runs(lowered.data() + p.stages[si].g_begin, p.stages[si].g_count, precision, out.gates, out.terms);
        out.steps[si].g_begin = begin;
        out.steps[si].g_count = static_cast<uint32_t>(out.gates.size()) - begin;
    }
}

// The adjoint mirror of a forward tile pass: the same tile, registers and layouts walked backwards.
// It loads with the forward pass's store layout (so it reads the amplitudes where the forward pass left
// them, qubit moves included), runs the forward stages in reverse order, each with its gates inverted
// and in reverse order, and stores with the forward pass's load layout, so it leaves every qubit on the
// index bit it had before the forward pass.  Gate i of the pass gets its derivative record from
// deriv_of_rec[p.src[i]] (null: none), as lower_tile_pass does.
void lower_adjoint_mirror(const TilePass& p, int precision, uint32_t nvec, const std::vector<const GateRec*>& deriv_of_rec,
                          TileProgram& out)
{
    if (p.sweep) {
        throw InvalidArg {"adjoint mirror: sweep passes have no tile"};
    }
    TileProgram fwd;
    lower_tile_pass(p, precision, nvec, fwd);
    out = TileProgram {};
    out.k = fwd.k;
    out.n_thr_bits = fwd.n_thr_bits;
    out.freepos = fwd.freepos;
    // layouts: the forward store layout first (the mirror's load), the rest in their order
    const size_t nl = fwd.layouts.size();
    std::vector<uint32_t> idx(nl);
    out.layouts.push_back(fwd.layouts[fwd.out_layout]);
    idx[fwd.out_layout] = 0;
    for (size_t i = 0; i < nl; ++i) {
        if (i != fwd.out_layout) {
            idx[i] = static_cast<uint32_t>(out.layouts.size());
            out.layouts.push_back(fwd.layouts[i]);
        }
    }
    out.out_layout = idx[0];
    out.steps.resize(p.stages.size());
    for (size_t k = 0; k < p.stages.size(); ++k) {
        const size_t si = p.stages.size() - 1 - k;
        const TileStage& st = p.stages[si];
        std::vector<TileGate> lowered;
        lowered.reserve(st.g_count);
        for (uint32_t j = st.g_count; j-- > 0;) {
            const size_t i = st.g_begin + j;
            TileGate t = lower_gate(inverse_of(p.gates[i]), precision, false);
            const GateRec* dr = p.src[i] < deriv_of_rec.size() ? deriv_of_rec[p.src[i]] : nullptr;
            if (dr) {
                GateRec dg = *dr;
                dg.target = p.gates[i].target;
                dg.control = p.gates[i].control;
                const bool diag = dg.m[2] == 0.0 && dg.m[3] == 0.0 && dg.m[4] == 0.0 && dg.m[5] == 0.0;
                dg.kind = diag ? GK_DIAG : GK_DENSE;
                TileGate td = lower_gate(dg, precision, true);
                td.pad = static_cast<uint32_t>(out.derivs.size());
                out.derivs.push_back(td);
                out.deriv_src.push_back(p.src[i]);
                t.pad = static_cast<uint32_t>(out.derivs.size());
            }
            lowered.push_back(t);
        }
        const uint32_t begin = static_cast<uint32_t>(out.gates.size());
        merge_diagonal_runs(lowered.data(), static_cast<uint32_t>(lowered.size()), precision, out.gates, out.terms);
        out.steps[k].g_begin = begin;
        out.steps[k].g_count = static_cast<uint32_t>(out.gates.size()) - begin;
        out.steps[k].layout = idx[fwd.steps[si].layout];
    }
}

}  // namespace kq
