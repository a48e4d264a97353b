// Internal declarations shared by the libkq translation units (host side + kernel launchers).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <vector>

#include "../../include/kq.h"

namespace kq
{

// ---- errors -------------------------------------------------------------------------------
// Every C entry point runs its body through guarded(); these exception types map to status codes.
struct InvalidArg : std::runtime_error { using std::runtime_error::runtime_error; };
struct HipError : std::runtime_error { using std::runtime_error::runtime_error; };
struct AllocError : std::runtime_error { using std::runtime_error::runtime_error; };
// kettle-level errors keep kettle's exception type: std::runtime_error / std::out_of_range

void set_error(const std::string& msg);

#define KQ_HIP_CHECK(expr)                                                                    \
    do {                                                                                      \
        hipError_t kq_e_ = (expr);                                                            \
        if (kq_e_ != hipSuccess) {                                                            \
            throw ::kq::HipError {std::string {#expr} + ": " + hipGetErrorString(kq_e_)};     \
        }                                                                                     \
    } while (0)

// ---- device gate records --------------------------------------------------------------------
// A compiled gate acting on one target bit of the state vector (a density matrix is a vector of
// 2n qubits: each kettle gate becomes two records, U on bit t and conj(U) on bit n + t).
enum GateKind : uint32_t
{
    GK_DENSE = 0,   // (a, b) -> (m00 a + m01 b, m10 a + m11 b), general complex matrix
    GK_DIAG = 1,    // a -> d0 a, b -> d1 b          (m[0..1] = d0, m[6..7] = d1)
    GK_SWAP = 2,    // (a, b) -> (b, a)
    GK_REAL = 3,    // real matrix (H, RY): half the multiplies of GK_DENSE
    GK_AXIS = 4,    // real diagonal, imaginary off-diagonal (RX, Y): half the multiplies
    GK_DIAG1 = 5,   // diagonal with d0 == 1 exactly (Z, S, T, P and their controlled forms)
};

struct GateRec
{
    uint32_t kind;
    uint32_t target;
    int32_t control;    // -1: no control
    uint32_t pad;
    double m[8];        // m00, m01, m10, m11 as (re, im)
};

// One fused pass.  The tile is the 2^k amplitudes sharing the values of every qubit outside the
// tile set; local bit j of a tile index is global qubit `qubits[j]`, and local bits 0..c-1 are
// the contiguous global qubits 0..c-1.  Inside the tile each thread holds 2^rq amplitudes in
// registers (rq "register bits" of the local index); the pass runs as stages, each with its own
// register bits, and the tile moves through LDS only between stages whose register bits differ.
struct TileStage
{
    uint8_t reg[8];          // local bit held by register position p (ascending), p < rq
    uint32_t thr_mask;       // local bits that form the thread index (ascending)
    uint32_t g_begin;        // gate range of the stage in the pass's gate array
    uint32_t g_count;
    uint32_t pad;
};

// Gate records of a tile pass encode where each index bit comes from:
//   enc = (idx << 2) | src,  src 0: register position idx, 1: thread-index bit idx,
//                             2: global qubit idx (outside the tile, uniform per tile)
// target: the encoded target bit (non-diagonal gates always have src 0); control: -1 or encoded.
struct TilePass
{
    int k = 0;                   // tile qubits
    int c = 0;                   // low contiguous qubits 0..c-1 in the tile
    int rq = 0;                  // register bits per thread
    std::vector<int> qubits;     // qubits[j] = index bit of local bit j (ascending)
    std::vector<int> out_qubits; // index bit local bit j is stored to (empty: qubits[j]); in place a
                                 // permutation of `qubits` that moves qubits onto the low bits for the
                                 // next pass, out of place any bits (with out_free, a permutation of all)
    std::vector<int> out_free;   // out of place (non-empty): the pass stores into a second buffer, and
                                 // the i-th index bit outside the tile (ascending) goes to out_free[i]
    std::vector<GateRec> gates;  // tile-encoded records
    std::vector<TileStage> stages;
    bool load_direct = false;    // load HBM straight into stage 0's layout
    bool store_direct = false;   // store HBM straight from the last stage's layout
    TileStage io_layout {};      // layout used for HBM I/O when not direct
    bool sweep = false;          // single gate: run the plain sweep kernel on `single`
    GateRec single {};
    std::vector<uint32_t> src;   // input record index of each applied gate, in execution order
};

// ---- lowered tile program (what the tile kernel reads) ----------------------------------------
// Gate variants of the tile kernel.  A gate whose target is register position P (0..3) runs the
// straight-line code of variant (kind, P, C), C = 0 for no register control or 1 + the register
// position of its control; a diagonal gate whose target bit is a thread or global bit scales
// every slot (or the slots whose register control is set) by one factor per lane: variant
// TV_UNIFORM_BASE + C.
enum TileKind : uint32_t { TK_DENSE = 0, TK_REAL, TK_AXIS, TK_DIAG, TK_DIAG1, TK_SWAP, TK_COUNT };
constexpr uint32_t tile_code(uint32_t kind, uint32_t pos, uint32_t ctrl) { return (kind * 4 + pos) * 5 + ctrl; }
constexpr uint32_t TV_UNIFORM_BASE = TK_COUNT * 4 * 5;
// Merged diagonal run: variant TV_MERGE_BASE + t * 5 + C multiplies, per lane, the factors of a list
// of diagonal terms (TileTerm) sharing the target t (register position 0..3, or 4 = a thread or
// outside bit, i.e. every slot) and the register control C, then applies them in one step.
constexpr uint32_t TV_MERGE_BASE = TV_UNIFORM_BASE + 5;
constexpr uint32_t TV_END = TV_MERGE_BASE + 25;
// Derivative records only: Re <w| Q |v> for the fixed generators of the angle gates (dG/dtheta = Q G):
// variant TV_DGEN_BASE + gen * 20 + P * 5 + C with gen 0: (-i/2) X (RX), 1: (-i/2) Y (RY), 2: (-i/2) Z
// (RZ), 3: i |1><1| (P), on register position P with register control C; no coefficients needed.
constexpr uint32_t TV_DGEN_BASE = TV_END;
constexpr uint32_t TV_DGEN_END = TV_DGEN_BASE + 4 * 20;
// the generator a derivative matrix is (0..3, as above), or -1
int derivative_generator(const double m[8]);
// Uncontrolled one-qubit gates with their pivot entry divided out, for the interpreted tile kernel:
// variant TV_NORM_BASE + form * 4 + P on register position P; forms 0..3 real rotation / reflection
// (t in the first coefficient: [[1, t], [-t, 1]], [[1, t], [t, -1]], [[t, 1], [-1, t]], [[t, 1], [1, -t]]),
// 4..5 axis rotation ([[1, i t], [i t, 1]], [[t, i], [i, t]]), 6..7 complex ([[1, x], [y, z]],
// [[x, 1], [y, z]], x, y, z in the first six).
constexpr uint32_t TV_NORM_BASE = TV_DGEN_END;
constexpr uint32_t TV_NORM_END = TV_NORM_BASE + 8 * 4;

// one diagonal gate inside a merged run: its conditions, selection and the two diagonal entries
struct TileTerm
{
    uint32_t lane_ctrl;
    uint32_t lane_sel;
    uint64_t glb_ctrl;
    uint64_t glb_sel;
    double d[4];             // d0 (re, im), d1 (re, im): doubles, or 4 floats first (complex64)
    uint64_t pad;
};

struct TileGate
{
    uint32_t code;
    uint32_t lane_ctrl;      // thread-index bits that must all be 1 (controls on thread bits)
    uint32_t lane_sel;       // uniform-diagonal: the thread-index bit selecting d1 (0: see glb_sel);
                             // merged run: index of its first term
    uint32_t pad;            // adjoint: derivative record index + 1 (0: none)
    uint64_t glb_ctrl;       // amplitude-index bits outside the tile that must all be 1
    uint64_t glb_sel;        // uniform-diagonal: the index bit outside the tile selecting d1;
                             // merged run: number of terms
    double m[8];             // coefficients: 8 doubles (complex128) or 8 floats first (complex64)
};

// Where a register layout puts the tile: slot r of thread t holds local index S_r | T_t.  Both
// the LDS swizzle and the global index are XOR-linear in the local index, so the address of a
// slot is (thread part) ^ (slot part) with the slot parts uniform.
struct TileLayout
{
    uint32_t lds_slot[16];   // swizzled LDS byte offset of S_r
    uint32_t lds_thr[8];     // swizzled LDS byte offset of thread bit j
    uint64_t hbm_slot[16];   // amplitude offset of S_r in the state vector
    uint64_t hbm_thr[8];     // amplitude offset of thread bit j
};

struct TileStep
{
    uint32_t g_begin;
    uint32_t g_count;
    uint32_t layout;         // index into the pass's layouts (0 = load layout)
    uint32_t pad;
};

// One state of a batched tile launch (device memory; pointers are device addresses).
struct BatchTileItem
{
    uint64_t a_off;          // amplitude offset of the state (psi) in the first arena
    uint64_t b_off;          // amplitude offset of lambda in the second arena (adjoint launches)
    uint64_t steps, layouts, gates, derivs, partials, terms;
    uint64_t free_mask;
    int32_t n_thr_bits;
    int32_t n_steps;
    uint32_t out_layout;
    uint32_t n_part;
    uint32_t block_begin;
    uint32_t input;          // forward items: 0 load the state; 1 the state is |0...0>, not written yet (the
                             // workgroup of tile 0 builds it, the others store zeros); 2 skip the tiles whose
                             // outside bits under zmask are not all 0 (they hold zeros already)
    uint64_t zmask;
};

struct TileProgram
{
    int k = 0;
    int n_thr_bits = 0;
    uint32_t out_layout = 0;
    std::vector<uint8_t> freepos;     // global qubits outside the tile, ascending
    std::vector<TileGate> gates;
    std::vector<TileStep> steps;
    std::vector<TileLayout> layouts;
    std::vector<TileGate> derivs;     // adjoint passes: derivative records (pad = partial slot)
    std::vector<TileTerm> terms;      // terms of merged diagonal runs
    std::vector<uint32_t> deriv_src;  // input record index of each derivative slot
    std::vector<uint8_t> out_freepos; // out of place: destination bit of each freepos entry (empty: in place)
};

// Rewrites the uncontrolled one-qubit records of a lowered pass (for the interpreted tile kernel) in
// the TV_NORM forms, a diagonal as d0 = 1 with d1 / d0, and appends to its last step one uniform record that
// multiplies every slot by F, the product of the pivots taken out; inside the pass the vectors are the
// true ones divided by F, so a derivative record reduces Re <w|Q|v> / |F|^2: dscale (optional) receives
// |F|^2 for each derivative slot.  Always appends exactly one record (F = 1 when nothing is rewritten),
// so a program's size does not depend on its angles.
void normalise_program(TileProgram& tp, int precision, std::vector<double>* dscale);

// Lower a planned pass into the tile kernel's program (coefficients rounded to the precision).
// deriv_of_src (adjoint passes): for input record i, its derivative dG (or null); a gate with one
// gets TileGate::pad = derivative index + 1 and its derivative record the next partial slot.
void lower_tile_pass(const TilePass& p, int precision, uint32_t nvec, TileProgram& out,
                     const std::vector<const GateRec*>* deriv_of_src = nullptr);
// The adjoint reverse of a forward tile pass on the same tile: loads with its store layout, runs its
// stages backwards with inverted gates (and derivative records from deriv_of_rec, indexed by the
// pass's src), stores with its load layout.
void lower_adjoint_mirror(const TilePass& p, int precision, uint32_t nvec, const std::vector<const GateRec*>& deriv_of_rec,
                          TileProgram& out);
// merge runs of commuting diagonal gates of one step (appends to out / terms)
void merge_diagonal_runs(const TileGate* gates, uint32_t count, int precision, std::vector<TileGate>& out,
                         std::vector<TileTerm>& terms);
// Default LDS placement of a tile: entry of local index l (its own inverse; linear over GF(2)).
inline uint32_t lds_swizzle(uint32_t l) { return l ^ (((l >> 4) ^ (l >> 8)) & 15u); }

// throws unless every step and gate record of the program is one the tile kernel runs
void check_tile_records(const TileProgram& prog);
uint32_t tile_pass_blocks(const TileProgram& prog);
uint32_t tile_pass_waves(const TileProgram& prog);

struct PlanStats
{
    size_t n_gates = 0;
    size_t n_passes = 0;
    size_t n_tile_passes = 0;
};

// ---- states ---------------------------------------------------------------------------------
}  // namespace kq

struct kq_state
{
    uint32_t n = 0;          // qubits of the simulated system
    uint32_t nvec = 0;       // qubits of the stored vector (n, or 2n for a density matrix)
    int precision = KQ_C128;
    int is_dm = 0;
    uint64_t n_amps = 0;     // 2^nvec
    void* amps = nullptr;    // device buffer of n_amps complex values
    hipStream_t stream = nullptr;
    // per-state device scratch (reductions, CDFs), grown on demand
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // per-state device buffer for uploaded gate programs, and its pinned host staging buffer
    void* prog = nullptr;
    size_t prog_bytes = 0;
    void* stage = nullptr;
    size_t stage_bytes = 0;
    // per-state device buffer for adjoint partial sums
    void* aux = nullptr;
    size_t aux_bytes = 0;
    // per-state workspace of the adjoint gradient (lambda, the reverse-sweep psi, checkpoints, derivative
    // partials, reverse-pass programs), grown when a call needs more and kept for the next call
    void* adj = nullptr;
    size_t adj_bytes = 0;
    // qubit q of the state is stored on index bit bit_of[q] (empty: bit q); only gate passes of a
    // statevector leave it permuted, every other operation restores bit q first
    std::vector<uint8_t> bit_of;
    // a second buffer of the state's size: the destination of passes that store out of place (the
    // buffers then trade places); allocated by the first such pass and kept until the state is destroyed
    void* spare = nullptr;
    // >= 0: the state is the basis vector of this stored index, not yet written to `amps` (a reset
    // waits for the first operation: a gate run builds its first pass's tiles from the index, anything
    // else writes the vector first)
    int64_t lazy_basis = -1;
    // Known zeros: every amplitude whose stored index has (index & zero_mask) != zero_val is 0 (a basis
    // state: every bit known; each tile pass clears its tile's bits).  Gate runs skip the zero tiles;
    // any other operation forgets it (zero_mask = 0).
    uint64_t zero_mask = 0;
    uint64_t zero_val = 0;
};

namespace kq
{

inline size_t amp_bytes(int precision) { return precision == KQ_C64 ? 8 : 16; }

void* scratch(kq_state* s, size_t bytes);

// ---- pass profiling (HIP events on the state's stream) ------------------------------------------
struct PassProfiler
{
    bool enabled = false;
    kq_pass_stats cls[KQ_PASS_CLASSES] {};
    // events recorded but not yet read: (class, start, stop, bytes, gates)
    struct Pending { int cls; hipEvent_t a; hipEvent_t b; double bytes; uint64_t gates; };
    std::vector<Pending> pending;
};
PassProfiler& profiler();
// bracket one launch; `launch` is invoked between the two events
template <typename F>
void profiled(kq_state* s, int cls, double bytes, uint64_t gates, F&& launch)
{
    PassProfiler& p = profiler();
    if (!p.enabled) {
        launch();
        return;
    }
    hipEvent_t a, b;
    KQ_HIP_CHECK(hipEventCreate(&a));
    KQ_HIP_CHECK(hipEventCreate(&b));
    KQ_HIP_CHECK(hipEventRecord(a, s->stream));
    launch();
    KQ_HIP_CHECK(hipEventRecord(b, s->stream));
    p.pending.push_back({cls, a, b, bytes, gates});
}
void profiler_collect();   // read and release pending events (after a stream sync)

// ---- gate compilation and planning (kq_plan.cpp) --------------------------------------------
// kettle gate (kq_op) -> one (statevector) or two (density matrix) device records
void compile_gate(const kq_op& op, uint32_t n, bool is_dm, std::vector<GateRec>& out);
// statevector record whose kind depends on the gate type only for parameterised gates (aux >= 0)
void compile_gate_stable(const kq_op& op, uint32_t n, std::vector<GateRec>& out);
bool is_diag(const GateRec& g);
// Plan fused tile passes over a straight-line list of records.
// allow_sweep: a pass holding a single diagonal gate may run as a plain sweep (off for adjoint plans)
// With `emit`, each pass is handed over as soon as it is planned and the returned list is empty.
// final_bits (with bit_of): the placement the last pass should leave (qubit -> index bit).  The last
// pass then takes in the bits it needs to store every qubit there, in place when the displaced qubits'
// bits and destinations fit its tile, else (out_of_place) into a second buffer when the bits that hold
// the destinations' low qubits fit; otherwise the placement stays as planned.
std::vector<TilePass> plan_passes(const std::vector<GateRec>& gates, uint32_t nvec, int max_k,
                                  int low_c, int max_gates, int rq, int max_stages, bool allow_sweep = true,
                                  std::vector<uint8_t>* bit_of = nullptr,
                                  const std::function<void(TilePass&&)>* emit = nullptr,
                                  const std::vector<uint8_t>* final_bits = nullptr, bool out_of_place = false);
// gate-free tile passes that store every qubit q back on index bit q
std::vector<TilePass> plan_restore(std::vector<uint8_t> bit_of, uint32_t nvec, int max_k, int low_c, int rq);
// whether statevector gate passes may move qubits onto other index bits (kq_set_fusion_relabel)
bool fusion_relabel();
void set_fusion_relabel(bool on);
bool fusion_autotune();
void set_fusion_autotune(bool on);
int pass_remaps(const TilePass& p);
int fusion_max_stages();
// qubits 0..c-1 always belong to a tile (contiguous 2^c-amplitude chunks per access)
int fusion_low_qubits();
void set_fusion_low_qubits(int c);
int register_qubits(int precision);
PlanStats& last_stats();
int fusion_max_k(int precision, uint32_t nvec);
int fusion_max_k_for(int precision, uint32_t nvec, size_t lds_bytes);
// LDS bytes one workgroup may use on the current device (queried at initialisation)
size_t& device_lds_per_block();
constexpr size_t TILE_STATIC_LDS = 0;      // static LDS of the tile kernel besides the tile
constexpr size_t TILE_LDS_ENTRY = 8;       // LDS bytes per amplitude of the tile kernel's remaps (complex128: halves)
// derivative staging of the adjoint tile kernel: TILE_STAGE derivatives x 64 lanes of doubles and their
// slots, per whole wavefront, after the tile in LDS
constexpr uint32_t TILE_STAGE = 8;
constexpr uint32_t TILE_STAGE_BYTES = TILE_STAGE * 64 * 8 + 256;
size_t adjoint_tile_lds(int k);
int fusion_max_gates();
void set_fusion(int max_k, int max_gates, int max_stages, double max_cost);
double fusion_max_cost();
// arithmetic cost of a gate in a tile pass, in units of one general complex 2x2 gate
double gate_cost(const GateRec& g);
int fusion_setting_k();

// ---- kernel launchers (kq_kernels.hip) --------------------------------------------------------
void launch_gate_sweep(kq_state* s, const GateRec& g);
// d_prog: the uploaded [gates][steps][layouts][terms] of `prog`
void launch_tile_pass(kq_state* s, const TileProgram& prog, const void* d_gates, const void* d_steps,
                      const void* d_layouts, const void* d_terms);
// adjoint reverse-sweep pass over (psi, lam); partials: [slot][blocks * waves] doubles
void launch_adjoint_tile_pass(kq_state* s, void* psi, void* lam, const TileProgram& prog, const void* d_gates,
                              const void* d_steps, const void* d_layouts, const void* d_derivs, const void* d_terms,
                              double* d_partials);
// many states in one launch (k: tile qubits of every item; items sorted by block_begin)
void launch_tile_batch(hipStream_t stream, int precision, bool adjoint, void* arena_a, void* arena_b,
                       const BatchTileItem* d_items, uint32_t n_items, uint32_t n_blocks, int k);
// out[i] = sum of rows[offs[i] .. offs[i] + counts[i]) in a fixed order (one row per slot)
void launch_sum_rows(hipStream_t st, const double* rows, const uint64_t* d_offs, const uint32_t* d_counts, uint32_t n,
                     double* d_out);
void launch_set_basis(kq_state* s, uint64_t index);
// weights: SV -> (sum |a|^2 over bit q = 0, = 1); DM -> block Frobenius norms (kettle's quirk)
void launch_qubit_weights(kq_state* s, uint32_t q, double out[2]);
void launch_collapse(kq_state* s, uint32_t q, int kept, double scale);
// probabilities into a device array of N doubles (|a|^2 or Re diag rho); noise per qubit
void launch_probabilities(kq_state* s, double* d_probs, const double* noise_host);
void launch_noise(hipStream_t st, double* d_probs, uint32_t n, uint32_t q, double noise);
// inclusive scan of d_probs in place; returns last value and the largest value < last
void launch_cdf(hipStream_t st, double* d_cdf, uint64_t n, void* scratch_ptr, size_t scratch_bytes,
                double* last, double* second_last);
// per-shot index; cdf_scratch: the scratch launch_cdf used (its positive-probability bits)
void launch_lower_bound(hipStream_t st, const double* d_cdf, uint64_t n, const double* d_u, uint64_t shots, void* cdf_scratch,
                        uint64_t* d_out);
size_t cdf_scratch_bytes(uint64_t n);

// Fixed device workspace of the allocation-free sampler (kq_sample_probabilities_noalloc): up to
// NOALLOC_MAX_PROBS probabilities (scanned into their CDF in place), the scan partials, and up to
// NOALLOC_MAX_SHOTS uniform draws and drawn indices.  One copy per device, held until the process ends.
constexpr uint64_t NOALLOC_MAX_PROBS = uint64_t {1} << 20;
constexpr uint64_t NOALLOC_MAX_SHOTS = uint64_t {1} << 20;
constexpr size_t NOALLOC_CDF_BYTES = (NOALLOC_MAX_PROBS / 8) + 8192;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_MAX_PROBS * 8 + NOALLOC_CDF_BYTES + NOALLOC_MAX_SHOTS * 16;
struct NoallocWorkspace
{
    void* ptr;               // device address of the workspace on the current device
    hipStream_t stream;      // the stream its calls run on (one per device)
};
// the current device's workspace: its address (hipGetSymbolAddress) and stream, looked up and created
// at the first call on that device
NoallocWorkspace noalloc_workspace();

struct PauliGroup
{
    uint64_t x;          // flip mask
    uint32_t first;      // first string of the group in the string table
    uint32_t count;
};
struct PauliZ
{
    uint64_t z;          // sign mask
    double w[2];         // total complex weight (coefficient * phase * canonical factor)
};
// Pauli operator in device form: strings grouped by X/Y flip mask (kq_pauli_expectation's packing)
void build_pauli(uint32_t n, size_t n_strings, const uint32_t* off, const uint32_t* q, const uint8_t* k,
                 const uint8_t* phases, const double* coeffs, std::vector<PauliGroup>& groups,
                 std::vector<PauliZ>& strings);
// adjoint helpers: G^dagger, and dG/dtheta of an angle gate (false for gates without an angle)
GateRec inverse_of(const GateRec& g);
bool derivative_of(const kq_op& op, const GateRec& g, GateRec& d);

// One state of a batched Pauli / read-out launch (device memory).
struct BatchPauliItem
{
    uint64_t a_off;          // amplitude offset of psi
    uint64_t b_off;          // amplitude offset of lambda
    uint64_t n_amps;
    uint64_t groups, zs;     // device addresses of the state's PauliGroup / PauliZ tables
    uint64_t part_off;       // first partial of the state
    uint32_t n_groups;
    uint32_t block_begin;
};
uint32_t batch_blocks(uint64_t n_amps);   // blocks of one state in a batched Pauli / read-out launch
// lambda = H psi per state and per-block partials of Re <psi|lambda>
void launch_pauli_batch(hipStream_t st, int in_precision, int out_precision, const void* arena_a, void* arena_b,
                        const BatchPauliItem* d_items, uint32_t n_items, uint32_t n_blocks, double* d_part);
// complex128 -> complex64 copy of n amplitudes
void launch_narrow(hipStream_t st, const void* in, void* out, uint64_t n);
// per-block partials of max |psi_i - delta_{i0}|
void launch_basis_distance_batch(hipStream_t st, int precision, const void* arena, const BatchPauliItem* d_items,
                                 uint32_t n_items, uint32_t n_blocks, double* d_part);
// every state of the arena to |0...0> (zero fill, then a one at each state's offset)
void launch_batch_reset(hipStream_t st, int precision, void* arena, uint64_t n_amps, const uint64_t* d_offs, uint32_t n);
void launch_max_rows(hipStream_t st, const double* rows, const uint64_t* d_offs, const uint32_t* d_counts, uint32_t n,
                     double* d_out);

// sum_i conj(a_i) * sum_g (sum_{s in g} w_s (-1)^{|i & z_s|}) a_{i ^ x_g}
void launch_pauli_expectation(kq_state* s, const std::vector<PauliGroup>& groups, const std::vector<PauliZ>& strings,
                              double out[2]);
// b_i = sum_g (sum_{s in g} w_s (-1)^{|i & z_s|}) a_{i ^ x_g}  (b = H a), same precision as a
// H psi runs over tiles of 2^PAULI_TILE_BITS amplitudes once the state holds at least 16 of them
constexpr int PAULI_TILE_BITS = 12;
// stable order of the groups by their flip mask's bits above the tile (the partner tile), so that
// the tiled H psi stages each distinct partner tile once; the sums' terms only change order
inline void sort_pauli_groups_by_tile(std::vector<PauliGroup>& groups)
{
    std::stable_sort(groups.begin(), groups.end(), [](const PauliGroup& l, const PauliGroup& r) {
        return (l.x >> PAULI_TILE_BITS) < (r.x >> PAULI_TILE_BITS);
    });
}
void launch_pauli_apply_sum(kq_state* s, const void* d_in, void* d_out, const std::vector<PauliGroup>& groups,
                            const std::vector<PauliZ>& strings);
// in-place (c, x, z) operator: a_i <- w (-1)^{|i & z|} a_{i ^ x}
void launch_pauli_apply_one(kq_state* s, uint64_t x, uint64_t z, const double w[2]);

void launch_dm_trace(kq_state* s, double out[2]);
// tr(rho^2) = sum_{r,c} rho(r,c) rho(c,r), FP64
void launch_dm_trace_of_square(kq_state* s, double out[2]);
void launch_dm_partial_trace(kq_state* s, const std::vector<uint32_t>& traced, double* out_host);
void launch_dm_from_sv(kq_state* dm, const kq_state* sv);

void launch_pair_map_1q(uint32_t n, uint32_t t, uint64_t* d_s0, uint64_t* d_s1);
void launch_pair_map_2q(uint32_t n, uint32_t c, uint32_t t, int four_way, uint64_t* d_out);

// adjoint sweep helpers (kq_adjoint.cpp uses these)
// One gate on two vectors: psi <- G^dag psi, then <lam| dG psi> accumulated (if dG != null),
// then lam <- G^dag lam.  Result written to d_partial[slot] as (re, im) after reduction.
void launch_adjoint_gate(kq_state* s, void* psi, void* lam, const GateRec& inv, const GateRec* deriv,
                         double* d_acc);
void launch_copy(kq_state* s, void* dst, const void* src, size_t bytes);
void launch_copy_kernel(kq_state* s, void* dst, const void* src, uint64_t n_amps);
// GB/s of plain 16-byte streams over 2^n_qubits elements: read, write, in place, copy
void stream_probe(uint32_t n_qubits, double gbps[24]);   // [launch shape][read, write, in place, copy]
// psi[r, :] <- U psi[r, :] over index bits 0..3 (kq_dense.hip); d_u: 16 x 16 complex in the state's precision
void launch_dense4(kq_state* s, const void* d_u, bool use_mfma);
// one wave, one 16x16x4 matrix instruction (FP64 for KQ_C128, FP32 otherwise): d[4 l + e] = element e of lane l
void launch_mfma16_layout(hipStream_t st, int precision, const void* a, const void* b, void* d);   // [launch shape][read, write, in place, copy]

// Tables of the one-pass index-bit permutation (k_permute_bits): tile local bit j is source index bit
// src_of[j] (a single-bit mask) and destination bit dst_of[j]; block bit i is source bit free_src[i]
// and destination bit free_dst[i]; the load and store layouts give the local bit of each thread bit
// and register bit.
struct PermuteTables
{
    uint64_t src_of[12];
    uint64_t dst_of[12];
    uint8_t free_src[52];
    uint8_t free_dst[52];
    uint32_t n_free;
    uint8_t ld_thr[8];
    uint8_t ld_reg[4];
    uint8_t st_thr[8];
    uint8_t st_reg[4];
};
// dst[pi(i)] = src[i], pi the bit permutation the tables describe (s->n_amps amplitudes, nvec >= 12)
void launch_permute_bits(kq_state* s, const void* src, void* dst, const PermuteTables& t);
double launch_max_abs_diff(kq_state* s, const void* a, const void* b);
// Device-side drift check of the adjoint sweep: *d_drift = max |psi_i - ckpt_i| (block partials in
// d_part, RED_BLOCKS entries), then psi <- ckpt when it exceeds tol, counted in *d_count; no host sync.
constexpr int REDUCTION_BLOCKS = 2048;
void launch_drift_restore(kq_state* s, void* psi, const void* ckpt, double tol, double2* d_part, double* d_drift,
                          uint32_t* d_count);
void launch_dot(kq_state* s, const void* a, const void* b, double out[2]);
// <a|b> in FP64 into *d_out on the device (block partials in d_part, REDUCTION_BLOCKS entries), no sync
void launch_dot_to_device(kq_state* s, const void* a, const void* b, double2* d_part, double2* d_out);

// ---- tile passes compiled at run time (kq_jit.cpp) ---------------------------------------------------
struct JitStats
{
    size_t compiled = 0;       // kernels compiled
    size_t hits = 0;           // passes served from the cache
    size_t failures = 0;       // compilations that failed (the pass then runs interpreted)
    double compile_seconds = 0.0;
    size_t uncapped = 0;       // kernels that spilled under the register cap, compiled again without it
    std::string last_log;
};
JitStats& jit_stats();
void set_jit_mode(int mode);   // 0 off, 1 states of >= 18 qubits, 2 every tile pass
int jit_mode();
// min_waves_per_simd > 0 caps the generated kernels' registers so that many waves fit per SIMD;
// blocks_per_cu > 0 launches that many blocks per CU, each walking several tiles (0: one block per tile)
struct JitTuning
{
    int min_waves_per_simd = 3;   // register cap (<= 170 VGPRs): at least 3 waves per SIMD
    int blocks_per_cu = 0;
    bool normalise = true;   // uncontrolled one-qubit gates with their pivot divided out
    int priority = 1;        // wave priority around HBM traffic: 0 off, 1 loads, 2 loads and stores
    int nontemporal = 3;     // nontemporal hint on HBM traffic: 0 off, 1 loads, 2 loads and stores,
                             // 3 loads and stores of vectors larger than the 256 MiB Infinity Cache
};
// the nontemporal mode (0..2) a pass over vectors of 2^nvec amplitudes of `amp_bytes` bytes runs with
inline int nontemporal_mode(int mode, uint32_t nvec, size_t amp_bytes)
{
    return mode < 3 ? mode : ((amp_bytes << nvec) > (size_t {256} << 20) ? 2 : 0);
}
void set_jit_tuning(const JitTuning& t);
JitTuning jit_tuning();
uint32_t device_cu_count();
uint32_t& device_cus();
bool jit_wanted(const kq_state* s);
bool jit_available();
// What a forward pass kernel starts from: the state (load); a basis state not yet written, built from
// its index (basis: the kernel's last argument); or the state with known-zero tiles, those whose outside
// bits under a mask differ from given values (skip_zero: the last two arguments, zmask and zval), which
// are skipped (in place: the buffer holds the zeros) or stored as zeros (out of place).
enum class PassInput : int { load = 0, basis = 1, skip_zero = 2 };
// kernel source of one forward pass; coefs receives its coefficient buffer
std::string jit_source(const TileProgram& tp, int precision, std::vector<double>& coefs, PassInput input = PassInput::load);
// The known zeros of a state followed through a run of passes in launch order: every amplitude whose
// stored index has (index & mask) != val is 0; basis: the state is that basis vector, not written yet.
struct KnownZeros
{
    uint64_t mask = 0;
    uint64_t val = 0;
    bool basis = false;
    // the input of a tile pass with this program (first: the run's first pass; compiled: it runs as a
    // compiled kernel without the persistent form); zmask / zval: the known outside bits for skip_zero
    PassInput input_for(const TileProgram& tp, bool first, bool compiled, uint64_t& zmask, uint64_t& zval) const;
    // after a tile pass: its tile bits are unknown, an out-of-place pass moves the outside bits
    void after(const TileProgram& tp);
    // after anything else
    void forget()
    {
        mask = 0;
        val = 0;
        basis = false;
    }
};
// compiled kernels for the sources (null where compilation failed); compiles missing ones in parallel
std::vector<hipFunction_t> jit_kernels(const std::vector<std::string>& sources);
// the compiled kernel of a source if it is cached (a cache hit), else null; never compiles
hipFunction_t jit_lookup(const std::string& source);
// compile without loading (no device needed): number of failures, last log in *log
size_t jit_compile_only(const std::vector<std::string>& sources, std::string* log);
// out-of-place programs (tp.out_freepos) store into s->spare, which must be allocated; extra: the values
// of the kernel's input mode (PassInput::basis: the stored basis index; skip_zero: zmask, zval)
void launch_jit_pass(kq_state* s, hipFunction_t fn, const TileProgram& tp, const double* d_coefs,
                     const uint64_t* extra = nullptr, size_t n_extra = 0);
// kernel source of one adjoint reverse pass over (psi, lambda) with derivative reductions (the records of
// an adjoint-lowered program); coefs receives its coefficient buffer
// What the generated kernel of an adjoint reverse pass reads besides its structure: the coefficient
// buffer; per derivative record, the factor its reduced partials need (uncontrolled one-qubit gates run
// with their pivot divided out, so inside the pass psi and lambda are the true vectors divided by the
// product F of the pivots so far, and Re <lambda|Q|psi> comes out divided by |F|^2); and the choice of
// form of every such gate (which depends on the coefficients, and on which the source depends).
struct AdjointCoefs
{
    std::vector<double> coefs;
    std::vector<double> dscale;     // per derivative record: multiply its reduced partials by this
    std::vector<std::string> forms; // per record: its normalised statements (empty: the general form)
};
std::string jit_adjoint_source(const TileProgram& tp, int precision, AdjointCoefs& out);
// the coefficients, derivative factors and forms alone (what jit_adjoint_source also returns)
void adjoint_coefs(const TileProgram& tp, int precision, AdjointCoefs& out, std::vector<size_t>* offs = nullptr);
// everything of an adjoint reverse pass its generated source depends on besides the coefficients
std::vector<uint64_t> adjoint_structure_key(const TileProgram& tp, int precision, const AdjointCoefs& c);
// partials: [slot][blocks * waves] doubles, as the interpreted adjoint pass writes them
void launch_jit_adjoint_pass(kq_state* s, hipFunction_t fn, const TileProgram& tp, void* psi, void* lam,
                             const double* d_coefs, double* d_partials);

// ---- batched energy + adjoint gradient (kq_batch.cpp) ------------------------------------------------
// precision: the reverse sweep's (and lambda's); forward_precision: the forward sweep's and the
// energy's (complex128 forward with a complex64 reverse sweep gives FP64-exact energies)
kq_batch* batch_create(int precision, int forward_precision, hipStream_t stream);
void batch_destroy(kq_batch* b);
void batch_add(kq_batch* b, uint32_t n_qubits, const kq_op* ops, size_t n_ops, size_t n_params, size_t n_strings,
               const uint32_t* off, const uint32_t* q, const uint8_t* k, const uint8_t* phases, const double* coeffs,
               size_t* index);
void batch_set_parameters(kq_batch* b, size_t index, const double* values);
void batch_set_all_parameters(kq_batch* b, const double* values);
void batch_run(kq_batch* b, bool use_graph, double* energies, double* grads, double* drift);
void batch_info(const kq_batch* b, size_t* n_instances, size_t* n_params, size_t* device_bytes, size_t* n_launches);
void batch_set_drift_recompute(kq_batch* b, double drift_tol);
void batch_plan_info(const kq_batch* b, size_t* fwd_passes, size_t* rev_passes, double* algorithmic_bytes);
void batch_recomputed(const kq_batch* b, size_t* last_run, size_t* total);

}  // namespace kq
