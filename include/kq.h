/*
 * kq.h -- C-ABI of libkq, the MI355X (gfx950) engine under kettle's simulate() hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No torch or C++ types
 * cross this boundary.  Every entry point returns KQ_OK (0) on success; on failure it returns a
 * non-zero status and kq_last_error() gives a thread-local message.  Host-side validation
 * failures that kettle reports with std::runtime_error return KQ_ERR_RUNTIME with kettle's
 * message, so a C++ caller can rethrow the same exception type and text.
 *
 * Indexing is kettle's: little endian, qubit q <-> bit q of the amplitude index
 * (reference: source/kettle_internal/state/qubit_state_conversion.cpp:37-83).
 * Amplitudes cross the boundary as interleaved (re, im) doubles, which is the memory layout of
 * kettle's std::vector<std::complex<double>> (include/kettle/state/statevector.hpp:94-96).
 * Density matrices cross it column-major, (row r, col c) at r + c*N, which is the layout of the
 * Eigen::MatrixXcd kettle stores (include/kettle/state/density_matrix.hpp).
 */
#ifndef KQ_ABI_H_INCLUDED
#define KQ_ABI_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define KQ_OK            0
#define KQ_ERR_INVALID   1   /* bad argument (null pointer, out-of-range qubit, bad size)      */
#define KQ_ERR_HIP       2   /* a HIP runtime call failed                                      */
#define KQ_ERR_ALLOC     3   /* device allocation failed or exceeds the configured cap         */
#define KQ_ERR_RUNTIME   4   /* kettle-level std::runtime_error (message is kettle's text)     */
#define KQ_ERR_OUT_OF_RANGE 5 /* kettle-level std::out_of_range                                */

/* ---- precisions -------------------------------------------------------------------------- */
#define KQ_C128 0            /* complex<double> amplitudes (kettle's only precision)           */
#define KQ_C64  1            /* complex<float> amplitudes; norms, probabilities, CDFs,
                                expectations and gradients are still accumulated in FP64     */

/* ---- gate codes: the values of kettle's `enum class Gate : std::uint8_t`
 *      (include/kettle/gates/primitive_gate.hpp:19-52), in the same order ------------------ */
enum kq_gate {
    KQ_H = 0, KQ_X, KQ_Y, KQ_Z, KQ_S, KQ_SDAG, KQ_T, KQ_TDAG, KQ_SX, KQ_SXDAG,
    KQ_RX = 10, KQ_RY, KQ_RZ, KQ_P,
    KQ_CH = 14, KQ_CX, KQ_CY, KQ_CZ, KQ_CS, KQ_CSDAG, KQ_CT, KQ_CTDAG, KQ_CSX, KQ_CSXDAG,
    KQ_CRX = 24, KQ_CRY, KQ_CRZ, KQ_CP,
    KQ_U = 28, KQ_CU = 29,
    KQ_M = 30,
    /* program-structure ops (kettle's non-gate CircuitElements, circuit_element.hpp:23-94) */
    KQ_OP_IF = 64,           /* q0 = predicate offset, q1 = length of the body that follows    */
    KQ_OP_IF_ELSE = 65,      /* q0 = predicate offset, q1 = if-body length, aux = else length  */
    KQ_OP_LOG_STATE = 66,    /* StatevectorCircuitLogger: snapshot the state (callback)       */
    KQ_OP_LOG_CREG = 67      /* ClassicalRegisterCircuitLogger: snapshot the register         */
};

/*
 * One element of a flattened circuit.  Fields follow kettle's GateInfo
 * (primitive_gate.hpp:71-79; packing in gates/primitive_gate/gate_create.cpp:34-244):
 *   1-target gates (H..P, U):  q0 = target
 *   1-control-1-target gates:  q0 = control, q1 = target
 *   M:                         q0 = qubit,   q1 = classical bit
 *   angle gates:               angle = the evaluated angle (literal or parameter value)
 *   U / CU:                    u = {re00, im00, re01, im01, re10, im10, re11, im11}
 *   aux:                       parameterised gates: index of the parameter (-1 = literal);
 *                              KQ_OP_IF_ELSE: length of the else body
 * The body of an IF / IF_ELSE is the next q1 (and then aux) ops of the same array, so nested
 * sub-circuits flatten recursively.
 */
typedef struct kq_op {
    int32_t gate;
    int32_t q0;
    int32_t q1;
    int32_t aux;
    double  angle;
    double  u[8];
} kq_op;

/*
 * Control-flow predicates (kettle's ControlFlowPredicate, control_flow_predicate.cpp:12-69)
 * are packed in one int32 array.  A predicate at offset p is
 *   preds[p] = kind (0 = IF, 1 = IF_NOT), preds[p+1] = n, then n pairs (bit_index, expected).
 */

/* classical register cells: -1 = not measured, else 0 / 1 (kettle's optional<int>) */

/* Logger callback: kind is KQ_OP_LOG_STATE or KQ_OP_LOG_CREG.  For LOG_STATE, `amps` is a host
 * copy of the state (interleaved complex doubles; 2*N doubles for a statevector, 2*N*N for a
 * density matrix).  For LOG_CREG, `creg` is the register.  Both pointers are valid only for the
 * duration of the call. */
typedef void (*kq_logger_fn)(void* user, int32_t kind, const double* amps, const int8_t* creg,
                             size_t n_bits);

/* ---- library ------------------------------------------------------------------------------ */
const char* kq_last_error(void);
int kq_init(int device);                 /* selects the HIP device; idempotent                */
int kq_device_info(char* name, size_t name_len, int* n_cu, size_t* total_mem);
int kq_synchronize(void);

/* ---- states (statevector or density matrix) ----------------------------------------------- */
typedef struct kq_state kq_state;

/* Statevector of n_qubits in |0...0> (kettle: Statevector(n), state/state.cpp:21-28). */
int kq_sv_create(kq_state** out, uint32_t n_qubits, int precision);
/* Density matrix |0..0><0..0| of n_qubits, stored as a 2n-qubit vector vec(rho). */
int kq_dm_create(kq_state** out, uint32_t n_qubits, int precision);
int kq_state_destroy(kq_state* s);
int kq_state_info(const kq_state* s, uint32_t* n_qubits, int* precision, int* is_density);
/* Raw device pointer and byte size of the amplitude buffer (for zero-copy interop), qubit q on index
 * bit q.  Valid until the next call that applies gates to the state: a gate run may leave qubits on
 * other index bits, and putting them back may move the amplitudes to a new buffer. */
int kq_state_device_ptr(kq_state* s, void** ptr, size_t* bytes);
/* Host <-> device copies of the whole state, interleaved complex doubles (2*N or 2*N*N). */
int kq_state_upload(kq_state* s, const double* host);
int kq_state_download(const kq_state* s, double* host);
/* Computational basis state |index> (or |index><index| for a density matrix).  Nothing is written
 * yet (nor by kq_sv_create / kq_dm_create for their |0...0>): the next gate run builds the state in its
 * first pass, and any other call on the state writes it first.  A device pointer fetched earlier does
 * not see it until it is written. */
int kq_state_set_basis(kq_state* s, uint64_t index);
int kq_state_copy(kq_state* dst, const kq_state* src);

/* ---- simulate() ---------------------------------------------------------------------------
 * ket::simulate(circuit, state, seed) for a flattened circuit (simulate.cpp:417-597 for a
 * statevector, simulate_density_matrix.cpp:354-548 for a density matrix).  `creg` must hold
 * n_bits cells; they are reset to -1 first.  has_seed = 0 draws mid-circuit measurements from
 * std::random_device like kettle's get_prng_(nullopt) (common/prng.cpp:9-19).
 * A density matrix rejects logger ops with kettle's error, as simulate_density_matrix.cpp:391. */
int kq_simulate(kq_state* s, uint32_t circuit_n_qubits, const kq_op* ops, size_t n_ops,
                const int32_t* preds, size_t n_preds, int has_seed, int32_t seed,
                int8_t* creg, size_t n_bits, kq_logger_fn logger, void* logger_user);

/* Straight-line gate list (no M, no control flow): the fused device path alone. */
int kq_apply_gates(kq_state* s, const kq_op* ops, size_t n_ops);

/* Gate fusion control: max_tile_qubits = 0 disables fusion (one sweep per gate), -1 picks the
 * largest tile that fits LDS; max_gates_per_pass and max_stages_per_pass (register stages, each
 * costing one LDS remap of the tile) bound a pass; values <= 0 leave a setting unchanged. */
int kq_set_fusion(int max_tile_qubits, int max_gates_per_pass, int max_stages_per_pass,
                  double max_cost_per_pass);
int kq_get_fusion(int* max_tile_qubits, int* max_gates_per_pass, int* max_stages_per_pass,
                  double* max_cost_per_pass);
/* max_cost_per_pass bounds a pass's arithmetic in units of one general complex 2x2 gate (real or
 * axis-rotation matrices count 1/2, one-sided diagonals 1/4, swaps 1/10): above roughly the HBM
 * time of one sweep a pass stops being memory-bound. */
/* Qubits 0..low_qubits-1 belong to every tile pass (default 5), so that each wavefront access
 * covers 2^low_qubits consecutive amplitudes; fewer leave more tile slots to other qubits. */
int kq_set_fusion_low_qubits(int low_qubits);
/* on != 0 (default): a statevector's gate passes may store qubits on other index bits of their tile,
 * chosen so that the next pass takes the most gates; the state records where each qubit is, and
 * every operation other than gates first stores qubit q back on bit q (gate-free tile passes). */
int kq_set_fusion_relabel(int on);
/* on != 0 (default): for statevector gate runs through the run-time compiled passes, the planner
 * tries tiles with 4-7 low qubits and 12-28 register stages per pass on the host, keeps the plan a
 * cost model prices lowest (passes, plus LDS remaps above six per pass) and caches that choice per
 * gate structure (not the angles).  kq_set_fusion with max_stages_per_pass > 0 and
 * kq_set_fusion_low_qubits switch it off (explicit settings win); this call switches it back on.
 * No counterpart in the reference (kettle has no fusion). */
int kq_set_fusion_autotune(int on);
/* The tuner's state and its last choice; plans_tried counts candidate plans built so far. */
int kq_fusion_autotune_info(int* on, int* low_qubits, int* max_stages, size_t* plans_tried, size_t* cached);
/* Tile passes compiled at run time for their exact gate structure (straight-line code, no per-gate
 * dispatch; coefficients stay data, so repeated structures with other angles reuse the kernel):
 * mode 0 = off (every pass interpreted), 1 = states of >= 18 qubits (default), 2 = every tile pass.
 * A pass whose kernel fails to compile runs interpreted; kq_jit_info counts it. */
int kq_set_jit(int mode);
/* Code shape of the run-time pass kernels: min_waves_per_simd > 0 caps their registers so that at
 * least that many waves fit per SIMD (default 3; 0: the compiler's choice; a kernel that would spill
 * under the cap is compiled again without it); blocks_per_cu > 0 launches that many blocks per CU,
 * each walking several tiles and loading the next tile while the current one computes (0: one block
 * per tile);
 * normalise != 0 (default) divides each uncontrolled one-qubit gate by its larger first-row entry
 * and multiplies the product of those pivots back in once per pass. */
int kq_set_jit_tuning(int min_waves_per_simd, int blocks_per_cu, int normalise);
/* Wave priority in the run-time pass kernels: 0 leaves it alone, 1 (default) raises it while a wave
 * issues its tile's HBM loads, 2 also while it issues its stores, so that waves starting or finishing
 * a tile go ahead of the waves computing one.  No counterpart in the reference. */
int kq_set_jit_priority(int mode);
/* Nontemporal hint on the run-time pass kernels' HBM traffic: 0 off, 1 on the tile loads, 2 on the
 * loads and the stores, 3 (default) as 2 for vectors larger than 256 MiB (the Infinity Cache) and as
 * 0 for smaller ones.  No counterpart in the reference. */
int kq_set_jit_nontemporal(int mode);
/* Counters of the run-time compiled kernels; uncapped counts kernels compiled again without the
 * register cap because they spilled under it. */
int kq_jit_info(int* mode, size_t* compiled, size_t* cache_hits, size_t* failures, double* compile_seconds,
                size_t* uncapped);
/* Fusion plan statistics of the last kq_apply_gates / kq_simulate call. */
int kq_last_plan_stats(size_t* n_gates, size_t* n_passes, size_t* n_tile_passes);

/* ---- mid-circuit measurement (measure.hpp:33-57, measure.cpp:22-72;
 *      density matrix: measure_density_matrix.cpp:21-95) ---------------------------------- */
int kq_qubit_weights(kq_state* s, uint32_t qubit, double out[2]);
int kq_collapse(kq_state* s, uint32_t qubit, int kept_bit, double scale);
int kq_measure(kq_state* s, uint32_t qubit, int has_seed, int32_t seed, int* outcome);

/* ---- probabilities and sampling (calculations/probabilities.cpp:49-124,
 *      calculations/measurements.cpp:29-312) --------------------------------------------- */
/* |amp_i|^2 (diag(rho) for a density matrix) with optional per-qubit bit-flip noise
 * (noise = NULL or n_qubits values in [0,1]).  out: N doubles on the host. */
int kq_probabilities(kq_state* s, const double* noise, double* out);
/* perform_measurements_as_memory(state, shots, noise, seed): state indices per shot. */
int kq_sample(kq_state* s, const double* noise, uint64_t n_shots, int has_seed, int32_t seed,
              uint64_t* out_indices);
/* perform_measurements_as_counts_marginal(circuit, state, shots, ...) (measurements.cpp:178-208):
 * every shot runs the circuit on a copy of `initial` (mid-circuit outcomes drawn afresh per shot,
 * as simulate() without a seed does) and draws one index from its final probabilities (with noise)
 * with a sampler seeded by `seed` (has_seed; then every shot of one outcome history gets the same
 * index, as in the reference) or from std::random_device.  Shots with the same outcome history
 * share one simulated branch, so the device work grows with the number of distinct histories, not
 * with n_shots; branch states are cached up to max_cached_bytes (a branch that does not fit is
 * rebuilt from its nearest cached ancestor when needed).  Loggers are ignored.  out_indices gets
 * n_shots indices; *n_branches (optional) the number of branch nodes simulated. */
int kq_sample_circuit_shots(const kq_state* initial, uint32_t circuit_n_qubits, const kq_op* ops, size_t n_ops,
                            const int32_t* preds, size_t n_preds, size_t n_bits, const double* noise, uint64_t n_shots,
                            int has_seed, int32_t seed, uint64_t max_cached_bytes, uint64_t* out_indices,
                            uint64_t* n_branches);
/* The same sampler over caller-provided raw probabilities (perform_measurements_as_memory(
 * probabilities_raw, ...)); probabilities are uploaded and scanned on the device. */
int kq_sample_probabilities(const double* probs, uint64_t n_probs, uint64_t n_shots,
                            int has_seed, int32_t seed, uint64_t* out_indices);

/* kq_sample_probabilities without any device allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The probabilities, their CDF, the draws and the indices live in a
 * fixed workspace of the library, 24.2 MiB per device, held from the first use on a device until the
 * process ends; the first call on a device also creates the stream these calls run on.  Limits:
 * n_probs <= 2^20 and n_shots <= 2^20, else KQ_ERR_INVALID.  Every call on one device shares that
 * device's workspace: calls that overlap on one GPU (from several host threads) race. */
int kq_sample_probabilities_noalloc(const double* probs, uint64_t n_probs, uint64_t n_shots,
                                    int has_seed, int32_t seed, uint64_t* out_indices);

/* ---- Pauli strings (operator/pauli/pauli_operator.cpp:76-105, simulation/simulate_pauli.cpp)
 * A Pauli operator of S strings is given as:
 *   term_offsets[S+1]: string s owns terms [term_offsets[s], term_offsets[s+1])
 *   term_qubits[], term_kinds[] (1 = X, 2 = Y, 3 = Z; 0 = I is ignored), in kettle's terms() order
 *   phases[S] (PauliPhase: 0 = +1, 1 = +i, 2 = -1, 3 = -i), coeffs[2*S] (re, im)
 * Expectation = sum_s coeff_s * phase_s * <psi| P_s |psi>, accumulated in FP64. */
int kq_pauli_expectation(kq_state* s, size_t n_strings, const uint32_t* term_offsets,
                         const uint32_t* term_qubits, const uint8_t* term_kinds,
                         const uint8_t* phases, const double* coeffs, double out[2]);
/* ket::simulate(SparsePauliString, state): applies one string's terms in place. */
int kq_apply_pauli_string(kq_state* s, size_t n_terms, const uint32_t* term_qubits,
                          const uint8_t* term_kinds);

/* ---- density-matrix read-out (state/density_matrix.cpp:150-278) ------------------------- */
int kq_dm_trace(kq_state* s, double out[2]);
/* DensityMatrix::trace_of_square (density_matrix.hpp:100-104, density_matrix.cpp:164-168): the trace
 * of rho * rho, sum_{r,c} rho(r,c) rho(c,r), accumulated in FP64; kettle's is_pure(tol)
 * (density_matrix.cpp:151-155) is std::norm(trace_of_square - 1) <= tol on top of it. */
int kq_dm_trace_of_square(kq_state* s, double out[2]);
/* partial_trace(rho, qubits): traces out `qubits`; out: (2^(n-k))^2 complex, column-major. */
int kq_dm_partial_trace(kq_state* s, const uint32_t* qubits, size_t n, double* out);
/* |psi><psi| of a statevector (statevector_to_density_matrix, density_matrix.cpp:170-185). */
int kq_dm_from_statevector(kq_state* dm, const kq_state* sv);

/* ---- adjoint energy gradient (no kettle counterpart: kettle's VQE is derivative-free) ------
 * For a straight-line parameterised circuit (ops with aux >= 0 name parameter slots) and a
 * Pauli operator H (same packing as kq_pauli_expectation, real weights), computes
 * E = <psi|H|psi> and dE/dtheta_p for p < n_params by one forward and one reverse sweep.
 * `state` is the work state (its precision is the compute precision); it ends holding psi(theta).
 * The forward sweep is one plan of fused passes; checkpoints of psi sit at pass boundaries, every
 * ceil(sqrt(P)) of its P passes (checkpoint_every = 0) or at the first boundary past every
 * checkpoint_every gates, at most 32 GiB of them.  They bound inverse-gate drift: at each checkpoint
 * the uncomputed state is compared with the stored one on the device and replaced by it when the
 * max-abs deviation exceeds drift_tol.  n_restores (nullable) reports how often that happened.
 * Buffers (lambda, checkpoints, partials) stay in the state's workspace between calls. */
int kq_energy_gradient(kq_state* state, const kq_op* ops, size_t n_ops,
                       size_t n_strings, const uint32_t* term_offsets, const uint32_t* term_qubits,
                       const uint8_t* term_kinds, const uint8_t* phases, const double* coeffs,
                       size_t n_params, double* energy, double* grad,
                       uint32_t checkpoint_every, double drift_tol, size_t* n_restores);

/* ---- batched energy + adjoint gradient (no kettle counterpart) --------------------------------
 * Many independent straight-line parameterised circuits, each with its own qubit count, Pauli
 * operator (same packing as kq_pauli_expectation, real weights) and parameters, evaluated together:
 * kettle's VQE cost function (vqe_basic.cpp:61-82) for thousands of instances per call.  Instances
 * are planned once (at the first run after the last kq_batch_add); kq_batch_set_parameters only
 * changes angles.  kq_batch_run computes, for every instance i in add order, energies[i] and its
 * gradient at grads + (sum of n_params of the instances before i); drift[i] (nullable) is the
 * max-abs distance of the uncomputed state from |0...0> after the reverse sweep, a measure of
 * inverse-gate drift.  use_graph != 0 replays the launch sequence as a captured hipGraph from the
 * second run on.  A batch owns its stream and device memory; calls on one batch are serialised.
 * precision is that of the reverse sweep and of lambda; forward_precision that of the forward
 * sweep and the energies: equal to precision, or KQ_C128 with KQ_C64 (FP64-exact energies, the
 * gradient's reverse sweep at complex64 traffic). */
typedef struct kq_batch kq_batch;
int kq_batch_create(kq_batch** out, int precision, int forward_precision);
int kq_batch_add(kq_batch* b, uint32_t n_qubits, const kq_op* ops, size_t n_ops, size_t n_params,
                 size_t n_strings, const uint32_t* term_offsets, const uint32_t* term_qubits,
                 const uint8_t* term_kinds, const uint8_t* phases, const double* coeffs, size_t* index);
int kq_batch_set_parameters(kq_batch* b, size_t index, const double* values);
/* Every instance's parameters at once (an optimiser step): values holds them in add order, n_params
 * per instance back to back.  The next run lowers the changed instances again on host threads. */
int kq_batch_set_all_parameters(kq_batch* b, const double* values);
int kq_batch_run(kq_batch* b, int use_graph, double* energies, double* grads, double* drift);
int kq_batch_info(const kq_batch* b, size_t* n_instances, size_t* n_params_total, size_t* device_bytes,
                  size_t* n_launches);
/* Plan totals over all instances (after the first run): forward and reverse tile passes, and the
 * algorithmic bytes of one run (one read and/or write of every amplitude of every vector that each
 * launch updates: reset, forward passes, lambda = H psi, reverse passes on psi and lambda, drift). */
int kq_batch_plan_info(const kq_batch* b, size_t* fwd_passes, size_t* rev_passes, double* algorithmic_bytes);
/* Drift-triggered recompute: after every later kq_batch_run, each instance whose drift exceeds
 * drift_tol (or is NaN) is evaluated again by kq_energy_gradient on a complex128 work state of its
 * qubit count (held by the batch, with a restore at every checkpoint), and that energy and
 * gradient replace the batched ones in the outputs; drift[] still reports the batched sweep's.
 * drift_tol < 0 (the default) turns it off; 0 recomputes every instance with any drift.
 * kq_batch_recomputed reports how many instances the last run recomputed, and in total. */
int kq_batch_set_drift_recompute(kq_batch* b, double drift_tol);
int kq_batch_recomputed(const kq_batch* b, size_t* last_run, size_t* total);
int kq_batch_destroy(kq_batch* b);

/* ---- profiling ------------------------------------------------------------------------------
 * While enabled, every gate pass launched by kq_simulate / kq_apply_gates / kq_energy_gradient
 * and every Pauli-operator launch of kq_energy_gradient is bracketed by HIP events on the state's
 * own stream; durations and algorithmic bytes accumulate per class:
 *   0 = tile pass whose tile is the low contiguous qubits (LDS-fused low-index gates)
 *   1 = tile pass that includes higher qubits (strided high-index gates)
 *   2 = single-gate sweep (lone diagonal gates)
 *   3 = adjoint reverse tile pass (two vectors and derivative partials)
 *   4 = lambda = H psi (read psi + write lambda counted; the gathers per X-mask group re-read psi)
 * Bytes: one read + one write of every amplitude of every vector a launch updates. */
#define KQ_PASS_CLASSES 5
typedef struct kq_pass_stats {
    double   ms;          /* summed kernel time */
    uint64_t launches;
    double   bytes;       /* summed algorithmic bytes */
    uint64_t gates;       /* gates applied by these launches */
    double   min_ms;      /* shortest and longest launch */
    double   max_ms;
} kq_pass_stats;
int kq_profile_begin(void);
/* Device bandwidth reference: GB/s of plain streams of 16-byte elements over two 2^n_qubits
 * buffers (16 <= n_qubits <= 32, 2 x 16 * 2^n bytes of device memory for the call): read only,
 * write only, read-modify-write in place, copy, each the best of six launch shapes.
 * kq_stream_probe_shapes gives every shape: gbps[4 * shape + mode], shapes 0 / 1 = a persistent
 * grid of 2 / 8 blocks of 256 threads per CU, 2 = one block per 4096-element chunk (the tile
 * passes' shape), 3 = shape 2 with nontemporal loads and stores, 4 / 5 = shapes 2 / 3 with the
 * wave priority raised while the loads issue (as the tile passes do).  No counterpart in the
 * reference. */
int kq_stream_probe(uint32_t n_qubits, double gbps[4]);
int kq_stream_probe_shapes(uint32_t n_qubits, double gbps[24]);
int kq_profile_end(kq_pass_stats out[KQ_PASS_CLASSES]);

/* ---- debug / parity probes ------------------------------------------------------------------
 * Device-computed pair maps in kettle's generator order (gate_pair_generator.hpp:24-160):
 * flat pair index k -> (s0, s1).  For the 2-qubit map with four_way != 0, writes 4 indices per
 * k in next_unset_and_set() order (c0t0, c0t1, c1t0, c1t1). */
int kq_debug_pair_map_1q(uint32_t n_qubits, uint32_t target, uint64_t* s0, uint64_t* s1);
int kq_debug_pair_map_2q(uint32_t n_qubits, uint32_t control, uint32_t target, int four_way,
                         uint64_t* out);

/* Host-only view of the fusion plan of a straight-line gate list (no device work, no GPU
 * needed).  Each op compiles to one device gate record (two for a density matrix: U on the row
 * bits, then conj(U) on the column bits, records 2i and 2i+1).  For the j-th record applied,
 * order[j] is its record index and pass_of[j] the pass that applies it; for each pass,
 * pass_kind is its profiling class (0, 1, 2 as above) and pass_cost its arithmetic cost in the
 * units of kq_set_fusion's max_cost_per_pass.  order and pass_of hold one entry per record,
 * pass_kind and pass_cost at least as many; lds_bytes is the per-workgroup LDS the tile is
 * fitted to (0 = the value of the initialised device, 64 KiB before kq_init). */
/* Host-only check of the planner's qubit moves: applies a straight-line gate list (<= 24 qubits) to
 * `amps` (interleaved complex doubles, in place) pass by pass as planned with moves, each gate on the
 * index bits holding its qubits at that point, moves the amplitudes as each pass's store would, then
 * applies the restore passes; the result must equal the plain simulation. */
int kq_debug_check_relabel(uint32_t n_qubits, const kq_op* ops, size_t n_ops, double* amps, size_t* n_passes,
                           size_t* n_restore_passes);
/* Host-only: generate and compile (without loading) the run-time pass kernels of a straight-line
 * gate list; fails with the compiler log if any does not compile.  No GPU needed. */
int kq_debug_jit_compile(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                         size_t* n_kernels, size_t* failures);
/* Host-only: pass pass_index of a straight-line gate list as a statevector runs it (qubit moves, the
 * last pass's return of every qubit to its own bit and any restore passes included).  info[0] = 1 for
 * a sweep of one gate (coefs: its 8 matrix values, info[3] = target bit, info[4] = control bit + 1 or
 * 0), else 0 for a tile pass with its generated kernel source, or 2 for one that stores out of place
 * (its kernel takes a fourth argument, the destination buffer, which then holds the state);
 * info[1] = workgroups, info[2] = threads per workgroup, coefs: its coefficient buffer. */
int kq_debug_pass(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes, size_t pass_index,
                  char* src, size_t src_cap, size_t* src_len, double* coefs, size_t coef_cap, size_t* n_coefs,
                  uint64_t info[5], size_t* n_passes);
/* Host-only: as kq_debug_pass, for a run that starts from the basis state |basis> not yet written
 * (kq_state_set_basis): info[5] = the pass kernel's input (0: it loads the state; 1: it builds its
 * tiles from the basis index, its fourth argument after the destination buffer of an out-of-place
 * pass, and reads nothing; 2: it skips the tiles whose outside index bits under info[6] differ from
 * info[7], taking those two as its last arguments, and stores zeros there when it stores out of
 * place), info[6] the known-bit mask, info[7] the known values (the basis index for input 1). */
int kq_debug_pass_from_basis(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                             size_t pass_index, uint64_t basis, char* src, size_t src_cap, size_t* src_len, double* coefs,
                             size_t coef_cap, size_t* n_coefs, uint64_t info[8], size_t* n_passes);
/* Host-only: generate and compile (without loading) the run-time adjoint reverse-pass kernels of a
 * straight-line parameterised gate list (ops with aux >= 0 carry derivatives), the whole list as one
 * reverse segment; fails with the compiler log if any does not compile.  No GPU needed. */
int kq_debug_adjoint_jit_compile(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                                 size_t* n_kernels, size_t* failures);
/* Host-only: the generated source of tile pass pass_index (NUL-terminated, truncated to cap - 1);
 * *len is the full length and *n_kernels the number of tile passes. */
int kq_debug_jit_source(uint32_t n_qubits, int precision, const kq_op* ops, size_t n_ops, size_t lds_bytes,
                        size_t pass_index, char* buf, size_t cap, size_t* len, size_t* n_kernels);
int kq_debug_plan(uint32_t n_qubits, int precision, int is_density, const kq_op* ops, size_t n_ops,
                  size_t lds_bytes, uint32_t* order, uint32_t* pass_of, int32_t* pass_kind,
                  double* pass_cost, size_t* n_passes);
/* A/B probe of the vector ALUs against the matrix cores: psi[r, :] <- U psi[r, :] for every row r of
 * 16 amplitudes (U on qubits 0..3, 16 x 16 complex, row-major, interleaved FP64; rounded to FP32 for a
 * complex64 state), by one launch of the VALU form (use_mfma = 0) or of the MFMA form (16x16x4
 * matrix instructions); *ms (nullable) is that launch's duration from HIP events.  Statevectors of
 * at least 8 qubits.  No counterpart in the reference (kettle has no multi-qubit dense gate). */
int kq_debug_dense_block(kq_state* state, const double* u, int use_mfma, double* ms);
/* Lane layout of the 16x16x4 matrix instruction (FP64 for KQ_C128, FP32 for KQ_C64): one wave runs
 * D = A x B once with lane l holding operands a[l], b[l]; d[4 l + e] is element e of lane l's D. */
int kq_debug_mfma16_layout(int precision, const double* a, const double* b, double* d);
/* Fault injection for tests of the callers' error paths (no counterpart in the reference): any single
 * device allocation of more than alloc_cap_bytes fails with KQ_ERR_ALLOC (0: no cap), and, when
 * fail_status is non-zero, the entry-point call after the next fail_after_calls ones returns
 * fail_status without doing anything, with the message "injected fault" (once).  Process-wide; this
 * call and kq_last_error are never failed.  kettle's C++ layer turns the statuses into its exception
 * types like any other failure (std::out_of_range for KQ_ERR_OUT_OF_RANGE, else std::runtime_error). */
int kq_debug_set_faults(size_t alloc_cap_bytes, int fail_status, int fail_after_calls);

#ifdef __cplusplus
}
#endif

#endif /* KQ_ABI_H_INCLUDED */
