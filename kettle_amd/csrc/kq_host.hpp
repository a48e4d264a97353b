// Host-side helpers shared by the C-ABI translation units of libkq (kq_api.cpp, kq_run.cpp,
// kq_measure.cpp, kq_pauli.cpp, kq_adjoint.cpp, kq_debug.cpp).
#pragma once

#include <cstdint>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "kq_internal.hpp"

namespace kq
{

// Runs the body of a C entry point: maps the library's exception types to kq.h status codes and
// keeps the message for kq_last_error().
// the status an injected fault makes this call return (0: none); see kq_debug_set_faults
int take_injected_fault();

template <typename F>
int guarded(F&& f)
{
    if (const int injected = take_injected_fault(); injected != 0) {
        set_error("injected fault");
        return injected;
    }
    try {
        f();
        return KQ_OK;
    }
    catch (const InvalidArg& e) {
        set_error(e.what());
        return KQ_ERR_INVALID;
    }
    catch (const HipError& e) {
        set_error(e.what());
        return KQ_ERR_HIP;
    }
    catch (const AllocError& e) {
        set_error(e.what());
        return KQ_ERR_ALLOC;
    }
    catch (const std::out_of_range& e) {
        set_error(e.what());
        return KQ_ERR_OUT_OF_RANGE;
    }
    catch (const std::exception& e) {
        set_error(e.what());
        return KQ_ERR_RUNTIME;
    }
    catch (...) {
        set_error("unknown exception");
        return KQ_ERR_RUNTIME;
    }
}

// ---- library and states (kq_api.cpp) ----------------------------------------------------------------
void ensure_init();                 // selects the library's device (initialises it once)
// a usable state; a pending basis state (kq_state::lazy_basis) is written first
void check_state(const kq_state* s);
// a usable state, a pending basis state left pending (callers that overwrite the state or build it)
void check_state_lazy(const kq_state* s);
void* device_alloc(size_t bytes);   // hipMalloc or AllocError
void* grow(void*& buf, size_t& have, size_t need);   // device buffer of at least `need` bytes
// pending_zero: |0...0> is left pending (kq_state::lazy_basis) instead of written now; callers that
// write the buffer directly create states without it
kq_state* create_state(uint32_t n, int precision, bool dm, bool pending_zero = false);

// ---- gate runs (kq_run.cpp) ---------------------------------------------------------------------------
// device records of a straight-line list through the fused passes (qubit moves where allowed)
void run_records(kq_state* s, const std::vector<GateRec>& recs, bool allow_relabel = true);
void run_ops(kq_state* s, const kq_op* ops, size_t n);
// puts qubit q back on index bit q (after gate passes that moved qubits)
void normalise_layout(kq_state* s);
void normalise_layout(const kq_state* s);
// writes a pending basis state (kq_state::lazy_basis) into the state's buffer
void materialise_basis(kq_state* s);
// SWAP triples dropped as relabels: sigma receives the qubit permutation they leave
std::vector<GateRec> exchanges_as_relabels(const std::vector<GateRec>& recs, uint32_t nvec, std::vector<uint8_t>& sigma);
PermuteTables permute_tables(const std::vector<uint8_t>& bit_of);
// The fused passes of a straight-line statevector gate list from qubit q on bit q, as a gate run plans
// them (plan tuner, qubit moves when enabled); bit_of receives the placement the passes leave.
std::vector<TilePass> plan_statevector_passes(kq_state* s, const std::vector<GateRec>& recs, bool allow_sweep,
                                              std::vector<uint8_t>& bit_of);
// planned passes in order on the state's stream, with after[j].second enqueued right after pass
// after[j].first (after sorted by pass); with sync, returns once the stream has drained, else once
// everything is enqueued (the pinned staging of the pass coefficients must not be reused before the
// stream drains: the caller synchronises before the state runs another gate list)
void execute_passes_with(kq_state* s, const std::vector<TilePass>& passes,
                         const std::vector<std::pair<size_t, std::function<void()>>>& after, bool sync = true);

// ---- measurement, interpreter, sampling (kq_measure.cpp) ---------------------------------------------
std::mt19937 get_prng(bool has_seed, int32_t seed);
int measure(kq_state* s, uint32_t q, bool has_seed, int32_t seed);
std::vector<double> download_host(const kq_state* s);
void simulate(kq_state* s, uint32_t circuit_n, const kq_op* ops, size_t n_ops, const int32_t* preds, size_t n_preds,
              bool has_seed, int32_t seed, int8_t* creg, size_t n_bits, kq_logger_fn logger, void* logger_user);
void sample_from_device_probs(hipStream_t st, double* d_p, uint64_t n, uint64_t shots, bool has_seed, int32_t seed,
                              uint64_t* out);
void check_noise(const double* noise, uint32_t n);
void sample_circuit_shots(const kq_state* init, uint32_t circuit_n, const kq_op* ops, size_t n_ops, const int32_t* preds,
                          size_t n_preds, size_t n_bits, const double* noise, uint64_t n_shots, bool has_seed, int32_t seed,
                          uint64_t max_cached_bytes, uint64_t* out, uint64_t* n_branches);

// ---- adjoint gradient (kq_adjoint.cpp) -----------------------------------------------------------------
void energy_gradient(kq_state* s, const kq_op* ops, size_t n_ops, size_t n_strings, const uint32_t* off,
                     const uint32_t* q, const uint8_t* k, const uint8_t* phases, const double* coeffs, size_t n_params,
                     double* energy, double* grad, uint32_t checkpoint_every, double drift_tol, size_t* n_restores);

}  // namespace kq
