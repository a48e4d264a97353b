// The C-ABI entry points of libkq (include/kq.h) for the library and the state handles: device
// initialisation, state creation and copies, profiling switches and the batched-gradient handles.
// The gate runs, measurement and sampling, Pauli operators, the adjoint gradient and the debug entry
// points live in kq_run.cpp, kq_measure.cpp, kq_pauli.cpp, kq_adjoint.cpp and kq_debug.cpp.
#include <atomic>
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

thread_local std::string g_last_error;
std::mutex g_init_mutex;
bool g_initialised = false;
int g_device = 0;

}  // namespace

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

void check_state_lazy(const kq_state* s)
{
    if (s == nullptr || s->amps == nullptr) {
        throw InvalidArg {"null or destroyed state handle"};
    }
}

void check_state(const kq_state* s)
{
    check_state_lazy(s);
    auto* w = const_cast<kq_state*>(s);
    materialise_basis(w);
    w->zero_mask = 0;   // the call may change the state: no known zeros are kept
}

namespace
{
// kq_debug_set_faults: a cap on single device allocations, and a status the call after the next
// g_fail_after ones returns
std::atomic<size_t> g_alloc_cap {0};
std::atomic<int> g_fail_status {0};
std::atomic<int> g_fail_after {0};
}  // namespace

int take_injected_fault()
{
    if (g_fail_status.load(std::memory_order_relaxed) == 0) {
        return 0;
    }
    if (g_fail_after.fetch_sub(1) > 0) {
        return 0;
    }
    return g_fail_status.exchange(0);
}

void* device_alloc(size_t bytes)
{
    const size_t cap = g_alloc_cap.load(std::memory_order_relaxed);
    if (cap != 0 && bytes > cap) {
        throw AllocError {"device allocation of " + std::to_string(bytes) + " bytes exceeds the cap of " + std::to_string(cap)
                          + " bytes"};
    }
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

kq_state* create_state(uint32_t n, int precision, bool dm, bool pending_zero)
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
        if (pending_zero) {
            s->lazy_basis = 0;   // |0...0>, written by the first operation
        }
        else {
            launch_set_basis(s, 0);
            KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
        }
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
        kq_pass_stats& c = p.cls[e.cls];
        c.min_ms = c.launches == 0 ? ms : std::min(c.min_ms, static_cast<double>(ms));
        c.max_ms = std::max(c.max_ms, static_cast<double>(ms));
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

int kq_debug_set_faults(size_t alloc_cap_bytes, int fail_status, int fail_after_calls)
{
    if (fail_after_calls < 0 || fail_status < 0) {
        set_error("fault injection: negative call count or status");
        return KQ_ERR_INVALID;
    }
    g_alloc_cap = alloc_cap_bytes;
    g_fail_after = fail_after_calls;
    g_fail_status = fail_status;
    return KQ_OK;
}

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
        *out = create_state(n_qubits, precision, false, true);
    });
}

int kq_dm_create(kq_state** out, uint32_t n_qubits, int precision)
{
    return guarded([&] {
        if (!out) {
            throw InvalidArg {"null output pointer"};
        }
        *out = create_state(n_qubits, precision, true, true);
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
        if (s->spare) {
            (void) hipFree(s->spare);
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
        if (s->adj) {
            (void) hipFree(s->adj);
        }
        delete s;
    });
}

int kq_state_info(const kq_state* s, uint32_t* n_qubits, int* precision, int* is_density)
{
    return guarded([&] {
        check_state_lazy(s);
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
        check_state_lazy(s);
        if (!host) {
            throw InvalidArg {"null host buffer"};
        }
        ensure_init();
        s->bit_of.clear();   // overwritten: qubit q on bit q
        s->lazy_basis = -1;
        s->zero_mask = 0;
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
        check_state_lazy(s);
        const uint64_t N = uint64_t {1} << s->n;
        if (index >= N) {
            throw InvalidArg {"basis index out of range"};
        }
        ensure_init();
        // written by the next operation (a gate run's first pass builds it without reading the buffer)
        s->bit_of.clear();
        s->lazy_basis = static_cast<int64_t>(s->is_dm ? index + index * N : index);
    });
}

int kq_state_copy(kq_state* dst, const kq_state* src)
{
    return guarded([&] {
        check_state_lazy(dst);
        check_state(src);
        if (dst->n_amps != src->n_amps || dst->precision != src->precision || dst->is_dm != src->is_dm) {
            throw InvalidArg {"state copy: shapes differ"};
        }
        ensure_init();
        normalise_layout(src);
        dst->bit_of.clear();
        dst->lazy_basis = -1;
        dst->zero_mask = 0;
        KQ_HIP_CHECK(hipStreamSynchronize(src->stream));
        launch_copy_kernel(dst, dst->amps, src->amps, dst->n_amps);
        KQ_HIP_CHECK(hipStreamSynchronize(dst->stream));
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

int kq_profile_end(kq_pass_stats out[KQ_PASS_CLASSES])
{
    return guarded([&] {
        PassProfiler& p = profiler();
        profiler_collect();
        if (out) {
            for (int i = 0; i < KQ_PASS_CLASSES; ++i) {
                out[i] = p.cls[i];
            }
        }
        p.enabled = false;
    });
}

int kq_stream_probe(uint32_t n_qubits, double gbps[4])
{
    return guarded([&] {
        if (n_qubits < 16 || n_qubits > 32 || !gbps) {
            throw InvalidArg {"stream probe: 16 <= n_qubits <= 32 and a 4-value output"};
        }
        ensure_init();
        double all[24];
        stream_probe(n_qubits, all);
        for (int m = 0; m < 4; ++m) {
            gbps[m] = 0.0;
            for (int sh = 0; sh < 6; ++sh) {
                gbps[m] = std::max(gbps[m], all[4 * sh + m]);
            }
        }
    });
}

int kq_stream_probe_shapes(uint32_t n_qubits, double gbps[24])
{
    return guarded([&] {
        if (n_qubits < 16 || n_qubits > 32 || !gbps) {
            throw InvalidArg {"stream probe: 16 <= n_qubits <= 32 and a 24-value output"};
        }
        ensure_init();
        stream_probe(n_qubits, gbps);
    });
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

int kq_set_jit_nontemporal(int mode)
{
    return guarded([&] {
        JitTuning t = jit_tuning();
        t.nontemporal = mode;
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

int kq_batch_set_all_parameters(kq_batch* b, const double* values)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        batch_set_all_parameters(b, values);
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

int kq_batch_plan_info(const kq_batch* b, size_t* fwd_passes, size_t* rev_passes, double* algorithmic_bytes)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        batch_plan_info(b, fwd_passes, rev_passes, algorithmic_bytes);
    });
}

int kq_batch_set_drift_recompute(kq_batch* b, double drift_tol)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        batch_set_drift_recompute(b, drift_tol);
    });
}

int kq_batch_recomputed(const kq_batch* b, size_t* last_run, size_t* total)
{
    return guarded([&] {
        if (!b) {
            throw InvalidArg {"batch: null handle"};
        }
        batch_recomputed(b, last_run, total);
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

}  // extern "C"
