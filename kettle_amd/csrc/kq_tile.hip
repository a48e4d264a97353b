// The fused tile pass: the hot kernel of simulate() for gfx950.
//
// One workgroup owns one tile: the 2^k amplitudes that share the values of every qubit outside the
// pass's tile set.  The tile is read from HBM once and written once; all gates of the pass are
// applied in between without touching HBM.
//
// Register blocking.  Each thread holds 16 amplitudes in VGPRs: slot r holds local index
// S_r | T_t, where the four "register bits" of the local index are spread over r and the other
// k - 4 bits come from the thread index.  A pass is a sequence of steps; within a step every
// non-diagonal gate acts on a register bit, so it is straight-line register arithmetic on the
// pairs (v[r], v[r | 1 << P]).  Between steps whose register bits differ the tile goes through LDS
// once (write, barrier, read).
//
// Gate variants.  The host lowers every gate to a variant code (kind, register position P of the
// target, register position of the control or none) plus lane / workgroup conditions for controls
// on thread bits or outside the tile; each variant is one straight-line block whose slot pattern is
// fixed at compile time, so no per-slot branch exists and the amplitudes stay in the same VGPRs.
// Diagonal gates whose target is a thread or outside bit scale the slots by one factor per lane.
//
// Addressing.  The LDS swizzle (l ^ (((l >> 4) ^ (l >> 8)) & 15), which spreads lanes whose local
// indices differ by 16 or 256 over different banks) and the local -> global index map are both
// XOR-linear in the local index, so every address is (per-thread part) ^ (per-slot part), the
// per-slot parts being uniform values the host precomputes per layout.  HBM I/O uses a layout whose
// thread bits are the lowest local bits, so each wavefront access covers 64 consecutive
// amplitudes.
#include <hip/hip_runtime.h>

#include <cstring>

#include "kq_internal.hpp"

namespace kq
{
namespace
{

constexpr int TILE_BLOCK = 256;
constexpr int MAX_FREE_Q = 48;

template <typename R> struct AmpT;
template <> struct AmpT<double> { using T = double2; };
template <> struct AmpT<float> { using T = float2; };
template <typename R> using Amp = typename AmpT<R>::T;

struct TileArgs
{
    int nfree;
    int n_thr_bits;
    int n_steps;
    uint32_t out_layout;
    uint8_t freepos[MAX_FREE_Q];
};

template <typename R>
struct Cx
{
    R re, im;
};

template <typename R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, R br, R bi)
{
    return {a.re * br - a.im * bi, a.re * bi + a.im * br};
}

// compile-time loop over register slots
template <int I, int N>
struct Unroll
{
    template <typename F>
    __device__ __forceinline__ static void run(F&& f)
    {
        f.template operator()<I>();
        Unroll<I + 1, N>::run(f);
    }
};
template <int N>
struct Unroll<N, N>
{
    template <typename F>
    __device__ __forceinline__ static void run(F&&)
    {
    }
};

// Variant (KIND, P, C): pairs (v[r], v[r | 1 << P]) for every r with bit P clear and, if C > 0,
// bit C - 1 set.  The zero entries of real, axis-rotation and diagonal matrices are skipped.
template <typename R, int KIND, int P, int C>
__device__ __forceinline__ void apply(Cx<R> (&v)[16], const R* __restrict__ m)
{
    if constexpr (C == P + 1) {
        return;   // never emitted by the host
    }
    else {
        const R m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5], m6 = m[6], m7 = m[7];
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            if constexpr (((r >> P) & 1) == 0 && (C == 0 || ((r >> (C - 1)) & 1) == 1)) {
                constexpr int r1 = r | (1 << P);
                const Cx<R> a = v[r];
                const Cx<R> b = v[r1];
                if constexpr (KIND == TK_SWAP) {
                    v[r] = b;
                    v[r1] = a;
                }
                else if constexpr (KIND == TK_REAL) {
                    v[r] = {a.re * m0 + b.re * m2, a.im * m0 + b.im * m2};
                    v[r1] = {a.re * m4 + b.re * m6, a.im * m4 + b.im * m6};
                }
                else if constexpr (KIND == TK_AXIS) {
                    // m00 = m0, m01 = i m3, m10 = i m5, m11 = m6
                    v[r] = {a.re * m0 - b.im * m3, a.im * m0 + b.re * m3};
                    v[r1] = {b.re * m6 - a.im * m5, b.im * m6 + a.re * m5};
                }
                else if constexpr (KIND == TK_DIAG1) {
                    v[r1] = cmul(b, m6, m7);
                }
                else if constexpr (KIND == TK_DIAG) {
                    v[r] = cmul(a, m0, m1);
                    v[r1] = cmul(b, m6, m7);
                }
                else {
                    v[r] = {a.re * m0 - a.im * m1 + b.re * m2 - b.im * m3, a.re * m1 + a.im * m0 + b.re * m3 + b.im * m2};
                    v[r1] = {a.re * m4 - a.im * m5 + b.re * m6 - b.im * m7, a.re * m5 + a.im * m4 + b.re * m7 + b.im * m6};
                }
            }
        });
    }
}

// Uniform diagonal: every slot (C = 0) or the slots with register bit C - 1 set, times (fr, fi).
template <typename R, int C>
__device__ __forceinline__ void scale(Cx<R> (&v)[16], R fr, R fi)
{
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (C == 0 || ((r >> (C - 1)) & 1) == 1) {
            v[r] = cmul(v[r], fr, fi);
        }
    });
}

#define KQ_TILE_CASE(K, P, C) \
    case tile_code(K, P, C): apply<R, K, P, C>(v, m); break;
#define KQ_TILE_CASES_C(K, P) \
    KQ_TILE_CASE(K, P, 0) KQ_TILE_CASE(K, P, 1) KQ_TILE_CASE(K, P, 2) KQ_TILE_CASE(K, P, 3) KQ_TILE_CASE(K, P, 4)
#define KQ_TILE_CASES_P(K) \
    KQ_TILE_CASES_C(K, 0) KQ_TILE_CASES_C(K, 1) KQ_TILE_CASES_C(K, 2) KQ_TILE_CASES_C(K, 3)

template <typename R>
__device__ __forceinline__ void apply_gate(Cx<R> (&v)[16], uint32_t code, const R* __restrict__ m)
{
    switch (code) {
        KQ_TILE_CASES_P(TK_DENSE)
        KQ_TILE_CASES_P(TK_REAL)
        KQ_TILE_CASES_P(TK_AXIS)
        KQ_TILE_CASES_P(TK_DIAG)
        KQ_TILE_CASES_P(TK_DIAG1)
        KQ_TILE_CASES_P(TK_SWAP)
        default: break;
    }
}

template <typename R>
__device__ __forceinline__ void apply_uniform(Cx<R> (&v)[16], uint32_t c, R fr, R fi)
{
    switch (c) {
        case 0: scale<R, 0>(v, fr, fi); break;
        case 1: scale<R, 1>(v, fr, fi); break;
        case 2: scale<R, 2>(v, fr, fi); break;
        case 3: scale<R, 3>(v, fr, fi); break;
        case 4: scale<R, 4>(v, fr, fi); break;
        default: break;
    }
}

__device__ __forceinline__ uint32_t lds_thread_part(const TileLayout& L, uint32_t tid, int nbits)
{
    uint32_t x = 0;
    for (int j = 0; j < nbits; ++j) {
        x ^= ((tid >> j) & 1u) ? L.lds_thr[j] : 0u;
    }
    return x;
}

__device__ __forceinline__ uint64_t hbm_thread_part(const TileLayout& L, uint32_t tid, int nbits)
{
    uint64_t x = 0;
    for (int j = 0; j < nbits; ++j) {
        x |= ((tid >> j) & 1u) ? L.hbm_thr[j] : 0ull;
    }
    return x;
}

template <typename R>
__global__ __launch_bounds__(TILE_BLOCK) void k_tile(Amp<R>* __restrict__ a, const TileArgs d,
                                                     const TileStep* __restrict__ steps,
                                                     const TileLayout* __restrict__ layouts,
                                                     const TileGate* __restrict__ gates)
{
    extern __shared__ unsigned char smem_raw[];
    const uint32_t tid = threadIdx.x;
    const int tb = d.n_thr_bits;

    uint64_t base = 0;
    {
        const uint64_t b = blockIdx.x;
        for (int j = 0; j < d.nfree; ++j) {
            base |= ((b >> j) & 1ull) << d.freepos[j];
        }
    }

    Cx<R> v[16];
    {
        const TileLayout& L = layouts[0];
        const Amp<R>* p = a + (base | hbm_thread_part(L, tid, tb));
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            const Amp<R> x = p[L.hbm_slot[r]];
            v[r] = {x.x, x.y};
        });
    }

    uint32_t cur = 0;
    const auto remap = [&](uint32_t to) {
        const TileLayout& A = layouts[cur];
        const TileLayout& B = layouts[to];
        const uint32_t xa = lds_thread_part(A, tid, tb);
        __syncthreads();
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            Amp<R> x;
            x.x = v[r].re;
            x.y = v[r].im;
            *reinterpret_cast<Amp<R>*>(smem_raw + (xa ^ A.lds_slot[r])) = x;
        });
        const uint32_t xb = lds_thread_part(B, tid, tb);
        __syncthreads();
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            const Amp<R> x = *reinterpret_cast<const Amp<R>*>(smem_raw + (xb ^ B.lds_slot[r]));
            v[r] = {x.x, x.y};
        });
        cur = to;
    };

    for (int si = 0; si < d.n_steps; ++si) {
        const TileStep st = steps[si];
        if (st.layout != cur) {
            remap(st.layout);
        }
        const uint32_t g_end = st.g_begin + st.g_count;
        for (uint32_t gi = st.g_begin; gi < g_end; ++gi) {
            const TileGate& g = gates[gi];
            if ((base & g.glb_ctrl) != g.glb_ctrl) {
                continue;   // a control (or a d0 = 1 diagonal's target) outside the tile is 0
            }
            if ((tid & g.lane_ctrl) != g.lane_ctrl) {
                continue;   // a control on a thread bit is 0 for this lane
            }
            const R* m = reinterpret_cast<const R*>(g.m);
            const uint32_t code = g.code;
            if (code >= TV_UNIFORM_BASE) {
                const bool sel = g.lane_sel ? ((tid & g.lane_sel) != 0) : ((base & g.glb_sel) != 0);
                const R fr = sel ? m[6] : m[0];
                const R fi = sel ? m[7] : m[1];
                apply_uniform<R>(v, code - TV_UNIFORM_BASE, fr, fi);
            }
            else {
                apply_gate<R>(v, code, m);
            }
        }
    }

    if (d.out_layout != cur) {
        remap(d.out_layout);
    }
    {
        const TileLayout& L = layouts[cur];
        Amp<R>* p = a + (base | hbm_thread_part(L, tid, tb));
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            Amp<R> x;
            x.x = v[r].re;
            x.y = v[r].im;
            p[L.hbm_slot[r]] = x;
        });
    }
}

template <typename R>
void set_lds_attr()
{
    static bool done = false;
    if (!done) {
        const int maxdyn = static_cast<int>(device_lds_per_block() - TILE_STATIC_LDS);
        (void) hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile<R>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   maxdyn);
        (void) hipGetLastError();
        done = true;
    }
}

}  // namespace

void launch_tile_pass(kq_state* s, const TileProgram& prog, const void* d_gates, const void* d_steps,
                      const void* d_layouts)
{
    if (prog.k < 4 || prog.k > 12 || prog.n_thr_bits != prog.k - 4 || prog.steps.empty() || prog.layouts.empty()
        || prog.out_layout >= prog.layouts.size()) {
        throw InvalidArg {"tile pass: bad program shape"};
    }
    if (prog.freepos.size() > static_cast<size_t>(MAX_FREE_Q) || prog.freepos.size() + prog.k != s->nvec) {
        throw InvalidArg {"tile pass: free qubits do not complete the state"};
    }
    for (const TileStep& st : prog.steps) {
        if (st.layout >= prog.layouts.size() || st.g_begin + st.g_count > prog.gates.size()) {
            throw InvalidArg {"tile pass: step out of range"};
        }
    }
    TileArgs d {};
    d.nfree = static_cast<int>(prog.freepos.size());
    d.n_thr_bits = prog.n_thr_bits;
    d.n_steps = static_cast<int>(prog.steps.size());
    d.out_layout = prog.out_layout;
    for (size_t j = 0; j < prog.freepos.size(); ++j) {
        d.freepos[j] = prog.freepos[j];
    }
    const uint64_t blocks = uint64_t {1} << d.nfree;
    if (blocks > 0x7fffffffull) {
        throw InvalidArg {"tile pass: grid too large"};
    }
    const size_t lds = (size_t {1} << prog.k) * amp_bytes(s->precision);
    if (lds + TILE_STATIC_LDS > device_lds_per_block()) {
        throw InvalidArg {"tile pass: tile does not fit the workgroup's LDS"};
    }
    const unsigned threads = 1u << prog.n_thr_bits;
    const auto* g = static_cast<const TileGate*>(d_gates);
    const auto* st = static_cast<const TileStep*>(d_steps);
    const auto* ly = static_cast<const TileLayout*>(d_layouts);
    if (s->precision == KQ_C64) {
        set_lds_attr<float>();
        k_tile<float><<<static_cast<unsigned>(blocks), threads, lds, s->stream>>>(static_cast<float2*>(s->amps), d, st, ly, g);
    }
    else {
        set_lds_attr<double>();
        k_tile<double><<<static_cast<unsigned>(blocks), threads, lds, s->stream>>>(static_cast<double2*>(s->amps), d, st, ly, g);
    }
    KQ_HIP_CHECK(hipGetLastError());
}

}  // namespace kq
