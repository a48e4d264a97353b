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
#include "kq_tile_device.hpp"

namespace kq
{
namespace
{

using dev::Amp;
using dev::apply;
using dev::Coef;
using dev::Cx;
using dev::cmul;
using dev::ddot;
using dev::ddot_uniform;
using dev::scale;
using dev::Unroll;
static_assert(int {dev::K_DENSE} == int {TK_DENSE} && int {dev::K_REAL} == int {TK_REAL} && int {dev::K_AXIS} == int {TK_AXIS}
                  && int {dev::K_DIAG} == int {TK_DIAG} && int {dev::K_DIAG1} == int {TK_DIAG1}
                  && int {dev::K_SWAP} == int {TK_SWAP},
              "gate kind values");

constexpr int TILE_BLOCK = 256;
constexpr int MAX_FREE_Q = 48;

struct TileArgs
{
    uint64_t free_mask;      // global qubits outside the tile; block index bit j -> j-th set bit
    int n_thr_bits;
    int n_steps;
    uint32_t out_layout;
    uint32_t input;          // as BatchTileItem::input (0 for single-state passes)
    uint64_t zmask;
};

// a gate record's coefficients: doubles (complex128) or the floats packed in its first 32 bytes
template <typename R>
__device__ __forceinline__ Coef<R> coef_of(const TileGate& g)
{
    Coef<R> c;
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            c.m[i] = g.m[i];
        }
    }
    else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float2 f = __builtin_bit_cast(float2, g.m[i]);
            c.m[2 * i] = f.x;
            c.m[2 * i + 1] = f.y;
        }
    }
    return c;
}

#define KQ_TILE_CASE(K, P, C) \
    case tile_code(K, P, C): apply<R, K, P, C>(v, m); break;
#define KQ_TILE_CASES_C(K, P) \
    KQ_TILE_CASE(K, P, 0) KQ_TILE_CASE(K, P, 1) KQ_TILE_CASE(K, P, 2) KQ_TILE_CASE(K, P, 3) KQ_TILE_CASE(K, P, 4)
#define KQ_TILE_CASES_P(K) \
    KQ_TILE_CASES_C(K, 0) KQ_TILE_CASES_C(K, 1) KQ_TILE_CASES_C(K, 2) KQ_TILE_CASES_C(K, 3)

template <typename R>
__device__ __forceinline__ void apply_gate(Cx<R> (&v)[16], uint32_t code, const Coef<R>& m)
{
    switch (code) {
        KQ_TILE_CASES_P(TK_DENSE)
        KQ_TILE_CASES_P(TK_REAL)
        KQ_TILE_CASES_P(TK_AXIS)
        KQ_TILE_CASES_P(TK_DIAG)
        KQ_TILE_CASES_P(TK_DIAG1)
        KQ_TILE_CASES_P(TK_SWAP)
        // the host emits only the codes above: no fall-through edge, so no path merges the
        // unchanged amplitudes with a changed copy (which would cost register moves per gate)
        default: __builtin_unreachable();
    }
}

#define KQ_NORM_CASES(P) \
    case 0 * 4 + (P): orth<R, P, 0>(v, m.m[0]); break; \
    case 1 * 4 + (P): orth<R, P, 1>(v, m.m[0]); break; \
    case 2 * 4 + (P): orth<R, P, 2>(v, m.m[0]); break; \
    case 3 * 4 + (P): orth<R, P, 3>(v, m.m[0]); break; \
    case 4 * 4 + (P): axis<R, P, 0>(v, m.m[0]); break; \
    case 5 * 4 + (P): axis<R, P, 1>(v, m.m[0]); break; \
    case 6 * 4 + (P): dense1<R, P, 0>(v, m.m[0], m.m[1], m.m[2], m.m[3], m.m[4], m.m[5]); break; \
    case 7 * 4 + (P): dense1<R, P, 1>(v, m.m[0], m.m[1], m.m[2], m.m[3], m.m[4], m.m[5]); break;

// an uncontrolled one-qubit gate with its pivot divided out (TV_NORM_BASE variants): fp = form * 4 + P
template <typename R>
__device__ __forceinline__ void apply_norm(Cx<R> (&v)[16], uint32_t fp, const Coef<R>& m)
{
    switch (fp) {
        KQ_NORM_CASES(0) KQ_NORM_CASES(1) KQ_NORM_CASES(2) KQ_NORM_CASES(3)
        default: __builtin_unreachable();
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
        default: __builtin_unreachable();
    }
}

#define KQ_DDOT_CASE(P, C) \
    case (P) * 5 + (C): return ddot<R, P, C>(v, w, m);

template <typename R>
__device__ __forceinline__ double ddot_gate(const Cx<R> (&v)[16], const Cx<R> (&w)[16], uint32_t pc, const Coef<R>& m)
{
    switch (pc) {
        KQ_DDOT_CASE(0, 0) KQ_DDOT_CASE(0, 2) KQ_DDOT_CASE(0, 3) KQ_DDOT_CASE(0, 4)
        KQ_DDOT_CASE(1, 0) KQ_DDOT_CASE(1, 1) KQ_DDOT_CASE(1, 3) KQ_DDOT_CASE(1, 4)
        KQ_DDOT_CASE(2, 0) KQ_DDOT_CASE(2, 1) KQ_DDOT_CASE(2, 2) KQ_DDOT_CASE(2, 4)
        KQ_DDOT_CASE(3, 0) KQ_DDOT_CASE(3, 1) KQ_DDOT_CASE(3, 2) KQ_DDOT_CASE(3, 3)
        default: break;
    }
    return 0.0;
}

#define KQ_DGEN_CASE(G, P, C) \
    case (G) * 20 + (P) * 5 + (C): return ddot_gen<R, G, P, C>(v, w);
#define KQ_DGEN_ALL(G) \
    KQ_DGEN_CASE(G, 0, 0) KQ_DGEN_CASE(G, 0, 2) KQ_DGEN_CASE(G, 0, 3) KQ_DGEN_CASE(G, 0, 4) \
    KQ_DGEN_CASE(G, 1, 0) KQ_DGEN_CASE(G, 1, 1) KQ_DGEN_CASE(G, 1, 3) KQ_DGEN_CASE(G, 1, 4) \
    KQ_DGEN_CASE(G, 2, 0) KQ_DGEN_CASE(G, 2, 1) KQ_DGEN_CASE(G, 2, 2) KQ_DGEN_CASE(G, 2, 4) \
    KQ_DGEN_CASE(G, 3, 0) KQ_DGEN_CASE(G, 3, 1) KQ_DGEN_CASE(G, 3, 2) KQ_DGEN_CASE(G, 3, 3)

// an angle gate's fixed generator (TV_DGEN_BASE variants): gc = gen * 20 + P * 5 + C
template <typename R>
__device__ __forceinline__ double ddot_gen_gate(const Cx<R> (&v)[16], const Cx<R> (&w)[16], uint32_t gc)
{
    switch (gc) {
        KQ_DGEN_ALL(0) KQ_DGEN_ALL(1) KQ_DGEN_ALL(2) KQ_DGEN_ALL(3)
        default: break;
    }
    return 0.0;
}

template <typename R>
__device__ __forceinline__ double ddot_uniform_gate(const Cx<R> (&v)[16], const Cx<R> (&w)[16], uint32_t c, R fr, R fi)
{
    switch (c) {
        case 0: return ddot_uniform<R, 0>(v, w, fr, fi);
        case 1: return ddot_uniform<R, 1>(v, w, fr, fi);
        case 2: return ddot_uniform<R, 2>(v, w, fr, fi);
        case 3: return ddot_uniform<R, 3>(v, w, fr, fi);
        case 4: return ddot_uniform<R, 4>(v, w, fr, fi);
        default: break;
    }
    return 0.0;
}

// the four coefficients of a merged-run term (d0, d1), doubles or the floats packed first
template <typename R>
__device__ __forceinline__ void term_coefs(const TileTerm& t, R& d0r, R& d0i, R& d1r, R& d1i)
{
    if constexpr (sizeof(R) == 8) {
        d0r = t.d[0];
        d0i = t.d[1];
        d1r = t.d[2];
        d1i = t.d[3];
    }
    else {
        const float2 a = __builtin_bit_cast(float2, t.d[0]);
        const float2 b = __builtin_bit_cast(float2, t.d[1]);
        d0r = a.x;
        d0i = a.y;
        d1r = b.x;
        d1i = b.y;
    }
}

// Per-lane product of a merged run's terms: f1 for slots whose target bit is 1 (or every slot for a
// uniform target), f0 for target bit 0.
template <typename R>
__device__ __forceinline__ void merged_factors(const TileTerm* __restrict__ terms, uint32_t first, uint32_t count,
                                               bool uniform, uint64_t base, uint32_t tid, R& f0r, R& f0i, R& f1r,
                                               R& f1i)
{
    f0r = 1;
    f0i = 0;
    f1r = 1;
    f1i = 0;
    for (uint32_t j = 0; j < count; ++j) {
        const TileTerm t = terms[first + j];
        R d0r, d0i, d1r, d1i;
        term_coefs<R>(t, d0r, d0i, d1r, d1i);
        const bool ok = (base & t.glb_ctrl) == t.glb_ctrl && (t.lane_ctrl == 0 || (tid & t.lane_ctrl) == t.lane_ctrl);
        if (uniform) {
            const bool sel = t.lane_sel ? ((tid & t.lane_sel) != 0) : ((base & t.glb_sel) != 0);
            const R fr = sel ? d1r : d0r;
            const R fi = sel ? d1i : d0i;
            const R nr = f1r * fr - f1i * fi;
            const R ni = f1r * fi + f1i * fr;
            f1r = ok ? nr : f1r;
            f1i = ok ? ni : f1i;
        }
        else {
            const R n0r = f0r * d0r - f0i * d0i;
            const R n0i = f0r * d0i + f0i * d0r;
            const R n1r = f1r * d1r - f1i * d1i;
            const R n1i = f1r * d1i + f1i * d1r;
            f0r = ok ? n0r : f0r;
            f0i = ok ? n0i : f0i;
            f1r = ok ? n1r : f1r;
            f1i = ok ? n1i : f1i;
        }
    }
}

#define KQ_DIAG_CASE(P, C) \
    case (P) * 5 + (C): apply<R, TK_DIAG, P, C>(v, m); break;

// a diagonal with per-lane factors on register position P with register control C (key = P * 5 + C)
template <typename R>
__device__ __forceinline__ void apply_diag_key(Cx<R> (&v)[16], uint32_t key, const Coef<R>& m)
{
    switch (key) {
        KQ_DIAG_CASE(0, 0) KQ_DIAG_CASE(0, 2) KQ_DIAG_CASE(0, 3) KQ_DIAG_CASE(0, 4)
        KQ_DIAG_CASE(1, 0) KQ_DIAG_CASE(1, 1) KQ_DIAG_CASE(1, 3) KQ_DIAG_CASE(1, 4)
        KQ_DIAG_CASE(2, 0) KQ_DIAG_CASE(2, 1) KQ_DIAG_CASE(2, 2) KQ_DIAG_CASE(2, 4)
        KQ_DIAG_CASE(3, 0) KQ_DIAG_CASE(3, 1) KQ_DIAG_CASE(3, 2) KQ_DIAG_CASE(3, 3)
        default: __builtin_unreachable();
    }
}

// apply a merged diagonal run to one register block
template <typename R>
__device__ __forceinline__ void apply_merged(Cx<R> (&v)[16], uint32_t key, R f0r, R f0i, R f1r, R f1i)
{
    if (key >= 20) {
        apply_uniform<R>(v, key - 20, f1r, f1i);
    }
    else {
        Coef<R> m;
        m.m[0] = f0r;
        m.m[1] = f0i;
        m.m[2] = m.m[3] = m.m[4] = m.m[5] = 0;
        m.m[6] = f1r;
        m.m[7] = f1i;
        apply_diag_key<R>(v, key, m);
    }
}

// Thread parts of a layout's addresses.  The per-bit tables are read unconditionally (one batch of
// scalar loads) and selected per lane.
__device__ __forceinline__ uint32_t lds_thread_part(const TileLayout& L, uint32_t tid, int nbits)
{
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t wj = L.lds_thr[j];
        x ^= (j < nbits && ((tid >> j) & 1u)) ? wj : 0u;
    }
    return x;
}

__device__ __forceinline__ uint64_t hbm_thread_part(const TileLayout& L, uint32_t tid, int nbits)
{
    uint64_t x = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint64_t wj = L.hbm_thr[j];
        x |= (j < nbits && ((tid >> j) & 1u)) ? wj : 0ull;
    }
    return x;
}

// Sum of `x` over the first `width` lanes of the wavefront (width a power of two <= 64); every
// lane of the wave must be active.
__device__ __forceinline__ double wave_sum(double x, int width)
{
    for (int off = 1; off < width; off <<= 1) {
        x += __shfl_xor(x, off, 64);
    }
    return x;
}

// ADJ = false: one state, gates applied in order.
// ADJ = true: the reverse sweep of the adjoint gradient over two states, psi (a) and lambda (b);
// the gate records are the inverse gates G^dag in reverse circuit order.  For each gate: if the record
// has a derivative record Q (pad = index + 1; dG = Q G), Re <lambda| Q |psi> over the tile is reduced
// per wavefront into partials[slot * n_part + block * waves + wave]; then psi <- G^dag psi and
// lambda <- G^dag lambda.
template <typename R, bool ADJ>
__device__ __forceinline__ void tile_body(Amp<R>* __restrict__ a, Amp<R>* __restrict__ b, const TileArgs& d,
                                          const TileStep* __restrict__ steps, const TileLayout* __restrict__ layouts,
                                          const TileGate* __restrict__ gates, const TileGate* __restrict__ derivs,
                                          const TileTerm* __restrict__ terms, double* __restrict__ partials,
                                          uint32_t n_part, uint32_t block)
{
    extern __shared__ unsigned char smem_raw[];
    const uint32_t tid = threadIdx.x;
    const int tb = d.n_thr_bits;

    // deposit the block index into the free qubits (uniform: scalar instructions only)
    uint64_t base = 0;
    {
        uint64_t blk = block;
        uint64_t mask = d.free_mask;
        while (mask != 0) {
            const int q = __builtin_ctzll(mask);
            base |= (blk & 1ull) << q;
            blk >>= 1;
            mask &= mask - 1;
        }
    }

    Cx<R> v[16];
    Cx<R> w[ADJ ? 16 : 1];
    if (!ADJ && d.input == 1) {
        // |0...0> not written yet: every tile but tile 0 is zero and stays zero under the pass's gates
        // (linear, within the tile), so its workgroup stores zeros where the pass stores and ends
        if (base != 0) {
            const TileLayout& S = layouts[d.out_layout];
            const uint64_t off = base | hbm_thread_part(S, tid, tb);
            Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
                Amp<R> z;
                z.x = 0;
                z.y = 0;
                a[off + S.hbm_slot[r]] = z;
            });
            return;
        }
        const TileLayout& L = layouts[0];
        const uint64_t off = hbm_thread_part(L, tid, tb);
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            v[r] = {(off + L.hbm_slot[r]) == 0 ? R(1) : R(0), R(0)};
        });
    }
    else {
        if (!ADJ && d.input == 2 && (base & d.zmask) != 0) {
            return;   // a zero tile, whose zeros the buffer holds
        }
        const TileLayout& L = layouts[0];
        const uint64_t off = base | hbm_thread_part(L, tid, tb);
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            const Amp<R> x = a[off + L.hbm_slot[r]];
            v[r] = {x.x, x.y};
            if constexpr (ADJ) {
                const Amp<R> y = b[off + L.hbm_slot[r]];
                w[r] = {y.x, y.y};
            }
        });
    }

    uint32_t cur = 0;
    const auto remap_one = [&](Cx<R> (&u)[16], const TileLayout& A, const TileLayout& B, uint32_t xa, uint32_t xb) {
        if constexpr (sizeof(R) == 8) {
            // complex128: the real parts, then the imaginary parts, through 8-byte entries (the layouts'
            // 16-byte offsets halved: the same swizzle in half the LDS, so more workgroups per CU)
            const uint32_t ha = xa >> 1, hb = xb >> 1;
            const auto half = [&]<bool IM>() __attribute__((always_inline)) {
                __syncthreads();
                Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
                    *reinterpret_cast<R*>(smem_raw + (ha ^ (A.lds_slot[r] >> 1))) = IM ? u[r].im : u[r].re;
                });
                __syncthreads();
                Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
                    const R x = *reinterpret_cast<const R*>(smem_raw + (hb ^ (B.lds_slot[r] >> 1)));
                    if constexpr (IM) {
                        u[r].im = x;
                    }
                    else {
                        u[r].re = x;
                    }
                });
            };
            half.template operator()<false>();
            half.template operator()<true>();
        }
        else {
            __syncthreads();
            Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
                Amp<R> x;
                x.x = u[r].re;
                x.y = u[r].im;
                *reinterpret_cast<Amp<R>*>(smem_raw + (xa ^ A.lds_slot[r])) = x;
            });
            __syncthreads();
            Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
                const Amp<R> x = *reinterpret_cast<const Amp<R>*>(smem_raw + (xb ^ B.lds_slot[r]));
                u[r] = {x.x, x.y};
            });
        }
    };
    const auto remap = [&](uint32_t to) {
        const TileLayout& A = layouts[cur];
        const TileLayout& B = layouts[to];
        const uint32_t xa = lds_thread_part(A, tid, tb);
        const uint32_t xb = lds_thread_part(B, tid, tb);
        remap_one(v, A, B, xa, xb);
        if constexpr (ADJ) {
            remap_one(w, A, B, xa, xb);
        }
        cur = to;
    };

    const int width = (1 << tb) < 64 ? (1 << tb) : 64;
    const uint32_t waves = (blockDim.x + 63) / 64;
    const uint32_t part_col = block * waves + tid / 64;
    // Derivative partials of whole wavefronts are staged: each lane writes its value into a wave-private
    // LDS area after the tile (TILE_STAGE derivatives x 64 lanes, and their slots); when it is full and at
    // the end the wave sums them together, 8 lanes per derivative over 8 values each, then 3 shuffle
    // steps, instead of a 6-step shuffle tree per derivative.  The fences order the area's writes and
    // reads within the wave.
    const bool staged = ADJ && width == 64;
    double* const stage = reinterpret_cast<double*>(smem_raw + (128u << tb) + (tid >> 6) * TILE_STAGE_BYTES);
    uint32_t* const stage_slot = reinterpret_cast<uint32_t*>(stage + TILE_STAGE * 64);
    const uint32_t lane = tid & 63u;
    uint32_t n_staged = 0;
    const auto flush = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const uint32_t j = lane >> 3;
        double x = 0.0;
        if (j < n_staged) {
            const double* q = stage + j * 64u + (lane & 7u) * 8u;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                x += q[i];
            }
        }
        x += __shfl_xor(x, 1, 64);
        x += __shfl_xor(x, 2, 64);
        x += __shfl_xor(x, 4, 64);
        if ((lane & 7u) == 0u && j < n_staged) {
            partials[static_cast<size_t>(stage_slot[j]) * n_part + part_col] = x;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        n_staged = 0;
    };

    for (int si = 0; si < d.n_steps; ++si) {
        const TileStep st = steps[si];
        if (st.layout != cur) {
            remap(st.layout);
        }
        const TileGate* __restrict__ gp = gates + st.g_begin;
        const uint32_t g_count = st.g_count;
        if (g_count == 0) {
            continue;   // a gate-free step (a pass that only moves qubits to other bits)
        }
        TileGate next = gp[0];
        for (uint32_t gi = 0; gi < g_count; ++gi) {
            // the record is copied into scalar registers once; the next one is fetched meanwhile
            const TileGate g = next;
            if (gi + 1 < g_count) {
                next = gp[gi + 1];
            }
            const Coef<R> m = coef_of<R>(g);
            const uint32_t code = g.code;
            const bool glb_ok = (base & g.glb_ctrl) == g.glb_ctrl;
            // lane condition only for controls on thread bits (uniform test first: no exec masking)
            const bool lane_ok = g.lane_ctrl == 0 || (tid & g.lane_ctrl) == g.lane_ctrl;
            const auto uniform_factor = [&](R& fr, R& fi) {
                const bool sel = g.lane_sel ? ((tid & g.lane_sel) != 0) : ((base & g.glb_sel) != 0);
                fr = sel ? m.m[6] : m.m[0];
                fi = sel ? m.m[7] : m.m[1];
            };
            if constexpr (!ADJ) {
                if (!glb_ok) {
                    continue;   // a control (or a d0 = 1 diagonal's target) outside the tile is 0
                }
                if (!lane_ok) {
                    continue;   // a control on a thread bit is 0 for this lane
                }
                if (code >= TV_NORM_BASE) {
                    apply_norm<R>(v, code - TV_NORM_BASE, m);
                }
                else if (code >= TV_MERGE_BASE) {
                    const uint32_t key = code - TV_MERGE_BASE;
                    R f0r, f0i, f1r, f1i;
                    merged_factors<R>(terms, g.lane_sel, static_cast<uint32_t>(g.glb_sel), key >= 20, base, tid, f0r, f0i,
                                      f1r, f1i);
                    apply_merged<R>(v, key, f0r, f0i, f1r, f1i);
                }
                else if (code >= TV_UNIFORM_BASE) {
                    R fr, fi;
                    uniform_factor(fr, fi);
                    apply_uniform<R>(v, code - TV_UNIFORM_BASE, fr, fi);
                }
                else {
                    apply_gate<R>(v, code, m);
                }
            }
            else {
                const bool ok = glb_ok && lane_ok;
                const bool norm = code >= TV_NORM_BASE;
                R fr = 0, fi = 0, f0r = 0, f0i = 0;
                if (!norm && code >= TV_MERGE_BASE) {
                    merged_factors<R>(terms, g.lane_sel, static_cast<uint32_t>(g.glb_sel), code - TV_MERGE_BASE >= 20, base,
                                      tid, f0r, f0i, fr, fi);
                }
                else if (!norm && code >= TV_UNIFORM_BASE) {
                    uniform_factor(fr, fi);
                }
                if (g.pad != 0) {
                    const TileGate dg = derivs[g.pad - 1];
                    const Coef<R> dm = coef_of<R>(dg);
                    double acc = 0.0;
                    if ((base & dg.glb_ctrl) == dg.glb_ctrl && (tid & dg.lane_ctrl) == dg.lane_ctrl) {
                        if (dg.code >= TV_DGEN_BASE) {
                            acc = ddot_gen_gate<R>(v, w, dg.code - TV_DGEN_BASE);
                        }
                        else if (dg.code >= TV_UNIFORM_BASE) {
                            const bool sel = dg.lane_sel ? ((tid & dg.lane_sel) != 0) : ((base & dg.glb_sel) != 0);
                            acc = ddot_uniform_gate<R>(v, w, dg.code - TV_UNIFORM_BASE, sel ? dm.m[6] : dm.m[0],
                                                       sel ? dm.m[7] : dm.m[1]);
                        }
                        else {
                            acc = ddot_gate<R>(v, w, dg.code % 20, dm);
                        }
                    }
                    if (staged) {
                        stage[n_staged * 64u + lane] = acc;
                        if (lane == 0) {
                            stage_slot[n_staged] = dg.pad;
                        }
                        if (++n_staged == TILE_STAGE) {
                            flush();
                        }
                    }
                    else {
                        acc = wave_sum(acc, width);
                        if ((tid & 63u) == 0) {
                            partials[static_cast<size_t>(dg.pad) * n_part + part_col] = acc;
                        }
                    }
                }
                if (ok) {
                    if (norm) {
                        apply_norm<R>(v, code - TV_NORM_BASE, m);
                    }
                    else if (code >= TV_MERGE_BASE) {
                        apply_merged<R>(v, code - TV_MERGE_BASE, f0r, f0i, fr, fi);
                    }
                    else if (code >= TV_UNIFORM_BASE) {
                        apply_uniform<R>(v, code - TV_UNIFORM_BASE, fr, fi);
                    }
                    else {
                        apply_gate<R>(v, code, m);
                    }
                }
                if (ok) {
                    if (norm) {
                        apply_norm<R>(w, code - TV_NORM_BASE, m);
                    }
                    else if (code >= TV_MERGE_BASE) {
                        apply_merged<R>(w, code - TV_MERGE_BASE, f0r, f0i, fr, fi);
                    }
                    else if (code >= TV_UNIFORM_BASE) {
                        apply_uniform<R>(w, code - TV_UNIFORM_BASE, fr, fi);
                    }
                    else {
                        apply_gate<R>(w, code, m);
                    }
                }
            }
        }
    }

    if constexpr (ADJ) {
        if (staged && n_staged > 0) {
            flush();
        }
    }
    if (d.out_layout != cur) {
        remap(d.out_layout);
    }
    {
        const TileLayout& L = layouts[cur];
        const uint64_t off = base | hbm_thread_part(L, tid, tb);
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            Amp<R> x;
            x.x = v[r].re;
            x.y = v[r].im;
            a[off + L.hbm_slot[r]] = x;
            if constexpr (ADJ) {
                Amp<R> y;
                y.x = w[r].re;
                y.y = w[r].im;
                b[off + L.hbm_slot[r]] = y;
            }
        });
    }
}

template <typename R, bool ADJ>
__global__ __launch_bounds__(TILE_BLOCK) void k_tile(Amp<R>* __restrict__ a, Amp<R>* __restrict__ b, const TileArgs d,
                                                     const TileStep* __restrict__ steps,
                                                     const TileLayout* __restrict__ layouts,
                                                     const TileGate* __restrict__ gates,
                                                     const TileGate* __restrict__ derivs, const TileTerm* __restrict__ terms,
                                                     double* __restrict__ partials, uint32_t n_part)
{
    tile_body<R, ADJ>(a, b, d, steps, layouts, gates, derivs, terms, partials, n_part, blockIdx.x);
}

// Many states in one launch: item i covers blocks [block_begin, next item's block_begin) and runs its
// own pass program on its own state (and lambda) in the arenas.  Items are sorted by block_begin.
template <typename R, bool ADJ>
__global__ __launch_bounds__(TILE_BLOCK) void k_tile_batch(Amp<R>* __restrict__ arena_a, Amp<R>* __restrict__ arena_b,
                                                           const BatchTileItem* __restrict__ items, uint32_t n_items)
{
    uint32_t lo = 0, hi = n_items;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (items[mid].block_begin <= blockIdx.x) {
            lo = mid;
        }
        else {
            hi = mid;
        }
    }
    const BatchTileItem& it = items[lo];
    const TileArgs d {it.free_mask, it.n_thr_bits, it.n_steps, it.out_layout, ADJ ? 0u : it.input, it.zmask};
    // the tables live in global memory: address-space-1 pointers keep their reads scalar
    using GStep = const __attribute__((address_space(1))) TileStep;
    using GLayout = const __attribute__((address_space(1))) TileLayout;
    using GGate = const __attribute__((address_space(1))) TileGate;
    using GTerm = const __attribute__((address_space(1))) TileTerm;
    using GDouble = __attribute__((address_space(1))) double;
    tile_body<R, ADJ>(arena_a + it.a_off, ADJ ? arena_b + it.b_off : nullptr, d,
                      (const TileStep*) (GStep*) it.steps, (const TileLayout*) (GLayout*) it.layouts,
                      (const TileGate*) (GGate*) it.gates, (const TileGate*) (GGate*) it.derivs,
                      (const TileTerm*) (GTerm*) it.terms, (double*) (GDouble*) it.partials, it.n_part,
                      blockIdx.x - it.block_begin);
}

template <typename R, bool ADJ>
void set_lds_attr()
{
    static bool done = false;
    if (!done) {
        const int maxdyn = static_cast<int>(device_lds_per_block() - TILE_STATIC_LDS);
        (void) hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile<R, ADJ>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, maxdyn);
        (void) hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_batch<R, ADJ>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, maxdyn);
        (void) hipGetLastError();
        done = true;
    }
}

TileArgs check_program(const kq_state* s, const TileProgram& prog)
{
    if (prog.k < 4 || prog.k > 12 || prog.n_thr_bits != prog.k - 4 || prog.steps.empty() || prog.layouts.empty()
        || prog.out_layout >= prog.layouts.size()) {
        throw InvalidArg {"tile pass: bad program shape"};
    }
    if (prog.freepos.size() > static_cast<size_t>(MAX_FREE_Q) || prog.freepos.size() + prog.k != s->nvec) {
        throw InvalidArg {"tile pass: free qubits do not complete the state"};
    }
    check_tile_records(prog);
    TileArgs d {};
    d.n_thr_bits = prog.n_thr_bits;
    d.n_steps = static_cast<int>(prog.steps.size());
    d.out_layout = prog.out_layout;
    for (const uint8_t q : prog.freepos) {
        if (q >= 64 || ((d.free_mask >> q) & 1ull)) {
            throw InvalidArg {"tile pass: bad free qubit list"};
        }
        d.free_mask |= uint64_t {1} << q;
    }
    if (prog.freepos.size() > 31) {
        throw InvalidArg {"tile pass: grid too large"};
    }
    if ((size_t {1} << prog.k) * TILE_LDS_ENTRY + TILE_STATIC_LDS > device_lds_per_block()) {
        throw InvalidArg {"tile pass: tile does not fit the workgroup's LDS"};
    }
    return d;
}

}  // namespace

void check_tile_records(const TileProgram& prog)
{
    for (const TileStep& st : prog.steps) {
        if (st.layout >= prog.layouts.size() || st.g_begin + st.g_count > prog.gates.size()) {
            throw InvalidArg {"tile pass: step out of range"};
        }
    }
    for (const TileGate& g : prog.gates) {
        if (g.pad > prog.derivs.size()) {
            throw InvalidArg {"tile pass: derivative index out of range"};
        }
        const bool norm = g.code >= TV_NORM_BASE && g.code < TV_NORM_END;
        if ((g.code >= TV_END && !norm) || (g.code < TV_UNIFORM_BASE && g.code % 5 == (g.code / 5) % 4 + 1)) {
            throw InvalidArg {"tile pass: unknown gate variant"};
        }
        if (!norm && g.code >= TV_MERGE_BASE && (static_cast<uint64_t>(g.lane_sel) + g.glb_sel > prog.terms.size())) {
            throw InvalidArg {"tile pass: merged run out of range"};
        }
    }
}

// dynamic LDS of an adjoint tile pass: the tile, then the derivative staging area of each whole wavefront
size_t adjoint_tile_lds(int k)
{
    const size_t threads = size_t {1} << (k - 4);
    return (size_t {1} << k) * TILE_LDS_ENTRY + (threads >= 64 ? threads / 64 * TILE_STAGE_BYTES : 0);
}

uint32_t tile_pass_blocks(const TileProgram& prog) { return 1u << prog.freepos.size(); }
uint32_t tile_pass_waves(const TileProgram& prog) { return ((1u << prog.n_thr_bits) + 63) / 64; }

void launch_tile_pass(kq_state* s, const TileProgram& prog, const void* d_gates, const void* d_steps,
                      const void* d_layouts, const void* d_terms)
{
    const TileArgs d = check_program(s, prog);
    if (!prog.derivs.empty()) {
        throw InvalidArg {"tile pass: derivative records need the adjoint launch"};
    }
    const unsigned blocks = tile_pass_blocks(prog);
    const unsigned threads = 1u << prog.n_thr_bits;
    const size_t lds = (size_t {1} << prog.k) * TILE_LDS_ENTRY;
    const auto* g = static_cast<const TileGate*>(d_gates);
    const auto* st = static_cast<const TileStep*>(d_steps);
    const auto* ly = static_cast<const TileLayout*>(d_layouts);
    const auto* tm = static_cast<const TileTerm*>(d_terms);
    if (s->precision == KQ_C64) {
        set_lds_attr<float, false>();
        k_tile<float, false><<<blocks, threads, lds, s->stream>>>(static_cast<float2*>(s->amps), nullptr, d, st, ly, g,
                                                                   nullptr, tm, nullptr, 0);
    }
    else {
        set_lds_attr<double, false>();
        k_tile<double, false><<<blocks, threads, lds, s->stream>>>(static_cast<double2*>(s->amps), nullptr, d, st, ly, g,
                                                                    nullptr, tm, nullptr, 0);
    }
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_tile_batch(hipStream_t stream, int precision, bool adjoint, void* arena_a, void* arena_b,
                       const BatchTileItem* d_items, uint32_t n_items, uint32_t n_blocks, int k)
{
    if (k < 4 || k > 12 || n_items == 0 || n_blocks == 0) {
        throw InvalidArg {"tile batch: bad launch shape"};
    }
    const unsigned threads = 1u << (k - 4);
    const size_t lds = adjoint ? adjoint_tile_lds(k) : (size_t {1} << k) * TILE_LDS_ENTRY;
    if (lds + TILE_STATIC_LDS > device_lds_per_block()) {
        throw InvalidArg {"tile batch: tile does not fit the workgroup's LDS"};
    }
    if (precision == KQ_C64) {
        if (adjoint) {
            set_lds_attr<float, true>();
            k_tile_batch<float, true><<<n_blocks, threads, lds, stream>>>(static_cast<float2*>(arena_a), static_cast<float2*>(arena_b),
                                                                          d_items, n_items);
        }
        else {
            set_lds_attr<float, false>();
            k_tile_batch<float, false><<<n_blocks, threads, lds, stream>>>(static_cast<float2*>(arena_a), nullptr, d_items, n_items);
        }
    }
    else {
        if (adjoint) {
            set_lds_attr<double, true>();
            k_tile_batch<double, true><<<n_blocks, threads, lds, stream>>>(static_cast<double2*>(arena_a),
                                                                           static_cast<double2*>(arena_b), d_items, n_items);
        }
        else {
            set_lds_attr<double, false>();
            k_tile_batch<double, false><<<n_blocks, threads, lds, stream>>>(static_cast<double2*>(arena_a), nullptr, d_items,
                                                                            n_items);
        }
    }
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_adjoint_tile_pass(kq_state* s, void* psi, void* lam, const TileProgram& prog, const void* d_gates,
                              const void* d_steps, const void* d_layouts, const void* d_derivs, const void* d_terms,
                              double* d_partials)
{
    const TileArgs d = check_program(s, prog);
    const unsigned blocks = tile_pass_blocks(prog);
    const unsigned threads = 1u << prog.n_thr_bits;
    const uint32_t n_part = blocks * tile_pass_waves(prog);
    const size_t lds = adjoint_tile_lds(prog.k);
    if (lds + TILE_STATIC_LDS > device_lds_per_block()) {
        throw InvalidArg {"tile pass: tile and derivative staging do not fit the workgroup's LDS"};
    }
    const auto* g = static_cast<const TileGate*>(d_gates);
    const auto* st = static_cast<const TileStep*>(d_steps);
    const auto* ly = static_cast<const TileLayout*>(d_layouts);
    const auto* dv = static_cast<const TileGate*>(d_derivs);
    const auto* tm = static_cast<const TileTerm*>(d_terms);
    if (s->precision == KQ_C64) {
        set_lds_attr<float, true>();
        k_tile<float, true><<<blocks, threads, lds, s->stream>>>(static_cast<float2*>(psi), static_cast<float2*>(lam), d, st,
                                                                  ly, g, dv, tm, d_partials, n_part);
    }
    else {
        set_lds_attr<double, true>();
        k_tile<double, true><<<blocks, threads, lds, s->stream>>>(static_cast<double2*>(psi), static_cast<double2*>(lam), d,
                                                                   st, ly, g, dv, tm, d_partials, n_part);
    }
    KQ_HIP_CHECK(hipGetLastError());
}

}  // namespace kq

