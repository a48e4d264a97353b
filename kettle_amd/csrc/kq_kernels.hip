// Device kernels of libkq for gfx950 (CDNA4, wave64) and their host-side launchers.
//
// Storage: a state is 2^nvec complex amplitudes, interleaved (re, im), FP64 (KQ_C128) or FP32
// (KQ_C64).  Qubit q is bit q of the amplitude index (kettle's little-endian convention,
// state/qubit_state_conversion.cpp:37-83).  A density matrix is stored column-major as a
// 2n-qubit vector: bit q is row qubit q, bit n+q is column qubit q.
//
// All reductions (norms, probabilities, expectations, gradients) accumulate in FP64 whatever the
// amplitude precision, and are deterministic: fixed per-block partials, then one fixed-order
// final reduction (no floating-point atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "kq_internal.hpp"

// The allocation-free sampler's workspace (kq::NOALLOC_SCRATCH_BYTES per device).  In host code this
// name is only a stand-in for the symbol: kernels and copies use the device address that
// kq::noalloc_workspace() obtains with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kq::NOALLOC_SCRATCH_BYTES];

namespace kq
{
namespace
{

constexpr int BLOCK = 256;
constexpr int RED_BLOCKS = REDUCTION_BLOCKS;   // upper bound on partial-sum blocks of a reduction

template <typename R> struct AmpT;
template <> struct AmpT<double> { using T = double2; };
template <> struct AmpT<float> { using T = float2; };
template <typename R> using Amp = typename AmpT<R>::T;

struct DevGate
{
    uint32_t kind;
    uint32_t target;
    int32_t control;
    uint32_t pad;
    double m[8];
};

// ---- small device helpers -----------------------------------------------------------------

__device__ __forceinline__ uint64_t insert_zero(uint64_t p, uint32_t j)
{
    const uint64_t low = p & ((uint64_t {1} << j) - 1);
    return ((p >> j) << (j + 1)) | low;
}

// insert zeros at two distinct positions lo < hi
__device__ __forceinline__ uint64_t insert_zero2(uint64_t p, uint32_t lo, uint32_t hi)
{
    return insert_zero(insert_zero(p, lo), hi);
}

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

template <typename R>
__device__ __forceinline__ void apply_dense(Cx<R>& a, Cx<R>& b, const R* m)
{
    const Cx<R> na {a.re * m[0] - a.im * m[1] + b.re * m[2] - b.im * m[3],
                    a.re * m[1] + a.im * m[0] + b.re * m[3] + b.im * m[2]};
    const Cx<R> nb {a.re * m[4] - a.im * m[5] + b.re * m[6] - b.im * m[7],
                    a.re * m[5] + a.im * m[4] + b.re * m[7] + b.im * m[6]};
    a = na;
    b = nb;
}

template <typename R>
__device__ __forceinline__ Cx<R> ld(const Amp<R>* p, uint64_t i)
{
    const Amp<R> v = p[i];
    return {v.x, v.y};
}

template <typename R>
__device__ __forceinline__ void st(Amp<R>* p, uint64_t i, Cx<R> v)
{
    Amp<R> w;
    w.x = v.re;
    w.y = v.im;
    p[i] = w;
}

// block-wide sum of a double2 (wave64 shuffles, then LDS across waves); result on thread 0
__device__ double2 block_sum2(double a, double b)
{
    __shared__ double2 red[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[wave] = make_double2(a, b);
    }
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) {
        for (int w = 0; w < BLOCK / 64; ++w) {
            r.x += red[w].x;
            r.y += red[w].y;
        }
    }
    __syncthreads();
    return r;
}

// final fixed-order reduction of block partials: one block, each thread a strided slice
__global__ void k_reduce_partials(const double2* part, int n, double2* out)
{
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        a += part[i].x;
        b += part[i].y;
    }
    const double2 r = block_sum2(a, b);
    if (threadIdx.x == 0) {
        *out = r;
    }
}

// ---- gate sweep: one gate over the whole vector ---------------------------------------------
// Enumeration is in coalesced order (low bits fastest): item p -> the amplitude index with the
// fixed bits (target 0, control 1) inserted.  The set of pairs is exactly the set kettle's
// generators yield (gate_pair_generator.hpp), so per-pair results are identical.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_gate_sweep(Amp<R>* a, uint64_t n_items, DevGate g)
{
    R m[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        m[i] = static_cast<R>(g.m[i]);
    }
    const uint32_t t = g.target;
    const bool ctrl = g.control >= 0;
    const uint32_t c = ctrl ? static_cast<uint32_t>(g.control) : 0;
    const uint32_t lo = ctrl ? min(c, t) : t;
    const uint32_t hi = ctrl ? max(c, t) : t;
    const uint64_t tbit = uint64_t {1} << t;
    const uint64_t cbit = ctrl ? (uint64_t {1} << c) : 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    if (g.kind == GK_DIAG || g.kind == GK_DIAG1) {
        // only amplitudes with control = 1 change; with d0 == 1 only those with target = 1.
        // The host decides (bit 0 of pad) from the FP64 coefficient and sizes n_items to match.
        const bool skip0 = (g.pad & 1u) != 0;
        for (uint64_t p = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; p < n_items; p += stride) {
            uint64_t i;
            if (ctrl && skip0) {
                i = insert_zero2(p, lo, hi) | cbit | tbit;
            }
            else if (ctrl) {
                i = insert_zero(p, c) | cbit;
            }
            else if (skip0) {
                i = insert_zero(p, t) | tbit;
            }
            else {
                i = p;
            }
            Cx<R> v = ld<R>(a, i);
            if (i & tbit) {
                v = cmul(v, m[6], m[7]);
            }
            else {
                v = cmul(v, m[0], m[1]);
            }
            st<R>(a, i, v);
        }
        return;
    }
    for (uint64_t p = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; p < n_items; p += stride) {
        const uint64_t i0 = (ctrl ? insert_zero2(p, lo, hi) : insert_zero(p, t)) | cbit;
        const uint64_t i1 = i0 | tbit;
        Cx<R> x = ld<R>(a, i0);
        Cx<R> y = ld<R>(a, i1);
        if (g.kind == GK_SWAP) {
            st<R>(a, i0, y);
            st<R>(a, i1, x);
        }
        else {
            apply_dense(x, y, m);
            st<R>(a, i0, x);
            st<R>(a, i1, y);
        }
    }
}

// ---- misc state kernels -----------------------------------------------------------------------
template <typename R>
__global__ void k_set_basis(Amp<R>* a, uint64_t n, uint64_t index)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
        Amp<R> v;
        v.x = i == index ? R(1) : R(0);
        v.y = R(0);
        a[i] = v;
    }
}

// SV (dm_n == 0): w_b = sum |a_i|^2 over bit q = b.  DM (dm_n = n): w_b = sum |x|^2 over vec
// indices with row bit q = col bit q = b (measure_density_matrix.cpp:21-51).
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_qubit_weights(const Amp<R>* a, uint64_t n_amps, uint32_t q, uint32_t dm_n,
                                                         double2* part)
{
    double w0 = 0.0, w1 = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n_amps; i += stride) {
        const Amp<R> v = a[i];
        const double p = static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
        const uint32_t br = (i >> q) & 1u;
        if (dm_n) {
            const uint32_t bc = (i >> (q + dm_n)) & 1u;
            if (br != bc) {
                continue;
            }
        }
        if (br) {
            w1 += p;
        }
        else {
            w0 += p;
        }
    }
    const double2 r = block_sum2(w0, w1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r;
    }
}

template <typename R>
__global__ void k_collapse(Amp<R>* a, uint64_t n_amps, uint32_t q, uint32_t dm_n, uint32_t kept, double scale)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n_amps; i += stride) {
        bool keep = ((i >> q) & 1u) == kept;
        if (dm_n) {
            keep = keep && (((i >> (q + dm_n)) & 1u) == kept);
        }
        Amp<R> v = a[i];
        if (keep) {
            v.x = static_cast<R>(v.x * scale);
            v.y = static_cast<R>(v.y * scale);
        }
        else {
            v.x = R(0);
            v.y = R(0);
        }
        a[i] = v;
    }
}

// probabilities: SV |a_i|^2 (std::norm); DM Re rho(i, i)
template <typename R>
__global__ void k_probabilities(const Amp<R>* a, uint64_t n_out, uint32_t dm_n, double* p)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n_out; i += stride) {
        if (dm_n) {
            p[i] = static_cast<double>(a[i + i * n_out].x);
        }
        else {
            const Amp<R> v = a[i];
            const double re = v.x, im = v.y;
            p[i] = re * re + im * im;
        }
    }
}

// apply_noise_ (probabilities.cpp:110-124) for one qubit: pairs are independent
__global__ void k_noise(double* p, uint64_t n_pairs, uint32_t q, double noise)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t k = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; k < n_pairs; k += stride) {
        const uint64_t i0 = insert_zero(k, q);
        const uint64_t i1 = i0 | (uint64_t {1} << q);
        const double p0 = p[i0], p1 = p[i1];
        p[i0] = ((1.0 - noise) * p0) + (noise * p1);
        p[i1] = ((1.0 - noise) * p1) + (noise * p0);
    }
}

// ---- inclusive scan (CDF) -----------------------------------------------------------------------
constexpr int SCAN_PER_THREAD = 8;
constexpr uint64_t SCAN_CHUNK = uint64_t {BLOCK} * SCAN_PER_THREAD;

__device__ double block_exclusive_scan(double v, double* total)
{
    __shared__ double wsum[BLOCK / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(incl, off, 64);
        if (lane >= off) {
            incl += o;
        }
    }
    if (lane == 63) {
        wsum[wave] = incl;
    }
    __syncthreads();
    double wave_off = 0.0, tot = 0.0;
    for (int w = 0; w < BLOCK / 64; ++w) {
        if (w < wave) {
            wave_off += wsum[w];
        }
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return wave_off + incl - v;
}

// Block sums of the scan, and one bit per entry of whether its probability is positive (8 entries per
// thread, one byte of `positive` each).
__global__ __launch_bounds__(BLOCK) void k_scan_block_sums(const double* p, uint64_t n, double* sums, uint8_t* positive)
{
    static_assert(SCAN_PER_THREAD == 8, "one byte of positive-probability bits per thread");
    const uint64_t base = blockIdx.x * SCAN_CHUNK + threadIdx.x * uint64_t {SCAN_PER_THREAD};
    double s = 0.0;
    uint32_t bits = 0;
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        if (base + j < n) {
            const double x = p[base + j];
            s += x;
            bits |= (x > 0.0 ? 1u : 0u) << j;
        }
    }
    if (base < n) {
        positive[base / 8] = static_cast<uint8_t>(bits);
    }
    double total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) {
        sums[blockIdx.x] = total;
    }
}

// exclusive scan of the block sums in one block (each thread a contiguous slice)
__global__ __launch_bounds__(BLOCK) void k_scan_sums(double* sums, uint64_t nb)
{
    const uint64_t per = (nb + BLOCK - 1) / BLOCK;
    const uint64_t b0 = threadIdx.x * per;
    double s = 0.0;
    for (uint64_t j = 0; j < per; ++j) {
        if (b0 + j < nb) {
            s += sums[b0 + j];
        }
    }
    double total;
    double run = block_exclusive_scan(s, &total);
    for (uint64_t j = 0; j < per; ++j) {
        if (b0 + j < nb) {
            const double v = sums[b0 + j];
            sums[b0 + j] = run;
            run += v;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_apply(double* p, uint64_t n, const double* offs)
{
    const uint64_t base = blockIdx.x * SCAN_CHUNK + threadIdx.x * uint64_t {SCAN_PER_THREAD};
    double v[SCAN_PER_THREAD];
    double s = 0.0;
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        v[j] = base + j < n ? p[base + j] : 0.0;
        s += v[j];
    }
    double total;
    double run = block_exclusive_scan(s, &total) + offs[blockIdx.x];
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        run += v[j];
        if (base + j < n) {
            p[base + j] = run;
        }
    }
}

// largest index j with cdf[j] < last (cumulative_end_offset_, measurements.cpp:235-256)
__global__ __launch_bounds__(BLOCK) void k_last_below(const double* cdf, uint64_t n, double last, long long* out)
{
    __shared__ long long best[BLOCK];
    long long b = -1;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i + 1 < n; i += stride) {
        if (cdf[i] < last) {
            b = static_cast<long long>(i);
        }
    }
    best[threadIdx.x] = b;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            best[threadIdx.x] = max(best[threadIdx.x], best[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(out, best[0]);
    }
}

// std::ranges::lower_bound per shot (measurements.cpp:297-312).  The device CDF sums in a tree order, so
// at a thread or block boundary it may differ from a sequential partial sum by an ulp, in either
// direction; a draw in that sliver may land on an entry of probability zero, which a sequential CDF
// never returns.  Such a draw belongs to the next entry of positive probability (the previous one when
// only zeros follow): the search steps there using the positive-probability bits of the scan.
__device__ uint64_t next_positive(const uint8_t* positive, uint64_t n, uint64_t i)
{
    for (uint64_t j = i; j < n; ++j) {
        if ((positive[j >> 3] >> (j & 7)) & 1u) {
            return j;
        }
    }
    for (uint64_t j = i; j-- > 0;) {
        if ((positive[j >> 3] >> (j & 7)) & 1u) {
            return j;
        }
    }
    return i;
}

__global__ void k_lower_bound(const double* cdf, uint64_t n, const double* u, uint64_t shots, const uint8_t* positive,
                              uint64_t* out)
{
    const uint64_t s = blockIdx.x * uint64_t {BLOCK} + threadIdx.x;
    if (s >= shots) {
        return;
    }
    const double x = u[s];
    uint64_t lo = 0, len = n;
    while (len > 0) {
        const uint64_t half = len >> 1;
        if (cdf[lo + half] < x) {
            lo = lo + half + 1;
            len = len - half - 1;
        }
        else {
            len = half;
        }
    }
    if (lo < n && !((positive[lo >> 3] >> (lo & 7)) & 1u)) {
        lo = next_positive(positive, n, lo);
    }
    out[s] = lo;
}

// ---- Pauli strings ------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void pauli_coef(uint64_t i, const PauliGroup& g, const PauliZ* zs, double& cr, double& ci)
{
    cr = 0.0;
    ci = 0.0;
    for (uint32_t s = 0; s < g.count; ++s) {
        const PauliZ z = zs[g.first + s];
        const double sg = (__popcll(i & z.z) & 1) ? -1.0 : 1.0;
        cr += sg * z.w[0];
        ci += sg * z.w[1];
    }
}

// (H a)_i = sum over groups (in order) of c_g(i) a[i ^ x_g], for one index (batched kernels).
template <typename R>
__device__ __forceinline__ void h_apply_row(const Amp<R>* a, uint64_t i, const PauliGroup* groups, uint32_t ng,
                                            const PauliZ* zs, double& hr, double& hi)
{
    hr = 0.0;
    hi = 0.0;
    for (uint32_t gi = 0; gi < ng; ++gi) {
        const PauliGroup g = groups[gi];
        double cr, ci;
        pauli_coef<R>(i, g, zs, cr, ci);
        const Amp<R> u = a[i ^ g.x];
        const double br = u.x, bi = u.y;
        hr += cr * br - ci * bi;
        hi += cr * bi + ci * br;
    }
}

// (H a)_i for the J indices i0 + j * T of one thread (T: threads of the grid), in the same order per index
// as h_apply_row: per group the J gathers go out together and each string's sign mask is read once for
// all J indices.  NARROW: every index and mask fits 32 bits (32-bit popcounts).
constexpr int PAULI_ROWS = 4;
template <typename R, bool NARROW>
__device__ __forceinline__ void h_apply_rows(const Amp<R>* a, const uint64_t (&idx)[PAULI_ROWS], const bool (&in)[PAULI_ROWS],
                                             const PauliGroup* groups, uint32_t ng, const PauliZ* zs,
                                             double (&hr)[PAULI_ROWS], double (&hi)[PAULI_ROWS])
{
#pragma unroll
    for (int j = 0; j < PAULI_ROWS; ++j) {
        hr[j] = 0.0;
        hi[j] = 0.0;
    }
    for (uint32_t gi = 0; gi < ng; ++gi) {
        const PauliGroup g = groups[gi];
        Amp<R> u[PAULI_ROWS];
#pragma unroll
        for (int j = 0; j < PAULI_ROWS; ++j) {
            u[j] = in[j] ? a[idx[j] ^ g.x] : Amp<R> {};
        }
        double cr[PAULI_ROWS], ci[PAULI_ROWS];
#pragma unroll
        for (int j = 0; j < PAULI_ROWS; ++j) {
            cr[j] = 0.0;
            ci[j] = 0.0;
        }
        for (uint32_t s = 0; s < g.count; ++s) {
            const PauliZ z = zs[g.first + s];
#pragma unroll
            for (int j = 0; j < PAULI_ROWS; ++j) {
                const uint32_t par = NARROW ? (__popc(static_cast<uint32_t>(idx[j]) & static_cast<uint32_t>(z.z)) & 1u)
                                            : (static_cast<uint32_t>(__popcll(idx[j] & z.z)) & 1u);
                const double sg = par ? -1.0 : 1.0;
                cr[j] += sg * z.w[0];
                ci[j] += sg * z.w[1];
            }
        }
#pragma unroll
        for (int j = 0; j < PAULI_ROWS; ++j) {
            const double br = u[j].x, bi = u[j].y;
            hr[j] += cr[j] * br - ci[j] * bi;
            hi[j] += cr[j] * bi + ci[j] * br;
        }
    }
}

template <typename R, bool NARROW>
__global__ __launch_bounds__(BLOCK) void k_pauli_expect(const Amp<R>* a, uint64_t n, const PauliGroup* groups, int ng,
                                                        const PauliZ* zs, double2* part)
{
    double er = 0.0, ei = 0.0;
    const uint64_t T = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i0 = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i0 < n; i0 += PAULI_ROWS * T) {
        uint64_t idx[PAULI_ROWS];
        bool in[PAULI_ROWS];
#pragma unroll
        for (int j = 0; j < PAULI_ROWS; ++j) {
            idx[j] = i0 + j * T;
            in[j] = idx[j] < n;
        }
        double hr[PAULI_ROWS], hi[PAULI_ROWS];
        h_apply_rows<R, NARROW>(a, idx, in, groups, static_cast<uint32_t>(ng), zs, hr, hi);
#pragma unroll
        for (int j = 0; j < PAULI_ROWS; ++j) {
            if (in[j]) {
                // conj(a) * (H a)_i
                const Amp<R> v = a[idx[j]];
                const double ar = v.x, ai = v.y;
                er += ar * hr[j] + ai * hi[j];
                ei += ar * hi[j] - ai * hr[j];
            }
        }
    }
    const double2 r = block_sum2(er, ei);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r;
    }
}

template <typename R, bool NARROW>
__global__ __launch_bounds__(BLOCK) void k_pauli_apply_sum(const Amp<R>* a, Amp<R>* b, uint64_t n, const PauliGroup* groups,
                                                           int ng, const PauliZ* zs)
{
    const uint64_t T = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i0 = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i0 < n; i0 += PAULI_ROWS * T) {
        uint64_t idx[PAULI_ROWS];
        bool in[PAULI_ROWS];
#pragma unroll
        for (int j = 0; j < PAULI_ROWS; ++j) {
            idx[j] = i0 + j * T;
            in[j] = idx[j] < n;
        }
        double hr[PAULI_ROWS], hi[PAULI_ROWS];
        h_apply_rows<R, NARROW>(a, idx, in, groups, static_cast<uint32_t>(ng), zs, hr, hi);
#pragma unroll
        for (int j = 0; j < PAULI_ROWS; ++j) {
            if (in[j]) {
                Amp<R> w;
                w.x = static_cast<R>(hr[j]);
                w.y = static_cast<R>(hi[j]);
                b[idx[j]] = w;
            }
        }
    }
}

// b = H a over tiles of 2^PAULI_TILE_BITS amplitudes, one workgroup per output tile t: a group with flip
// mask x reads tile t ^ (x >> PAULI_TILE_BITS) at offsets i ^ (x & (TILE - 1)).  The partner tile is
// staged in LDS and kept for the following groups while their high flip bits agree, so groups sorted
// by x >> PAULI_TILE_BITS (sort_pauli_groups_by_tile) read each distinct partner tile once per output
// tile instead of the state once per group.  The sign parity of a string splits into the tile's bits
// (uniform) and the lane's offset bits.  Needs a != b and n = n_tiles * TILE.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_pauli_apply_tiled(const Amp<R>* __restrict__ a, Amp<R>* __restrict__ b,
                                                             uint64_t n_tiles, const PauliGroup* groups, int ng,
                                                             const PauliZ* zs)
{
    constexpr uint32_t TILE = 1u << PAULI_TILE_BITS;
    constexpr int PER = TILE / BLOCK;
    __shared__ Amp<R> part[TILE];
    for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        double hr[PER], hi[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            hr[k] = 0.0;
            hi[k] = 0.0;
        }
        uint64_t staged = ~uint64_t {0};
        for (int gi = 0; gi < ng; ++gi) {
            const PauliGroup g = groups[gi];
            const uint64_t hx = g.x >> PAULI_TILE_BITS;
            if (hx != staged) {   // uniform over the workgroup
                __syncthreads();
                const Amp<R>* src = a + ((t ^ hx) << PAULI_TILE_BITS);
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    part[threadIdx.x + k * BLOCK] = src[threadIdx.x + k * BLOCK];
                }
                __syncthreads();
                staged = hx;
            }
            const uint32_t xl = static_cast<uint32_t>(g.x) & (TILE - 1);
            double cr[PER], ci[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                cr[k] = 0.0;
                ci[k] = 0.0;
            }
            for (uint32_t s = 0; s < g.count; ++s) {
                const PauliZ z = zs[g.first + s];
                const uint32_t tpar = static_cast<uint32_t>(__popcll(t & (z.z >> PAULI_TILE_BITS))) & 1u;
                const uint32_t zl = static_cast<uint32_t>(z.z) & (TILE - 1);
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const uint32_t il = threadIdx.x + k * BLOCK;
                    const double sg = ((__popc(il & zl) & 1u) ^ tpar) ? -1.0 : 1.0;
                    cr[k] += sg * z.w[0];
                    ci[k] += sg * z.w[1];
                }
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const Amp<R> u = part[(threadIdx.x + k * BLOCK) ^ xl];
                const double br = u.x, bi = u.y;
                hr[k] += cr[k] * br - ci[k] * bi;
                hi[k] += cr[k] * bi + ci[k] * br;
            }
        }
        Amp<R>* dst = b + (t << PAULI_TILE_BITS);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            Amp<R> w;
            w.x = static_cast<R>(hr[k]);
            w.y = static_cast<R>(hi[k]);
            dst[threadIdx.x + k * BLOCK] = w;
        }
    }
}



// in place a_i <- w (-1)^{|i & z|} a_{i ^ x}; for x != 0 each pair (i, i ^ x) is handled by the
// index whose highest bit of x is 0
template <typename R>
__global__ void k_pauli_apply_one(Amp<R>* a, uint64_t n, uint64_t x, uint64_t z, double wr, double wi)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    if (x == 0) {
        for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
            const double sg = (__popcll(i & z) & 1) ? -1.0 : 1.0;
            const Amp<R> v = a[i];
            const double vr = v.x, vi = v.y;
            Amp<R> w;
            w.x = static_cast<R>(sg * (wr * vr - wi * vi));
            w.y = static_cast<R>(sg * (wr * vi + wi * vr));
            a[i] = w;
        }
        return;
    }
    const uint32_t h = 63 - __clzll(x);
    for (uint64_t p = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; p < n / 2; p += stride) {
        const uint64_t i = insert_zero(p, h);
        const uint64_t j = i ^ x;
        const Amp<R> vi_ = a[i], vj_ = a[j];
        const double si = (__popcll(i & z) & 1) ? -1.0 : 1.0;
        const double sj = (__popcll(j & z) & 1) ? -1.0 : 1.0;
        Amp<R> ni, nj;
        ni.x = static_cast<R>(si * (wr * vj_.x - wi * vj_.y));
        ni.y = static_cast<R>(si * (wr * vj_.y + wi * vj_.x));
        nj.x = static_cast<R>(sj * (wr * vi_.x - wi * vi_.y));
        nj.y = static_cast<R>(sj * (wr * vi_.y + wi * vi_.x));
        a[i] = ni;
        a[j] = nj;
    }
}

// ---- density-matrix read-outs ----------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_dm_trace(const Amp<R>* a, uint64_t N, double2* part)
{
    double tr = 0.0, ti = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < N; i += stride) {
        const Amp<R> v = a[i + i * N];
        tr += v.x;
        ti += v.y;
    }
    const double2 r = block_sum2(tr, ti);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r;
    }
}

// trace_of_square (density_matrix.cpp:164-168): tr(rho^2) = sum_{r,c} rho(r,c) rho(c,r) in FP64, over pairs
// (I <= J) of T x T blocks of rho; a pair and its mirror (J, I) give the same products, so an
// off-diagonal pair counts twice.  A block's rho(c, r) tile is staged through LDS (transposed) so both
// tiles are read in coalesced columns.  Each workgroup walks the pairs with a grid stride in a fixed
// order and writes one partial; the partials are reduced in a fixed order (bit-reproducible).
constexpr uint32_t TSQ_TILE = 32;
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_dm_trace_sq(const Amp<R>* a, uint64_t N, uint32_t T, double2* part)
{
    __shared__ double2 bt[TSQ_TILE * (TSQ_TILE + 1)];
    const uint64_t nt = N / T;
    double sr = 0.0, si = 0.0;
    for (uint64_t q = blockIdx.x; q < nt * nt; q += gridDim.x) {
        const uint64_t I = q % nt, J = q / nt;
        if (I > J) {
            continue;   // uniform over the workgroup
        }
        for (uint32_t e = threadIdx.x; e < T * T; e += BLOCK) {
            const uint32_t cl = e % T, rl = e / T;
            const Amp<R> v = a[(J * T + cl) + (I * T + rl) * N];   // rho(c, r), c fastest
            bt[rl * (TSQ_TILE + 1) + cl] = make_double2(static_cast<double>(v.x), static_cast<double>(v.y));
        }
        __syncthreads();
        const double w = I == J ? 1.0 : 2.0;
        for (uint32_t e = threadIdx.x; e < T * T; e += BLOCK) {
            const uint32_t rl = e % T, cl = e / T;
            const Amp<R> x = a[(I * T + rl) + (J * T + cl) * N];   // rho(r, c), r fastest
            const double2 y = bt[rl * (TSQ_TILE + 1) + cl];
            const double xr = static_cast<double>(x.x), xi = static_cast<double>(x.y);
            sr += w * (xr * y.x - xi * y.y);
            si += w * (xr * y.y + xi * y.x);
        }
        __syncthreads();
    }
    const double2 r = block_sum2(sr, si);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r;
    }
}

// partial_trace (density_matrix.cpp:231-278): out(ir, ic) = sum_e rho(ir (+) e, ic (+) e) over
// assignments e of the traced qubits.  The sum runs as a binary counter over e with bit 0 of e on
// the highest traced qubit, which is the association kettle's one-qubit-at-a-time loop produces.
template <typename R>
__global__ void k_dm_partial_trace(const Amp<R>* a, uint32_t n, uint64_t kept_mask, uint64_t traced_desc_packed_lo,
                                   uint64_t traced_desc_packed_hi, uint32_t n_traced, uint64_t n_out_side, double2* out)
{
    const uint64_t idx = blockIdx.x * uint64_t {BLOCK} + threadIdx.x;
    if (idx >= n_out_side * n_out_side) {
        return;
    }
    const uint64_t ir = idx % n_out_side;
    const uint64_t ic = idx / n_out_side;
    const uint64_t N = uint64_t {1} << n;
    // deposit ir / ic into the kept positions (ascending)
    uint64_t r0 = 0, c0 = 0;
    {
        uint32_t j = 0;
        for (uint32_t q = 0; q < n; ++q) {
            if ((kept_mask >> q) & 1ull) {
                r0 |= ((ir >> j) & 1ull) << q;
                c0 |= ((ic >> j) & 1ull) << q;
                ++j;
            }
        }
    }
    uint8_t tq[32];
    for (uint32_t j = 0; j < n_traced; ++j) {
        const uint64_t word = j < 8 ? traced_desc_packed_lo : traced_desc_packed_hi;
        tq[j] = static_cast<uint8_t>((word >> (8 * (j & 7))) & 0xFF);
    }
    double sr[33], si[33];
    int top = 0;
    const uint64_t ne = uint64_t {1} << n_traced;
    for (uint64_t e = 0; e < ne; ++e) {
        uint64_t off = 0;
        for (uint32_t j = 0; j < n_traced; ++j) {
            off |= ((e >> j) & 1ull) << tq[j];
        }
        const Amp<R> v = a[(r0 | off) + (c0 | off) * N];
        double xr = v.x, xi = v.y;
        uint64_t cnt = e;
        while (cnt & 1ull) {
            --top;
            xr = sr[top] + xr;
            xi = si[top] + xi;
            cnt >>= 1;
        }
        sr[top] = xr;
        si[top] = xi;
        ++top;
    }
    out[idx] = make_double2(sr[0], si[0]);
}

template <typename R>
__global__ void k_dm_from_sv(Amp<R>* rho, const Amp<R>* psi, uint64_t N)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < N * N; i += stride) {
        const uint64_t r = i % N, c = i / N;
        const Amp<R> x = psi[r], y = psi[c];
        Amp<R> w;
        w.x = x.x * y.x + x.y * y.y;
        w.y = x.y * y.x - x.x * y.y;
        rho[i] = w;
    }
}

// ---- kettle generator order, for the index-map parity checks ---------------------------------------
__global__ void k_pair_map_1q(uint32_t n, uint32_t t, uint64_t* s0, uint64_t* s1)
{
    const uint64_t np = uint64_t {1} << (n - 1);
    const uint64_t i0_max = uint64_t {1} << t;
    const uint64_t i1_max = uint64_t {1} << (n - t - 1);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t k = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; k < np; k += stride) {
        const uint64_t i0 = k / i1_max;
        const uint64_t i1 = k - i0 * i1_max;
        const uint64_t a = i0 + 2 * i1 * i0_max;
        s0[k] = a;
        s1[k] = a + i0_max;
    }
}

__global__ void k_pair_map_2q(uint32_t n, uint32_t c, uint32_t t, int four, uint64_t* out)
{
    const uint32_t lo = min(c, t), hi = max(c, t);
    const uint64_t np = n < 2 ? 0 : (uint64_t {1} << (n - 2));
    const uint64_t i1_max = uint64_t {1} << (hi - lo - 1);
    const uint64_t i2_max = uint64_t {1} << (n - hi - 1);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t k = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; k < np; k += stride) {
        const uint64_t i0 = k / (i1_max * i2_max);
        const uint64_t r = k - i0 * i1_max * i2_max;
        const uint64_t i1 = r / i2_max;
        const uint64_t i2 = r - i1 * i2_max;
        const uint64_t c0t0 = i0 + i1 * (uint64_t {1} << (lo + 1)) + i2 * (uint64_t {1} << (hi + 1));
        const uint64_t tb = uint64_t {1} << t, cb = uint64_t {1} << c;
        if (four) {
            out[4 * k] = c0t0;
            out[4 * k + 1] = c0t0 + tb;
            out[4 * k + 2] = c0t0 + cb;
            out[4 * k + 3] = c0t0 + tb + cb;
        }
        else {
            out[2 * k] = c0t0 + cb;
            out[2 * k + 1] = c0t0 + cb + tb;
        }
    }
}

// ---- adjoint: one gate on (psi, lambda) plus the derivative overlap ------------------------------
// psi <- inv psi; acc += <lam | d psi>; lam <- inv lam   (inv = G^dagger, d = dG/dtheta)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_adjoint_gate(Amp<R>* psi, Amp<R>* lam, uint64_t n_items, DevGate inv,
                                                        DevGate d, int has_d, double2* part)
{
    R m[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        m[i] = static_cast<R>(inv.m[i]);
    }
    const uint32_t t = inv.target;
    const bool ctrl = inv.control >= 0;
    const uint32_t c = ctrl ? static_cast<uint32_t>(inv.control) : 0;
    const uint32_t lo = ctrl ? min(c, t) : t;
    const uint32_t hi = ctrl ? max(c, t) : t;
    const uint64_t tbit = uint64_t {1} << t;
    const uint64_t cbit = ctrl ? (uint64_t {1} << c) : 0;
    double accr = 0.0, acci = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t p = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; p < n_items; p += stride) {
        const uint64_t i0 = (ctrl ? insert_zero2(p, lo, hi) : insert_zero(p, t)) | cbit;
        const uint64_t i1 = i0 | tbit;
        Cx<R> x = ld<R>(psi, i0), y = ld<R>(psi, i1);
        Cx<R> u = ld<R>(lam, i0), v = ld<R>(lam, i1);
        if (has_d) {
            // Re <lambda| Q |psi> before the gate is undone (dG = Q G), in FP64
            const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
            const double m0r = xr * d.m[0] - xi * d.m[1] + yr * d.m[2] - yi * d.m[3];
            const double m0i = xr * d.m[1] + xi * d.m[0] + yr * d.m[3] + yi * d.m[2];
            const double m1r = xr * d.m[4] - xi * d.m[5] + yr * d.m[6] - yi * d.m[7];
            const double m1i = xr * d.m[5] + xi * d.m[4] + yr * d.m[7] + yi * d.m[6];
            const double ur = u.re, ui = u.im, vr = v.re, vi = v.im;
            accr += ur * m0r + ui * m0i + vr * m1r + vi * m1i;
            acci += ur * m0i - ui * m0r + vr * m1i - vi * m1r;
        }
        if (inv.kind == GK_SWAP) {
            const Cx<R> tmp = x;
            x = y;
            y = tmp;
        }
        else if (inv.kind == GK_DIAG || inv.kind == GK_DIAG1) {
            x = cmul(x, m[0], m[1]);
            y = cmul(y, m[6], m[7]);
        }
        else {
            apply_dense(x, y, m);
        }
        st<R>(psi, i0, x);
        st<R>(psi, i1, y);
        if (inv.kind == GK_SWAP) {
            const Cx<R> tmp = u;
            u = v;
            v = tmp;
        }
        else if (inv.kind == GK_DIAG || inv.kind == GK_DIAG1) {
            u = cmul(u, m[0], m[1]);
            v = cmul(v, m[6], m[7]);
        }
        else {
            apply_dense(u, v, m);
        }
        st<R>(lam, i0, u);
        st<R>(lam, i1, v);
    }
    if (has_d) {
        const double2 r = block_sum2(accr, acci);
        if (threadIdx.x == 0) {
            part[blockIdx.x] = r;
        }
    }
}

// plain copy of n amplitudes (also the calibration kernel for the HBM traffic counters)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_copy(Amp<R>* dst, const Amp<R>* src, uint64_t n)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
        dst[i] = src[i];
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_dot(const Amp<R>* a, const Amp<R>* b, uint64_t n, double2* part)
{
    double sr = 0.0, si = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
        const Amp<R> x = a[i], y = b[i];
        const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
        sr += xr * yr + xi * yi;
        si += xr * yi - xi * yr;
    }
    const double2 r = block_sum2(sr, si);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r;
    }
}

// ---- batched states -------------------------------------------------------------------------------
// lambda = H psi for many states in one launch, and per-block partials of Re <psi|lambda>.  Item i
// covers blocks [block_begin, next block_begin), BATCH_AMPS amplitudes per block.
constexpr uint32_t BATCH_AMPS = 4 * BLOCK;

template <typename R, typename RO>
__global__ __launch_bounds__(BLOCK) void k_pauli_batch(const Amp<R>* __restrict__ arena_a, Amp<RO>* __restrict__ arena_b,
                                                       const BatchPauliItem* __restrict__ items, uint32_t n_items,
                                                       double* __restrict__ part)
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
    const BatchPauliItem& it = items[lo];
    const uint32_t local = blockIdx.x - it.block_begin;
    const Amp<R>* a = arena_a + it.a_off;
    Amp<RO>* b = arena_b + it.b_off;
    using GGroup = const __attribute__((address_space(1))) PauliGroup;
    using GZ = const __attribute__((address_space(1))) PauliZ;
    const auto* groups = (const PauliGroup*) (GGroup*) it.groups;
    const auto* zs = (const PauliZ*) (GZ*) it.zs;
    // the block's BATCH_AMPS = PAULI_ROWS * BLOCK rows: thread t takes rows t + j * BLOCK, their gathers
    // issued together per group
    static_assert(BATCH_AMPS == PAULI_ROWS * BLOCK, "one row set per thread and block");
    uint64_t idx[PAULI_ROWS];
    bool in[PAULI_ROWS];
#pragma unroll
    for (int j = 0; j < PAULI_ROWS; ++j) {
        idx[j] = static_cast<uint64_t>(local) * BATCH_AMPS + threadIdx.x + static_cast<uint64_t>(j) * BLOCK;
        in[j] = idx[j] < it.n_amps;
    }
    double hr[PAULI_ROWS], hi_[PAULI_ROWS];
    if (it.n_amps <= (uint64_t {1} << 32)) {
        h_apply_rows<R, true>(a, idx, in, groups, it.n_groups, zs, hr, hi_);
    }
    else {
        h_apply_rows<R, false>(a, idx, in, groups, it.n_groups, zs, hr, hi_);
    }
    double e = 0.0;
#pragma unroll
    for (int j = 0; j < PAULI_ROWS; ++j) {
        if (in[j]) {
            const Amp<R> v = a[idx[j]];
            e += static_cast<double>(v.x) * hr[j] + static_cast<double>(v.y) * hi_[j];
            Amp<RO> w;
            w.x = static_cast<RO>(hr[j]);
            w.y = static_cast<RO>(hi_[j]);
            b[idx[j]] = w;
        }
    }
    const double2 r = block_sum2(e, 0.0);
    if (threadIdx.x == 0) {
        part[it.part_off + local] = r.x;
    }
}

// complex128 -> complex64 copy of a whole arena
__global__ __launch_bounds__(BLOCK) void k_narrow(const double2* __restrict__ in, float2* __restrict__ out, uint64_t n)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
        const double2 v = in[i];
        out[i] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_fill_zero(Amp<R>* __restrict__ a, uint64_t n)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
        Amp<R> z;
        z.x = 0;
        z.y = 0;
        a[i] = z;
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_set_ones(Amp<R>* __restrict__ a, const uint64_t* __restrict__ offs, uint32_t n)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) {
        Amp<R> one;
        one.x = 1;
        one.y = 0;
        a[offs[i]] = one;
    }
}

// max over |a_i - delta_{i,0}| per state: the distance of each state from |0...0>, per block
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_basis_distance_batch(const Amp<R>* __restrict__ arena,
                                                                const BatchPauliItem* __restrict__ items, uint32_t n_items,
                                                                double* __restrict__ part)
{
    __shared__ double red[BLOCK];
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
    const BatchPauliItem& it = items[lo];
    const uint32_t local = blockIdx.x - it.block_begin;
    const Amp<R>* a = arena + it.a_off;
    const uint64_t end = min(it.n_amps, static_cast<uint64_t>(local + 1) * BATCH_AMPS);
    double m = 0.0;
    for (uint64_t i = static_cast<uint64_t>(local) * BATCH_AMPS + threadIdx.x; i < end; i += BLOCK) {
        const Amp<R> v = a[i];
        const double re = static_cast<double>(v.x) - (i == 0 ? 1.0 : 0.0);
        m = fmax(m, sqrt(re * re + static_cast<double>(v.y) * static_cast<double>(v.y)));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[it.part_off + local] = red[0];
    }
}

// out[i] = max of rows[offs[i] .. offs[i] + counts[i])
__global__ __launch_bounds__(BLOCK) void k_max_rows(const double* __restrict__ rows, const uint64_t* __restrict__ offs,
                                                    const uint32_t* __restrict__ counts, double* __restrict__ out)
{
    __shared__ double red[BLOCK];
    const uint32_t row = blockIdx.x;
    const double* p = rows + offs[row];
    const uint32_t n = counts[row];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        acc = fmax(acc, p[i]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[row] = red[0];
    }
}

// out[i] = sum of rows[offs[i] .. offs[i] + counts[i]) (one block per row, fixed summation order)
__global__ __launch_bounds__(BLOCK) void k_sum_rows(const double* __restrict__ rows, const uint64_t* __restrict__ offs,
                                                    const uint32_t* __restrict__ counts, double* __restrict__ out)
{
    __shared__ double red[BLOCK];
    const uint32_t row = blockIdx.x;
    const double* p = rows + offs[row];
    const uint32_t n = counts[row];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        acc += p[i];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            red[threadIdx.x] += red[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[row] = red[0];
    }
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_max_abs_diff(const Amp<R>* a, const Amp<R>* b, uint64_t n, double2* part)
{
    double mx = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
        const Amp<R> x = a[i], y = b[i];
        const double dr = static_cast<double>(x.x) - y.x, di = static_cast<double>(x.y) - y.y;
        mx = fmax(mx, sqrt(dr * dr + di * di));
    }
    __shared__ double red[BLOCK];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = make_double2(red[0], 0.0);
    }
}

// ---- launch helpers ----------------------------------------------------------------------------------

unsigned grid_for(uint64_t items, unsigned per_thread = 1, unsigned cap = 1u << 20)
{
    const uint64_t threads = (items + per_thread - 1) / per_thread;
    uint64_t blocks = (threads + BLOCK - 1) / BLOCK;
    blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, cap));
    return static_cast<unsigned>(blocks);
}

DevGate to_dev(const GateRec& g)
{
    DevGate d {};
    d.kind = g.kind;
    d.target = g.target;
    d.control = g.control;
    d.pad = g.pad;
    std::memcpy(d.m, g.m, sizeof d.m);
    return d;
}

// two-stage deterministic reduction into a host double2
template <typename Launch>
void reduce2(kq_state* s, unsigned blocks, Launch&& launch, double out[2])
{
    auto* part = static_cast<double2*>(scratch(s, (RED_BLOCKS + 1) * sizeof(double2)));
    launch(part);
    KQ_HIP_CHECK(hipGetLastError());
    k_reduce_partials<<<1, BLOCK, 0, s->stream>>>(part, static_cast<int>(blocks), part + RED_BLOCKS);
    KQ_HIP_CHECK(hipGetLastError());
    double2 h;
    KQ_HIP_CHECK(hipMemcpyAsync(&h, part + RED_BLOCKS, sizeof h, hipMemcpyDeviceToHost, s->stream));
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    out[0] = h.x;
    out[1] = h.y;
}

}  // namespace

#define KQ_DISPATCH(s, KERNEL_CALL_DOUBLE, KERNEL_CALL_FLOAT) \
    do {                                                      \
        if ((s)->precision == KQ_C64) {                       \
            KERNEL_CALL_FLOAT;                                \
        }                                                     \
        else {                                                \
            KERNEL_CALL_DOUBLE;                               \
        }                                                     \
        KQ_HIP_CHECK(hipGetLastError());                      \
    } while (0)

void launch_gate_sweep(kq_state* s, const GateRec& g)
{
    DevGate d = to_dev(g);
    const bool ctrl = g.control >= 0;
    uint64_t items;
    if (g.kind == GK_DIAG || g.kind == GK_DIAG1) {
        const bool skip0 = g.m[0] == 1.0 && g.m[1] == 0.0;
        d.pad = skip0 ? 1u : 0u;
        items = s->n_amps >> ((ctrl ? 1 : 0) + (skip0 ? 1 : 0));
    }
    else {
        items = s->n_amps >> (ctrl ? 2 : 1);
    }
    if (items == 0) {
        return;
    }
    const unsigned blocks = grid_for(items, 4);
    KQ_DISPATCH(s,
                (k_gate_sweep<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<double2*>(s->amps), items, d)),
                (k_gate_sweep<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<float2*>(s->amps), items, d)));
}

void launch_set_basis(kq_state* s, uint64_t index)
{
    s->lazy_basis = -1;   // written now: no pending basis state may overwrite it later
    const unsigned blocks = grid_for(s->n_amps, 4);
    KQ_DISPATCH(s, (k_set_basis<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<double2*>(s->amps), s->n_amps, index)),
                (k_set_basis<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<float2*>(s->amps), s->n_amps, index)));
}

void launch_qubit_weights(kq_state* s, uint32_t q, double out[2])
{
    const unsigned blocks = grid_for(s->n_amps, 8, RED_BLOCKS);
    const uint32_t dm_n = s->is_dm ? s->n : 0;
    reduce2(s, blocks, [&](double2* part) {
        KQ_DISPATCH(s,
                    (k_qubit_weights<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(s->amps), s->n_amps, q, dm_n, part)),
                    (k_qubit_weights<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(s->amps), s->n_amps, q, dm_n, part)));
    }, out);
}

void launch_collapse(kq_state* s, uint32_t q, int kept, double scale)
{
    const unsigned blocks = grid_for(s->n_amps, 4);
    const uint32_t dm_n = s->is_dm ? s->n : 0;
    KQ_DISPATCH(s,
                (k_collapse<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<double2*>(s->amps), s->n_amps, q, dm_n, static_cast<uint32_t>(kept), scale)),
                (k_collapse<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<float2*>(s->amps), s->n_amps, q, dm_n, static_cast<uint32_t>(kept), scale)));
}

void launch_probabilities(kq_state* s, double* d_probs, const double* noise_host)
{
    const uint64_t n_out = uint64_t {1} << s->n;
    const unsigned blocks = grid_for(n_out, 4);
    const uint32_t dm_n = s->is_dm ? s->n : 0;
    KQ_DISPATCH(s,
                (k_probabilities<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(s->amps), n_out, dm_n, d_probs)),
                (k_probabilities<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(s->amps), n_out, dm_n, d_probs)));
    if (noise_host != nullptr) {
        for (uint32_t q = 0; q < s->n; ++q) {
            launch_noise(s->stream, d_probs, s->n, q, noise_host[q]);
        }
    }
}

void launch_noise(hipStream_t st, double* d_probs, uint32_t n, uint32_t q, double noise)
{
    const uint64_t np = uint64_t {1} << (n - 1);
    k_noise<<<grid_for(np, 4), BLOCK, 0, st>>>(d_probs, np, q, noise);
    KQ_HIP_CHECK(hipGetLastError());
}

// scratch of the CDF: block sums, the index of the last entry below the total, positive-probability bits
size_t cdf_scratch_bytes(uint64_t n)
{
    const uint64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    return nb * sizeof(double) + 64 + (n + 7) / 8;
}

namespace
{
uint8_t* cdf_positive_bits(void* scratch_ptr, uint64_t n)
{
    const uint64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    return static_cast<uint8_t*>(scratch_ptr) + nb * sizeof(double) + 64;
}
}  // namespace

void launch_cdf(hipStream_t st, double* d_cdf, uint64_t n, void* scratch_ptr, size_t scratch_bytes, double* last,
                double* second_last)
{
    const uint64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (scratch_bytes < cdf_scratch_bytes(n)) {
        throw InvalidArg {"cdf: scratch too small"};
    }
    auto* sums = static_cast<double*>(scratch_ptr);
    auto* best = reinterpret_cast<long long*>(static_cast<char*>(scratch_ptr) + nb * sizeof(double));
    k_scan_block_sums<<<static_cast<unsigned>(nb), BLOCK, 0, st>>>(d_cdf, n, sums, cdf_positive_bits(scratch_ptr, n));
    KQ_HIP_CHECK(hipGetLastError());
    k_scan_sums<<<1, BLOCK, 0, st>>>(sums, nb);
    KQ_HIP_CHECK(hipGetLastError());
    k_scan_apply<<<static_cast<unsigned>(nb), BLOCK, 0, st>>>(d_cdf, n, sums);
    KQ_HIP_CHECK(hipGetLastError());
    KQ_HIP_CHECK(hipMemcpyAsync(last, d_cdf + (n - 1), sizeof(double), hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    const long long init = -1;
    KQ_HIP_CHECK(hipMemcpyAsync(best, &init, sizeof init, hipMemcpyHostToDevice, st));
    k_last_below<<<grid_for(n, 8, 4096), BLOCK, 0, st>>>(d_cdf, n, *last, best);
    KQ_HIP_CHECK(hipGetLastError());
    long long j = -1;
    KQ_HIP_CHECK(hipMemcpyAsync(&j, best, sizeof j, hipMemcpyDeviceToHost, st));
    KQ_HIP_CHECK(hipStreamSynchronize(st));
    *second_last = 0.0;
    if (j >= 0) {
        KQ_HIP_CHECK(hipMemcpyAsync(second_last, d_cdf + j, sizeof(double), hipMemcpyDeviceToHost, st));
        KQ_HIP_CHECK(hipStreamSynchronize(st));
    }
}

static_assert((NOALLOC_MAX_PROBS + SCAN_CHUNK - 1) / SCAN_CHUNK * sizeof(double) + 64 + NOALLOC_MAX_PROBS / 8
                  <= NOALLOC_CDF_BYTES,
              "the allocation-free sampler's scan partials must fit their share of the workspace");

NoallocWorkspace noalloc_workspace()
{
    int dev = 0;
    KQ_HIP_CHECK(hipGetDevice(&dev));
    static std::mutex mu;
    static std::map<int, NoallocWorkspace> per_device;
    std::lock_guard<std::mutex> lock {mu};
    const auto it = per_device.find(dev);
    if (it != per_device.end()) {
        return it->second;
    }
    NoallocWorkspace ws {nullptr, nullptr};
    KQ_HIP_CHECK(hipGetSymbolAddress(&ws.ptr, HIP_SYMBOL(::scratch)));
    if (ws.ptr == nullptr) {
        throw HipError {"hipGetSymbolAddress returned no address for the sampler workspace"};
    }
    KQ_HIP_CHECK(hipStreamCreateWithFlags(&ws.stream, hipStreamNonBlocking));
    per_device.emplace(dev, ws);
    return ws;
}

void launch_lower_bound(hipStream_t st, const double* d_cdf, uint64_t n, const double* d_u, uint64_t shots, void* cdf_scratch,
                        uint64_t* d_out)
{
    if (shots == 0) {
        return;
    }
    k_lower_bound<<<static_cast<unsigned>((shots + BLOCK - 1) / BLOCK), BLOCK, 0, st>>>(d_cdf, n, d_u, shots,
                                                                                       cdf_positive_bits(cdf_scratch, n), d_out);
    KQ_HIP_CHECK(hipGetLastError());
}

namespace
{
// upload Pauli tables into the state's scratch after `reserve` bytes; returns device pointers
void upload_pauli(kq_state* s, size_t reserve, const std::vector<PauliGroup>& groups, const std::vector<PauliZ>& strings,
                  PauliGroup** dg, PauliZ** dz)
{
    const size_t gb = groups.size() * sizeof(PauliGroup);
    const size_t zb = strings.size() * sizeof(PauliZ);
    auto* base = static_cast<char*>(scratch(s, reserve + gb + zb + 256));
    *dg = reinterpret_cast<PauliGroup*>(base + reserve);
    *dz = reinterpret_cast<PauliZ*>(base + reserve + ((gb + 127) / 128) * 128);
    if (reserve + ((gb + 127) / 128) * 128 + zb > reserve + gb + zb + 256) {
        throw InvalidArg {"pauli: scratch layout"};
    }
    if (gb) {
        KQ_HIP_CHECK(hipMemcpyAsync(*dg, groups.data(), gb, hipMemcpyHostToDevice, s->stream));
    }
    if (zb) {
        KQ_HIP_CHECK(hipMemcpyAsync(*dz, strings.data(), zb, hipMemcpyHostToDevice, s->stream));
    }
}
}  // namespace

void launch_pauli_expectation(kq_state* s, const std::vector<PauliGroup>& groups, const std::vector<PauliZ>& strings,
                              double out[2])
{
    const size_t reserve = (RED_BLOCKS + 1) * sizeof(double2);
    PauliGroup* dg;
    PauliZ* dz;
    upload_pauli(s, reserve, groups, strings, &dg, &dz);
    auto* part = static_cast<double2*>(s->scratch);
    const unsigned blocks = grid_for(s->n_amps, 8, RED_BLOCKS);
    const int ng = static_cast<int>(groups.size());
    if (s->nvec <= 32) {
        KQ_DISPATCH(s,
                    (k_pauli_expect<double, true><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(s->amps), s->n_amps, dg, ng, dz, part)),
                    (k_pauli_expect<float, true><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(s->amps), s->n_amps, dg, ng, dz, part)));
    }
    else {
        KQ_DISPATCH(s,
                    (k_pauli_expect<double, false><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(s->amps), s->n_amps, dg, ng, dz, part)),
                    (k_pauli_expect<float, false><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(s->amps), s->n_amps, dg, ng, dz, part)));
    }
    k_reduce_partials<<<1, BLOCK, 0, s->stream>>>(part, static_cast<int>(blocks), part + RED_BLOCKS);
    KQ_HIP_CHECK(hipGetLastError());
    double2 h;
    KQ_HIP_CHECK(hipMemcpyAsync(&h, part + RED_BLOCKS, sizeof h, hipMemcpyDeviceToHost, s->stream));
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    out[0] = h.x;
    out[1] = h.y;
}

void launch_pauli_apply_sum(kq_state* s, const void* d_in, void* d_out, const std::vector<PauliGroup>& groups,
                            const std::vector<PauliZ>& strings)
{
    PauliGroup* dg;
    PauliZ* dz;
    upload_pauli(s, 0, groups, strings, &dg, &dz);
    const int ng = static_cast<int>(groups.size());
    const uint64_t n_tiles = s->n_amps >> PAULI_TILE_BITS;
    if (d_in != d_out && n_tiles >= 16) {
        const unsigned tb = static_cast<unsigned>(std::min<uint64_t>(n_tiles, uint64_t {1} << 20));
        KQ_DISPATCH(s,
                    (k_pauli_apply_tiled<double><<<tb, BLOCK, 0, s->stream>>>(static_cast<const double2*>(d_in), static_cast<double2*>(d_out), n_tiles, dg, ng, dz)),
                    (k_pauli_apply_tiled<float><<<tb, BLOCK, 0, s->stream>>>(static_cast<const float2*>(d_in), static_cast<float2*>(d_out), n_tiles, dg, ng, dz)));
        return;
    }
    const unsigned blocks = grid_for(s->n_amps, PAULI_ROWS);
    if (s->nvec <= 32) {
        KQ_DISPATCH(s,
                    (k_pauli_apply_sum<double, true><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(d_in), static_cast<double2*>(d_out), s->n_amps, dg, ng, dz)),
                    (k_pauli_apply_sum<float, true><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(d_in), static_cast<float2*>(d_out), s->n_amps, dg, ng, dz)));
    }
    else {
        KQ_DISPATCH(s,
                    (k_pauli_apply_sum<double, false><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(d_in), static_cast<double2*>(d_out), s->n_amps, dg, ng, dz)),
                    (k_pauli_apply_sum<float, false><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(d_in), static_cast<float2*>(d_out), s->n_amps, dg, ng, dz)));
    }
}

void launch_pauli_apply_one(kq_state* s, uint64_t x, uint64_t z, const double w[2])
{
    const unsigned blocks = grid_for(s->n_amps, 4);
    KQ_DISPATCH(s,
                (k_pauli_apply_one<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<double2*>(s->amps), s->n_amps, x, z, w[0], w[1])),
                (k_pauli_apply_one<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<float2*>(s->amps), s->n_amps, x, z, w[0], w[1])));
}

void launch_dm_trace(kq_state* s, double out[2])
{
    const uint64_t N = uint64_t {1} << s->n;
    const unsigned blocks = grid_for(N, 4, RED_BLOCKS);
    reduce2(s, blocks, [&](double2* part) {
        KQ_DISPATCH(s,
                    (k_dm_trace<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(s->amps), N, part)),
                    (k_dm_trace<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(s->amps), N, part)));
    }, out);
}

void launch_dm_trace_of_square(kq_state* s, double out[2])
{
    const uint64_t N = uint64_t {1} << s->n;
    const uint32_t T = static_cast<uint32_t>(std::min<uint64_t>(N, TSQ_TILE));
    const uint64_t pairs = (N / T) * (N / T);
    const unsigned blocks = static_cast<unsigned>(std::min<uint64_t>(pairs, RED_BLOCKS));
    reduce2(s, blocks, [&](double2* part) {
        KQ_DISPATCH(s,
                    (k_dm_trace_sq<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(s->amps), N, T, part)),
                    (k_dm_trace_sq<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(s->amps), N, T, part)));
    }, out);
}

void launch_dm_partial_trace(kq_state* s, const std::vector<uint32_t>& traced_desc, double* out_host)
{
    // traced_desc: traced qubits in descending order (bit 0 of the counter = highest qubit)
    const uint32_t n = s->n;
    const uint32_t nt = static_cast<uint32_t>(traced_desc.size());
    if (nt > 16) {
        throw InvalidArg {"partial trace: at most 16 traced qubits"};
    }
    uint64_t lo = 0, hi = 0, traced_mask = 0;
    for (uint32_t j = 0; j < nt; ++j) {
        traced_mask |= uint64_t {1} << traced_desc[j];
        if (j < 8) {
            lo |= static_cast<uint64_t>(traced_desc[j]) << (8 * j);
        }
        else {
            hi |= static_cast<uint64_t>(traced_desc[j]) << (8 * (j - 8));
        }
    }
    const uint64_t kept_mask = ((uint64_t {1} << n) - 1) & ~traced_mask;
    const uint64_t side = uint64_t {1} << (n - nt);
    const uint64_t n_out = side * side;
    auto* d_out = static_cast<double2*>(scratch(s, n_out * sizeof(double2)));
    const unsigned blocks = static_cast<unsigned>((n_out + BLOCK - 1) / BLOCK);
    KQ_DISPATCH(s,
                (k_dm_partial_trace<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(s->amps), n, kept_mask, lo, hi, nt, side, d_out)),
                (k_dm_partial_trace<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(s->amps), n, kept_mask, lo, hi, nt, side, d_out)));
    KQ_HIP_CHECK(hipMemcpyAsync(out_host, d_out, n_out * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
}

void launch_dm_from_sv(kq_state* dm, const kq_state* sv)
{
    const uint64_t N = uint64_t {1} << sv->n;
    const unsigned blocks = grid_for(N * N, 4);
    KQ_DISPATCH(dm,
                (k_dm_from_sv<double><<<blocks, BLOCK, 0, dm->stream>>>(static_cast<double2*>(dm->amps), static_cast<const double2*>(sv->amps), N)),
                (k_dm_from_sv<float><<<blocks, BLOCK, 0, dm->stream>>>(static_cast<float2*>(dm->amps), static_cast<const float2*>(sv->amps), N)));
}

void launch_pair_map_1q(uint32_t n, uint32_t t, uint64_t* d_s0, uint64_t* d_s1)
{
    const uint64_t np = uint64_t {1} << (n - 1);
    k_pair_map_1q<<<grid_for(np, 1), BLOCK>>>(n, t, d_s0, d_s1);
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_pair_map_2q(uint32_t n, uint32_t c, uint32_t t, int four_way, uint64_t* d_out)
{
    const uint64_t np = n < 2 ? 0 : (uint64_t {1} << (n - 2));
    if (np == 0) {
        return;
    }
    k_pair_map_2q<<<grid_for(np, 1), BLOCK>>>(n, c, t, four_way, d_out);
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_adjoint_gate(kq_state* s, void* psi, void* lam, const GateRec& inv, const GateRec* deriv, double* d_acc)
{
    const bool ctrl = inv.control >= 0;
    const uint64_t items = s->n_amps >> (ctrl ? 2 : 1);
    if (items == 0) {
        return;
    }
    const unsigned blocks = grid_for(items, 4, RED_BLOCKS);
    const DevGate di = to_dev(inv);
    const DevGate dd = deriv ? to_dev(*deriv) : DevGate {};
    auto* part = static_cast<double2*>(scratch(s, (RED_BLOCKS + 1) * sizeof(double2)));
    const int has_d = deriv ? 1 : 0;
    KQ_DISPATCH(s,
                (k_adjoint_gate<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<double2*>(psi), static_cast<double2*>(lam), items, di, dd, has_d, part)),
                (k_adjoint_gate<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<float2*>(psi), static_cast<float2*>(lam), items, di, dd, has_d, part)));
    if (deriv) {
        k_reduce_partials<<<1, BLOCK, 0, s->stream>>>(part, static_cast<int>(blocks), reinterpret_cast<double2*>(d_acc));
        KQ_HIP_CHECK(hipGetLastError());
    }
}

void launch_sum_rows(hipStream_t st, const double* rows, const uint64_t* d_offs, const uint32_t* d_counts, uint32_t n,
                     double* d_out)
{
    if (n == 0) {
        return;
    }
    k_sum_rows<<<n, BLOCK, 0, st>>>(rows, d_offs, d_counts, d_out);
    KQ_HIP_CHECK(hipGetLastError());
}

uint32_t batch_blocks(uint64_t n_amps) { return static_cast<uint32_t>((n_amps + BATCH_AMPS - 1) / BATCH_AMPS); }

void launch_pauli_batch(hipStream_t st, int in_precision, int out_precision, const void* arena_a, void* arena_b,
                        const BatchPauliItem* d_items, uint32_t n_items, uint32_t n_blocks, double* d_part)
{
    if (n_items == 0 || n_blocks == 0) {
        return;
    }
    if (in_precision == KQ_C64 && out_precision == KQ_C64) {
        k_pauli_batch<float, float><<<n_blocks, BLOCK, 0, st>>>(static_cast<const float2*>(arena_a), static_cast<float2*>(arena_b),
                                                                d_items, n_items, d_part);
    }
    else if (in_precision == KQ_C128 && out_precision == KQ_C64) {
        k_pauli_batch<double, float><<<n_blocks, BLOCK, 0, st>>>(static_cast<const double2*>(arena_a), static_cast<float2*>(arena_b),
                                                                 d_items, n_items, d_part);
    }
    else if (in_precision == KQ_C128 && out_precision == KQ_C128) {
        k_pauli_batch<double, double><<<n_blocks, BLOCK, 0, st>>>(static_cast<const double2*>(arena_a), static_cast<double2*>(arena_b),
                                                                  d_items, n_items, d_part);
    }
    else {
        throw InvalidArg {"batch: unsupported precision pair"};
    }
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_narrow(hipStream_t st, const void* in, void* out, uint64_t n)
{
    const unsigned blocks = grid_for(n, 4);
    k_narrow<<<blocks, BLOCK, 0, st>>>(static_cast<const double2*>(in), static_cast<float2*>(out), n);
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_basis_distance_batch(hipStream_t st, int precision, const void* arena, const BatchPauliItem* d_items,
                                 uint32_t n_items, uint32_t n_blocks, double* d_part)
{
    if (n_items == 0 || n_blocks == 0) {
        return;
    }
    if (precision == KQ_C64) {
        k_basis_distance_batch<float><<<n_blocks, BLOCK, 0, st>>>(static_cast<const float2*>(arena), d_items, n_items, d_part);
    }
    else {
        k_basis_distance_batch<double><<<n_blocks, BLOCK, 0, st>>>(static_cast<const double2*>(arena), d_items, n_items, d_part);
    }
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_batch_reset(hipStream_t st, int precision, void* arena, uint64_t n_amps, const uint64_t* d_offs, uint32_t n)
{
    const unsigned blocks = grid_for(n_amps, 4);
    const unsigned ob = (n + BLOCK - 1) / BLOCK;
    if (precision == KQ_C64) {
        k_fill_zero<float><<<blocks, BLOCK, 0, st>>>(static_cast<float2*>(arena), n_amps);
        k_set_ones<float><<<ob, BLOCK, 0, st>>>(static_cast<float2*>(arena), d_offs, n);
    }
    else {
        k_fill_zero<double><<<blocks, BLOCK, 0, st>>>(static_cast<double2*>(arena), n_amps);
        k_set_ones<double><<<ob, BLOCK, 0, st>>>(static_cast<double2*>(arena), d_offs, n);
    }
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_max_rows(hipStream_t st, const double* rows, const uint64_t* d_offs, const uint32_t* d_counts, uint32_t n,
                     double* d_out)
{
    if (n == 0) {
        return;
    }
    k_max_rows<<<n, BLOCK, 0, st>>>(rows, d_offs, d_counts, d_out);
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_copy(kq_state* s, void* dst, const void* src, size_t bytes)
{
    KQ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s->stream));
}

// ---- HBM stream probe ---------------------------------------------------------------------------------
// What this device's HBM delivers to plain streams of 16-byte elements, 16 per thread in flight:
// read only, write only, read-modify-write in place, copy; each in six launch shapes (a persistent
// grid of 2 or 8 blocks per CU; one block per 4096-element chunk like a tile pass, with plain or
// nontemporal loads and stores, and each of those with the wave priority raised while the loads
// issue, as the tile passes do).  The reference point for the tile passes, which read and write
// every amplitude once.
template <int MODE, bool NT, bool PRIO>
__global__ __launch_bounds__(BLOCK) void k_stream_probe(const double2* __restrict__ src, double2* __restrict__ dst,
                                                        uint64_t n, double* __restrict__ sink)
{
    constexpr int U = 16;
    const uint64_t chunk = uint64_t {BLOCK} * U;
    double acc = 0.0;
    for (uint64_t c = blockIdx.x * chunk; c + chunk <= n; c += uint64_t {gridDim.x} * chunk) {
        double2 v[U];
        if constexpr (PRIO) {
            __builtin_amdgcn_s_setprio(3);
        }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const uint64_t i = c + uint64_t {BLOCK} * j + threadIdx.x;
            if constexpr (MODE == 1) {
                v[j] = make_double2(static_cast<double>(j), 1.0);
            }
            else if constexpr (NT) {
                v[j].x = __builtin_nontemporal_load(&src[i].x);
                v[j].y = __builtin_nontemporal_load(&src[i].y);
            }
            else {
                v[j] = src[i];
            }
        }
        if constexpr (PRIO) {
            __builtin_amdgcn_s_setprio(0);
        }
        if constexpr (MODE == 0) {
#pragma unroll
            for (int j = 0; j < U; ++j) {
                acc += v[j].x + v[j].y;
            }
        }
        else {
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const uint64_t i = c + uint64_t {BLOCK} * j + threadIdx.x;
                if constexpr (NT) {
                    __builtin_nontemporal_store(v[j].x, &dst[i].x);
                    __builtin_nontemporal_store(v[j].y, &dst[i].y);
                }
                else {
                    dst[i] = v[j];
                }
            }
        }
    }
    if (MODE == 0 && acc == -1.0) {
        sink[0] = acc;   // keeps the reads; never true for the zero-filled probe buffers
    }
}

void stream_probe(uint32_t n_qubits, double gbps[24])
{
    const uint64_t n = uint64_t {1} << n_qubits;
    double2* a = nullptr;
    double2* b = nullptr;
    double* sink = nullptr;
    KQ_HIP_CHECK(hipMalloc(&a, n * sizeof(double2)));
    const auto release = [&] {
        (void) hipFree(a);
        (void) hipFree(b);
        (void) hipFree(sink);
    };
    try {
        KQ_HIP_CHECK(hipMalloc(&b, n * sizeof(double2)));
        KQ_HIP_CHECK(hipMalloc(&sink, 64));
        KQ_HIP_CHECK(hipMemset(a, 0, n * sizeof(double2)));
        KQ_HIP_CHECK(hipMemset(b, 0, n * sizeof(double2)));
        int dev = 0;
        KQ_HIP_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        KQ_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        const auto cus = static_cast<unsigned>(prop.multiProcessorCount);
        // n >= 2^16, so every shape has at least 16 blocks and the chunk loop covers all n elements
        const auto chunks = static_cast<unsigned>(n / (uint64_t {BLOCK} * 16));
        const unsigned grids[6] = {2 * cus, 8 * cus, chunks, chunks, chunks, chunks};
        hipEvent_t e0, e1;
        KQ_HIP_CHECK(hipEventCreate(&e0));
        KQ_HIP_CHECK(hipEventCreate(&e1));
        const double one = 16.0 * static_cast<double>(n);
        const auto time = [&](auto launch, double bytes) {
            constexpr int reps = 5;
            launch();
            KQ_HIP_CHECK(hipEventRecord(e0, nullptr));
            for (int i = 0; i < reps; ++i) {
                launch();
            }
            KQ_HIP_CHECK(hipEventRecord(e1, nullptr));
            KQ_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            KQ_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return bytes / (static_cast<double>(ms) / reps * 1.0e-3) / 1.0e9;
        };
        const auto shape = [&]<bool NT, bool PRIO>(unsigned grid, double* g) {
            g[0] = time([&] { k_stream_probe<0, NT, PRIO><<<grid, BLOCK>>>(a, b, n, sink); }, one);
            g[1] = time([&] { k_stream_probe<1, NT, PRIO><<<grid, BLOCK>>>(a, b, n, sink); }, one);
            g[2] = time([&] { k_stream_probe<2, NT, PRIO><<<grid, BLOCK>>>(a, a, n, sink); }, 2 * one);
            g[3] = time([&] { k_stream_probe<3, NT, PRIO><<<grid, BLOCK>>>(a, b, n, sink); }, 2 * one);
            KQ_HIP_CHECK(hipGetLastError());
        };
        shape.template operator()<false, false>(grids[0], gbps + 0);
        shape.template operator()<false, false>(grids[1], gbps + 4);
        shape.template operator()<false, false>(grids[2], gbps + 8);
        shape.template operator()<true, false>(grids[3], gbps + 12);
        shape.template operator()<false, true>(grids[4], gbps + 16);
        shape.template operator()<true, true>(grids[5], gbps + 20);
        (void) hipEventDestroy(e0);
        (void) hipEventDestroy(e1);
    }
    catch (...) {
        release();
        throw;
    }
    release();
}

void launch_copy_kernel(kq_state* s, void* dst, const void* src, uint64_t n_amps)
{
    const unsigned blocks = grid_for(n_amps, 4);
    KQ_DISPATCH(s,
                (k_copy<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<double2*>(dst), static_cast<const double2*>(src), n_amps)),
                (k_copy<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<float2*>(dst), static_cast<const float2*>(src), n_amps)));
}

// ---- index-bit permutation ---------------------------------------------------------------------------
// dst[pi(i)] = src[i] for every index i, pi a permutation of the index bits, in one pass.  A workgroup
// moves the 2^12 amplitudes that share the values of the bits outside a 12-bit tile set T: it loads
// them with lanes on the source's lowest bits (contiguous reads), exchanges them through LDS and
// stores them with lanes on the bits that land on the destination's lowest bits (contiguous writes).
template <typename A>
__global__ __launch_bounds__(256) void k_permute_bits(const A* __restrict__ src, A* __restrict__ dst, PermuteTables t)
{
    __shared__ A tile[4096];
    const uint32_t tid = threadIdx.x;
    uint64_t sb = 0, db = 0;
    // each XCD (workgroups dealt round robin) takes a contiguous eighth of the tiles
    const uint64_t blk = t.n_free >= 3 ? (uint64_t {blockIdx.x} >> 3) | (uint64_t {blockIdx.x & 7u} << (t.n_free - 3))
                                       : uint64_t {blockIdx.x};
    for (uint32_t i = 0; i < t.n_free; ++i) {
        if ((blk >> i) & 1u) {
            sb |= uint64_t {1} << t.free_src[i];
            db |= uint64_t {1} << t.free_dst[i];
        }
    }
    const auto lds = [](uint32_t l) { return l ^ (((l >> 4) ^ (l >> 8)) & 15u); };
    uint32_t lt = 0;
    uint64_t so = sb;
    for (int j = 0; j < 8; ++j) {
        if ((tid >> j) & 1u) {
            lt |= 1u << t.ld_thr[j];
            so |= t.src_of[t.ld_thr[j]];
        }
    }
    A v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        uint64_t o = so;
        for (int q = 0; q < 4; ++q) {
            if ((r >> q) & 1) {
                o |= t.src_of[t.ld_reg[q]];
            }
        }
        v[r] = src[o];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        uint32_t l = lt;
        for (int q = 0; q < 4; ++q) {
            if ((r >> q) & 1) {
                l |= 1u << t.ld_reg[q];
            }
        }
        tile[lds(l)] = v[r];
    }
    __syncthreads();
    lt = 0;
    uint64_t d0 = db;
    for (int j = 0; j < 8; ++j) {
        if ((tid >> j) & 1u) {
            lt |= 1u << t.st_thr[j];
            d0 |= t.dst_of[t.st_thr[j]];
        }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        uint32_t l = lt;
        uint64_t o = d0;
        for (int q = 0; q < 4; ++q) {
            if ((r >> q) & 1) {
                l |= 1u << t.st_reg[q];
                o |= t.dst_of[t.st_reg[q]];
            }
        }
        dst[o] = tile[lds(l)];
    }
}

void launch_permute_bits(kq_state* s, const void* src, void* dst, const PermuteTables& t)
{
    const unsigned blocks = static_cast<unsigned>(s->n_amps >> 12);
    KQ_DISPATCH(s,
                (k_permute_bits<double2><<<blocks, 256, 0, s->stream>>>(static_cast<const double2*>(src), static_cast<double2*>(dst), t)),
                (k_permute_bits<float2><<<blocks, 256, 0, s->stream>>>(static_cast<const float2*>(src), static_cast<float2*>(dst), t)));
}

// max over the block partials of k_max_abs_diff, into out[0] (one block, fixed order)
__global__ __launch_bounds__(BLOCK) void k_max_partials(const double2* part, int n, double* out)
{
    __shared__ double red[BLOCK];
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        mx = fmax(mx, part[i].x);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0];
    }
}

// dst <- src when *drift > tol (every workgroup reads the same flag), counting the restores
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_restore_if(Amp<R>* dst, const Amp<R>* src, uint64_t n, const double* drift,
                                                      double tol, uint32_t* count)
{
    if (!(*drift > tol)) {
        return;
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
    for (uint64_t i = blockIdx.x * uint64_t {BLOCK} + threadIdx.x; i < n; i += stride) {
        dst[i] = src[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(count, 1u);
    }
}

void launch_drift_restore(kq_state* s, void* psi, const void* ckpt, double tol, double2* d_part, double* d_drift,
                          uint32_t* d_count)
{
    const unsigned blocks = grid_for(s->n_amps, 8, RED_BLOCKS);
    KQ_DISPATCH(s,
                (k_max_abs_diff<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(psi), static_cast<const double2*>(ckpt), s->n_amps, d_part)),
                (k_max_abs_diff<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(psi), static_cast<const float2*>(ckpt), s->n_amps, d_part)));
    KQ_HIP_CHECK(hipGetLastError());
    k_max_partials<<<1, BLOCK, 0, s->stream>>>(d_part, static_cast<int>(blocks), d_drift);
    KQ_HIP_CHECK(hipGetLastError());
    const unsigned cblocks = grid_for(s->n_amps, 4);
    KQ_DISPATCH(s,
                (k_restore_if<double><<<cblocks, BLOCK, 0, s->stream>>>(static_cast<double2*>(psi), static_cast<const double2*>(ckpt), s->n_amps, d_drift, tol, d_count)),
                (k_restore_if<float><<<cblocks, BLOCK, 0, s->stream>>>(static_cast<float2*>(psi), static_cast<const float2*>(ckpt), s->n_amps, d_drift, tol, d_count)));
    KQ_HIP_CHECK(hipGetLastError());
}

double launch_max_abs_diff(kq_state* s, const void* a, const void* b)
{
    const unsigned blocks = grid_for(s->n_amps, 8, RED_BLOCKS);
    auto* part = static_cast<double2*>(scratch(s, (RED_BLOCKS + 1) * sizeof(double2)));
    KQ_DISPATCH(s,
                (k_max_abs_diff<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(a), static_cast<const double2*>(b), s->n_amps, part)),
                (k_max_abs_diff<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(a), static_cast<const float2*>(b), s->n_amps, part)));
    std::vector<double2> h(blocks);
    KQ_HIP_CHECK(hipMemcpyAsync(h.data(), part, blocks * sizeof(double2), hipMemcpyDeviceToHost, s->stream));
    KQ_HIP_CHECK(hipStreamSynchronize(s->stream));
    double mx = 0.0;
    for (const auto& v : h) {
        mx = std::max(mx, v.x);
    }
    return mx;
}

void launch_dot_to_device(kq_state* s, const void* a, const void* b, double2* d_part, double2* d_out)
{
    const unsigned blocks = grid_for(s->n_amps, 8, RED_BLOCKS);
    KQ_DISPATCH(s,
                (k_dot<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(a), static_cast<const double2*>(b), s->n_amps, d_part)),
                (k_dot<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(a), static_cast<const float2*>(b), s->n_amps, d_part)));
    KQ_HIP_CHECK(hipGetLastError());
    k_reduce_partials<<<1, BLOCK, 0, s->stream>>>(d_part, static_cast<int>(blocks), d_out);
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_dot(kq_state* s, const void* a, const void* b, double out[2])
{
    const unsigned blocks = grid_for(s->n_amps, 8, RED_BLOCKS);
    reduce2(s, blocks, [&](double2* part) {
        KQ_DISPATCH(s,
                    (k_dot<double><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const double2*>(a), static_cast<const double2*>(b), s->n_amps, part)),
                    (k_dot<float><<<blocks, BLOCK, 0, s->stream>>>(static_cast<const float2*>(a), static_cast<const float2*>(b), s->n_amps, part)));
    }, out);
}

}  // namespace kq
