// Dense 4-qubit block on index bits 0..3: psi[r, :] <- U psi[r, :] for every 16-amplitude row r,
// with U a 16 x 16 complex matrix.  Two forms of the same product, for an A/B of the vector ALUs
// against the matrix cores (a fused block of many gates on four qubits would run as this):
//   VALU: one row per thread, U's entries read as wave-uniform scalars (s_load), 4 FMAs per complex
//         multiply-accumulate, 1024 FMAs per row.
//   MFMA: one 16-row tile (256 amplitudes) per wave, new = Psi_tile x U^T as four real 16x16x16
//         products (re.re, -im.im, re.im, im.re), each four 16x16x4 matrix instructions
//         (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32); U^T's 4 x 16 slices stay in registers.
// Operand layouts of the 16x16x4 instructions (lane l of 64, measured with k_mfma16_layout on gfx950):
// A[i][k] from lane i + 16 k and B[k][j] from lane j + 16 k for both; the result row held in element e
// of lane j + 16 g differs: D[4 g + e][j] for f32, D[4 e + g][j] for f64 (rows interleaved).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kq_internal.hpp"

namespace kq
{
namespace
{

constexpr int DENSE_BLOCK = 256;

template <typename R> struct DenseAmp;
template <> struct DenseAmp<double> { using T = double2; };
template <> struct DenseAmp<float> { using T = float2; };

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename R> struct Mfma16;
template <> struct Mfma16<double>
{
    using V = v4d;
    static __device__ __forceinline__ V mac(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // result row of element e in lane group g
    static __device__ __forceinline__ uint32_t row(uint32_t g, int e) { return 4 * e + g; }
};
template <> struct Mfma16<float>
{
    using V = v4f;
    static __device__ __forceinline__ V mac(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ uint32_t row(uint32_t g, int e) { return 4 * g + e; }
};

// u: 16 x 16 complex, row-major, interleaved (re, im), in the amplitude precision
template <typename R>
__global__ __launch_bounds__(DENSE_BLOCK) void k_dense4_valu(typename DenseAmp<R>::T* __restrict__ a, uint64_t n_rows,
                                                             const R* __restrict__ u)
{
    using A = typename DenseAmp<R>::T;
    for (uint64_t r = uint64_t {blockIdx.x} * DENSE_BLOCK + threadIdx.x; r < n_rows; r += uint64_t {gridDim.x} * DENSE_BLOCK) {
        A* row = a + r * 16;
        A x[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            x[c] = row[c];
        }
        // one output at a time, so that only U's row i (64 scalar registers) is live
#pragma unroll 1
        for (int i = 0; i < 16; ++i) {
            R yr = 0, yi = 0;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const R ur = u[2 * (16 * i + c)];
                const R ui = u[2 * (16 * i + c) + 1];
                yr = fma(ur, x[c].x, yr);
                yr = fma(-ui, x[c].y, yr);
                yi = fma(ur, x[c].y, yi);
                yi = fma(ui, x[c].x, yi);
            }
            A y;
            y.x = yr;
            y.y = yi;
            row[i] = y;
        }
    }
}

// one wave per 256-amplitude tile; the tile loop is wave-uniform, so every MFMA runs on a full wave
template <typename R>
__global__ __launch_bounds__(DENSE_BLOCK) void k_dense4_mfma(typename DenseAmp<R>::T* __restrict__ a, uint64_t n_tiles,
                                                             const R* __restrict__ u)
{
    using A = typename DenseAmp<R>::T;
    using M = Mfma16<R>;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t lo = lane & 15;   // A: tile row; B: output index j; D: output index j
    const uint32_t hi = lane >> 4;   // A, B: k within the 4-wide slice; D: row group
    // B[k][j] = U^T[c][j] = U[j][c], c = 4 kb + hi, j = lo
    R br[4], bi[4], nbi[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        const uint32_t c = 4 * kb + hi;
        br[kb] = u[2 * (16 * lo + c)];
        bi[kb] = u[2 * (16 * lo + c) + 1];
        nbi[kb] = -bi[kb];
    }
    const uint64_t waves = uint64_t {gridDim.x} * (DENSE_BLOCK / 64);
    for (uint64_t t = uint64_t {blockIdx.x} * (DENSE_BLOCK / 64) + (threadIdx.x >> 6); t < n_tiles; t += waves) {
        A* tile = a + t * 256;
        A x[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            x[kb] = tile[16 * lo + 4 * kb + hi];   // A[i = lo][k = hi] of slice kb
        }
        typename M::V dr = {0, 0, 0, 0};
        typename M::V di = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            dr = M::mac(x[kb].x, br[kb], dr);
            dr = M::mac(x[kb].y, nbi[kb], dr);
            di = M::mac(x[kb].x, bi[kb], di);
            di = M::mac(x[kb].y, br[kb], di);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            A y;
            y.x = dr[e];
            y.y = di[e];
            tile[16 * M::row(hi, e) + lo] = y;   // D[row][j = lo]
        }
    }
}

// one wave: d[4 l + e] = element e of lane l after one 16x16x4 instruction with lane l's operands a[l], b[l]
template <typename R>
__global__ __launch_bounds__(64) void k_mfma16_layout(const R* __restrict__ a, const R* __restrict__ b, R* __restrict__ d)
{
    const uint32_t l = threadIdx.x;
    typename Mfma16<R>::V acc = {0, 0, 0, 0};
    acc = Mfma16<R>::mac(a[l], b[l], acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        d[4 * l + e] = acc[e];
    }
}

unsigned dense_grid(uint64_t work_items, uint64_t per_block)
{
    int dev = 0;
    (void) hipGetDevice(&dev);
    int cus = 256;
    (void) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint64_t need = (work_items + per_block - 1) / per_block;
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(need, uint64_t(cus) * 16)));
}

}  // namespace

void launch_mfma16_layout(hipStream_t st, int precision, const void* a, const void* b, void* d)
{
    if (precision == KQ_C128) {
        k_mfma16_layout<double><<<1, 64, 0, st>>>(static_cast<const double*>(a), static_cast<const double*>(b), static_cast<double*>(d));
    }
    else {
        k_mfma16_layout<float><<<1, 64, 0, st>>>(static_cast<const float*>(a), static_cast<const float*>(b), static_cast<float*>(d));
    }
    KQ_HIP_CHECK(hipGetLastError());
}

void launch_dense4(kq_state* s, const void* d_u, bool use_mfma)
{
    if (s->n_amps < 256) {
        throw InvalidArg {"dense block: the state needs at least 8 index bits"};
    }
    const uint64_t rows = s->n_amps / 16;
    const uint64_t tiles = s->n_amps / 256;
    if (s->precision == KQ_C128) {
        auto* a = static_cast<double2*>(s->amps);
        const auto* u = static_cast<const double*>(d_u);
        if (use_mfma) {
            k_dense4_mfma<double><<<dense_grid(tiles, DENSE_BLOCK / 64), DENSE_BLOCK, 0, s->stream>>>(a, tiles, u);
        }
        else {
            k_dense4_valu<double><<<dense_grid(rows, DENSE_BLOCK), DENSE_BLOCK, 0, s->stream>>>(a, rows, u);
        }
    }
    else {
        auto* a = static_cast<float2*>(s->amps);
        const auto* u = static_cast<const float*>(d_u);
        if (use_mfma) {
            k_dense4_mfma<float><<<dense_grid(tiles, DENSE_BLOCK / 64), DENSE_BLOCK, 0, s->stream>>>(a, tiles, u);
        }
        else {
            k_dense4_valu<float><<<dense_grid(rows, DENSE_BLOCK), DENSE_BLOCK, 0, s->stream>>>(a, rows, u);
        }
    }
    KQ_HIP_CHECK(hipGetLastError());
}

}  // namespace kq
