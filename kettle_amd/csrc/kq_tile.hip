// The fused tile pass: the hot kernel of simulate() for gfx950.
//
// One workgroup owns one tile: the 2^k amplitudes that share the values of every qubit outside
// the pass's tile set.  The tile is read from HBM once and written once; all gates of the pass
// are applied in between without touching HBM.
//
// Register blocking.  Each thread holds 2^RQ = 16 amplitudes in VGPRs.  Which RQ bits of the local tile index live in a
// thread's registers is the "layout"; the remaining bits form the thread index.  A pass is a
// sequence of stages; within a stage every non-diagonal gate acts on a register bit, so it is
// plain register arithmetic on pairs (v[r], v[r | 1 << p]).  Diagonal gates and controls may sit
// on any bit: register bits are known per register slot, thread bits per thread, and bits outside
// the tile are uniform over the workgroup.  Between stages whose layouts differ the tile goes
// through LDS once (write, barrier, read); LDS addresses are XOR-swizzled on the low four bits so
// lanes whose local indices differ by 16 or 256 land on different banks.
//
// HBM access: local bits 0..c-1 are the contiguous global qubits 0..c-1, and the I/O layout keeps
// local bits 0..2 on lanes, so every wavefront load or store covers runs of >= 8 consecutive
// amplitudes (>= 128 B for complex128).
#include <hip/hip_runtime.h>

#include <cstring>

#include "kq_internal.hpp"

namespace kq
{
namespace
{

constexpr int TILE_BLOCK_MAX = 256;
constexpr int MAX_TILE_Q = 16;
constexpr int MAX_FREE_Q = 48;

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

struct DevStage
{
    uint8_t reg[8];
    uint32_t thr_mask;
    uint32_t g_begin;
    uint32_t g_count;
    uint32_t pad;
};

struct TileArgs
{
    int k;
    int c;
    int nfree;
    int n_stages;
    int load_direct;
    int store_direct;
    uint8_t qpos[MAX_TILE_Q];
    uint8_t freepos[MAX_FREE_Q];
    DevStage io;
};

static_assert(sizeof(DevGate) == sizeof(GateRec), "device gate record layout");
static_assert(sizeof(DevStage) == sizeof(TileStage), "device stage record layout");

__device__ __forceinline__ uint32_t swz(uint32_t l)
{
    return l ^ (((l >> 4) ^ (l >> 8)) & 15u);
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

template <typename R, int RQ>
struct TileRegs
{
    Cx<R> v[1 << RQ];
    uint32_t reg_bit[RQ];   // local bit of register position p
    uint32_t tpart;         // local-index bits contributed by the thread index

    __device__ __forceinline__ void set_layout(const DevStage& st, uint32_t tid)
    {
#pragma unroll
        for (int p = 0; p < RQ; ++p) {
            reg_bit[p] = st.reg[p];
        }
        uint32_t t = 0, src = tid, mask = st.thr_mask;
        while (mask) {
            const uint32_t b = __ffs(mask) - 1;
            t |= (src & 1u) << b;
            src >>= 1;
            mask &= mask - 1;
        }
        tpart = t;
    }

    __device__ __forceinline__ bool same_layout(const DevStage& st) const
    {
        bool same = true;
#pragma unroll
        for (int p = 0; p < RQ; ++p) {
            same = same && (reg_bit[p] == st.reg[p]);
        }
        return same;
    }

    template <int r>
    __device__ __forceinline__ uint32_t local() const
    {
        uint32_t l = tpart;
#pragma unroll
        for (int p = 0; p < RQ; ++p) {
            if ((r >> p) & 1) {
                l |= 1u << reg_bit[p];
            }
        }
        return l;
    }
};

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

// One gate on register position P: pairs (v[r], v[r | 1 << P]).  KIND selects the arithmetic:
// the zero entries of real, axis-rotation and diagonal matrices are skipped at compile time.
// A control on a register bit (csrc == 0) is tested per slot with a uniform (scalar) branch.
template <typename R>
struct Coef
{
    R m[8];
};

template <typename R, int RQ, int P, int KIND>
__device__ __forceinline__ void reg_gate(TileRegs<R, RQ>& t, const Coef<R> c, int csrc, uint32_t cidx)
{
    const R* m = c.m;
    Unroll<0, (1 << RQ)>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (((r >> P) & 1) == 0) {
            constexpr int r1 = r | (1 << P);
            if (csrc != 0 || ((r >> cidx) & 1u)) {
                const Cx<R> a = t.v[r];
                const Cx<R> b = t.v[r1];
                if constexpr (KIND == GK_SWAP) {
                    t.v[r] = b;
                    t.v[r1] = a;
                }
                else if constexpr (KIND == GK_REAL) {
                    t.v[r] = {a.re * m[0] + b.re * m[2], a.im * m[0] + b.im * m[2]};
                    t.v[r1] = {a.re * m[4] + b.re * m[6], a.im * m[4] + b.im * m[6]};
                }
                else if constexpr (KIND == GK_AXIS) {
                    // m00 = m[0], m01 = i m[3], m10 = i m[5], m11 = m[6]
                    t.v[r] = {a.re * m[0] - b.im * m[3], a.im * m[0] + b.re * m[3]};
                    t.v[r1] = {b.re * m[6] - a.im * m[5], b.im * m[6] + a.re * m[5]};
                }
                else if constexpr (KIND == GK_DIAG1) {
                    t.v[r1] = cmul(b, m[6], m[7]);
                }
                else if constexpr (KIND == GK_DIAG) {
                    t.v[r] = cmul(a, m[0], m[1]);
                    t.v[r1] = cmul(b, m[6], m[7]);
                }
                else {
                    t.v[r] = {a.re * m[0] - a.im * m[1] + b.re * m[2] - b.im * m[3],
                              a.re * m[1] + a.im * m[0] + b.re * m[3] + b.im * m[2]};
                    t.v[r1] = {a.re * m[4] - a.im * m[5] + b.re * m[6] - b.im * m[7],
                               a.re * m[5] + a.im * m[4] + b.re * m[7] + b.im * m[6]};
                }
            }
        }
    });
}

template <typename R, int RQ, int KIND>
__device__ __forceinline__ void reg_gate_at(TileRegs<R, RQ>& t, const DevGate& g, uint32_t pos, int csrc, uint32_t cidx)
{
    Coef<R> m;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        m.m[i] = static_cast<R>(g.m[i]);
    }
    switch (pos) {
        case 0: reg_gate<R, RQ, 0, KIND>(t, m, csrc, cidx); break;
        case 1: reg_gate<R, RQ, 1, KIND>(t, m, csrc, cidx); break;
        case 2: reg_gate<R, RQ, 2, KIND>(t, m, csrc, cidx); break;
        case 3: reg_gate<R, RQ, 3, KIND>(t, m, csrc, cidx); break;
        default:
            if constexpr (RQ > 4) {
                reg_gate<R, RQ, (RQ > 4 ? 4 : 0), KIND>(t, m, csrc, cidx);
            }
            break;
    }
}

template <typename R, int RQ>
__global__ __launch_bounds__(TILE_BLOCK_MAX) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_tile_reg(Amp<R>* __restrict__ a, TileArgs d,
                                                             const DevStage* __restrict__ stages,
                                                             const DevGate* __restrict__ gates)
{
    constexpr int RA = 1 << RQ;
    extern __shared__ unsigned char smem_raw[];
    Amp<R>* tile = reinterpret_cast<Amp<R>*>(smem_raw);
    __shared__ uint64_t hoff[1 << 10];

    const uint32_t tid = threadIdx.x;
    const int k = d.k;
    const int c = d.c;
    const int nh = k - c;
    const uint32_t lowmask = (1u << c) - 1;

    uint64_t base = 0;
    {
        const uint64_t b = blockIdx.x;
        for (int j = 0; j < d.nfree; ++j) {
            base |= ((b >> j) & 1ull) << d.freepos[j];
        }
    }
    for (uint32_t h = tid; h < (1u << nh); h += blockDim.x) {
        uint64_t off = 0;
        for (int j = 0; j < nh; ++j) {
            off |= static_cast<uint64_t>((h >> j) & 1u) << d.qpos[c + j];
        }
        hoff[h] = off;
    }
    __syncthreads();

    TileRegs<R, RQ> t;
    const auto gidx = [&](uint32_t l) { return base | (l & lowmask) | hoff[l >> c]; };

    const DevStage s0 = stages[0];
    t.set_layout(d.load_direct ? s0 : d.io, tid);
    Unroll<0, RA>::run([&]<int r>() __attribute__((always_inline)) {
        const Amp<R> x = a[gidx(t.template local<r>())];
        t.v[r] = {x.x, x.y};
    });

    const auto remap = [&](const DevStage& to) {
        __syncthreads();
        Unroll<0, RA>::run([&]<int r>() __attribute__((always_inline)) {
            Amp<R> x;
            x.x = t.v[r].re;
            x.y = t.v[r].im;
            tile[swz(t.template local<r>())] = x;
        });
        __syncthreads();
        t.set_layout(to, tid);
        Unroll<0, RA>::run([&]<int r>() __attribute__((always_inline)) {
            const Amp<R> x = tile[swz(t.template local<r>())];
            t.v[r] = {x.x, x.y};
        });
    };

    for (int si = 0; si < d.n_stages; ++si) {
        const DevStage st = stages[si];
        if (!t.same_layout(st)) {
            remap(st);
        }
        for (uint32_t gi = st.g_begin; gi < st.g_begin + st.g_count; ++gi) {
            const DevGate& g = gates[gi];
            const uint32_t kind = g.kind;
            const int tsrc = static_cast<int>(g.target & 3u);
            const uint32_t tidx = g.target >> 2;
            const bool has_c = g.control >= 0;
            const int csrc = has_c ? static_cast<int>(static_cast<uint32_t>(g.control) & 3u) : 3;
            const uint32_t cidx = has_c ? (static_cast<uint32_t>(g.control) >> 2) : 0u;
            // controls on thread bits or outside the tile are uniform per thread / per tile
            if (csrc == 1 && !((tid >> cidx) & 1u)) {
                continue;
            }
            if (csrc == 2 && !((base >> cidx) & 1ull)) {
                continue;
            }
            if ((kind == GK_DIAG || kind == GK_DIAG1) && tsrc != 0) {
                // the target bit is uniform over the thread: one factor for every slot
                const uint32_t bit = tsrc == 1 ? ((tid >> tidx) & 1u) : static_cast<uint32_t>((base >> tidx) & 1ull);
                if (!bit && kind == GK_DIAG1) {
                    continue;
                }
                const R fr = static_cast<R>(bit ? g.m[6] : g.m[0]);
                const R fi = static_cast<R>(bit ? g.m[7] : g.m[1]);
                Unroll<0, RA>::run([&]<int r>() __attribute__((always_inline)) {
                    if (csrc != 0 || ((r >> cidx) & 1u)) {
                        t.v[r] = cmul(t.v[r], fr, fi);
                    }
                });
                continue;
            }
            // the target is register position tidx
            switch (kind) {
                case GK_SWAP: reg_gate_at<R, RQ, GK_SWAP>(t, g, tidx, csrc, cidx); break;
                case GK_REAL: reg_gate_at<R, RQ, GK_REAL>(t, g, tidx, csrc, cidx); break;
                case GK_AXIS: reg_gate_at<R, RQ, GK_AXIS>(t, g, tidx, csrc, cidx); break;
                case GK_DIAG1: reg_gate_at<R, RQ, GK_DIAG1>(t, g, tidx, csrc, cidx); break;
                case GK_DIAG: reg_gate_at<R, RQ, GK_DIAG>(t, g, tidx, csrc, cidx); break;
                default: reg_gate_at<R, RQ, GK_DENSE>(t, g, tidx, csrc, cidx); break;
            }
        }
    }

    if (!d.store_direct && !t.same_layout(d.io)) {
        remap(d.io);
    }
    Unroll<0, RA>::run([&]<int r>() __attribute__((always_inline)) {
        Amp<R> x;
        x.x = t.v[r].re;
        x.y = t.v[r].im;
        a[gidx(t.template local<r>())] = x;
    });
}

template <typename R, int RQ>
void set_lds_attr()
{
    static bool done = false;
    if (!done) {
        const int maxdyn = static_cast<int>(device_lds_per_block() - TILE_STATIC_LDS);
        (void) hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_reg<R, RQ>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, maxdyn);
        (void) hipGetLastError();
        done = true;
    }
}

}  // namespace

void launch_tile_pass(kq_state* s, const TilePass& p, const void* d_gates, const void* d_stages)
{
    const int rq_kernel = register_qubits(s->precision);
    if (p.rq != rq_kernel || p.k < rq_kernel) {
        throw InvalidArg {"tile pass: register bits do not match the kernel"};
    }
    if (p.k > MAX_TILE_Q || p.k - p.c > 10 || p.stages.empty()) {
        throw InvalidArg {"tile pass: bad tile shape"};
    }
    TileArgs d {};
    d.k = p.k;
    d.c = p.c;
    d.n_stages = static_cast<int>(p.stages.size());
    d.load_direct = p.load_direct ? 1 : 0;
    d.store_direct = p.store_direct ? 1 : 0;
    std::memcpy(&d.io, &p.io_layout, sizeof d.io);
    uint64_t qmask = 0;
    for (int j = 0; j < p.k; ++j) {
        d.qpos[j] = static_cast<uint8_t>(p.qubits[j]);
        qmask |= uint64_t {1} << p.qubits[j];
    }
    int nf = 0;
    for (uint32_t q = 0; q < s->nvec; ++q) {
        if (!((qmask >> q) & 1ull)) {
            if (nf >= MAX_FREE_Q) {
                throw InvalidArg {"tile pass: too many free qubits"};
            }
            d.freepos[nf++] = static_cast<uint8_t>(q);
        }
    }
    d.nfree = nf;
    const uint64_t blocks = uint64_t {1} << nf;
    if (blocks > 0x7fffffffull) {
        throw InvalidArg {"tile pass: grid too large"};
    }
    const size_t lds = (size_t {1} << p.k) * amp_bytes(s->precision);
    if (lds + TILE_STATIC_LDS > device_lds_per_block()) {
        throw InvalidArg {"tile pass: tile does not fit the workgroup's LDS"};
    }
    const unsigned threads = 1u << (p.k - p.rq);
    if (threads > TILE_BLOCK_MAX) {
        throw InvalidArg {"tile pass: too many threads per tile"};
    }
    const auto* g = static_cast<const DevGate*>(d_gates);
    const auto* st = static_cast<const DevStage*>(d_stages);
    if (s->precision == KQ_C64) {
        set_lds_attr<float, 4>();
        k_tile_reg<float, 4><<<static_cast<unsigned>(blocks), threads, lds, s->stream>>>(static_cast<float2*>(s->amps), d, st, g);
    }
    else {
        set_lds_attr<double, 4>();
        k_tile_reg<double, 4><<<static_cast<unsigned>(blocks), threads, lds, s->stream>>>(static_cast<double2*>(s->amps), d, st, g);
    }
    KQ_HIP_CHECK(hipGetLastError());
}

}  // namespace kq
