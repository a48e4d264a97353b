// Device templates of the tile pass, shared by the compiled tile kernels (kq_tile.hip) and the pass
// kernels generated at run time (kq_jit.cpp).  Self-contained: it compiles under hipcc after the HIP
// runtime header and under hiprtc, whose prelude supplies the fixed-width integer typedefs.
#pragma once

namespace kq
{
namespace dev
{

// gate kinds (the values of kq::TileKind)
enum : int
{
    K_DENSE = 0,
    K_REAL = 1,
    K_AXIS = 2,
    K_DIAG = 3,
    K_DIAG1 = 4,
    K_SWAP = 5
};

template <typename R> struct AmpT;
template <> struct AmpT<double> { using T = double2; };
template <> struct AmpT<float> { using T = float2; };
template <typename R> using Amp = typename AmpT<R>::T;

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

// Workgroup barrier ordering LDS accesses only: global loads still in flight (the next tile's
// amplitudes) are not waited for, which a full __syncthreads() fence would do.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Wave priority (0..3) for the instructions that follow: a wave that is about to issue its tile's HBM
// loads or stores goes ahead of the waves that are computing, so the memory pipeline stays fed.
template <int PRIO>
__device__ __forceinline__ void wave_priority()
{
    __builtin_amdgcn_s_setprio(PRIO);
}

// HBM access of one amplitude, with the nontemporal hint (streamed once per pass, not worth keeping in
// the caches) when NT is set; the two halves merge into one 8- or 16-byte access
template <bool NT, typename A>
__device__ __forceinline__ A amp_load(const A* p)
{
    if constexpr (NT) {
        A x;
        x.x = __builtin_nontemporal_load(&p->x);
        x.y = __builtin_nontemporal_load(&p->y);
        return x;
    }
    else {
        return *p;
    }
}
template <bool NT, typename A>
__device__ __forceinline__ void amp_store(A* p, const A& x)
{
    if constexpr (NT) {
        __builtin_nontemporal_store(x.x, &p->x);
        __builtin_nontemporal_store(x.y, &p->y);
    }
    else {
        *p = x;
    }
}

// the eight coefficients of a gate record, already in registers
template <typename R>
struct Coef
{
    R m[8];
};

// The coefficients of X G X (G with its rows and its columns exchanged) where `fl` holds, else G's own:
// a gate that meets controlled X gates still pending on its bit runs in this form, and they stay pending.
template <typename R>
__device__ __forceinline__ Coef<R> flip_coef(const Coef<R>& c, bool fl)
{
    return {{fl ? c.m[6] : c.m[0], fl ? c.m[7] : c.m[1], fl ? c.m[4] : c.m[2], fl ? c.m[5] : c.m[3],
             fl ? c.m[2] : c.m[4], fl ? c.m[3] : c.m[5], fl ? c.m[0] : c.m[6], fl ? c.m[1] : c.m[7]}};
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
__device__ __forceinline__ void apply(Cx<R> (&v)[16], const Coef<R>& m)
{
    if constexpr (C == P + 1) {
        __builtin_unreachable();   // never emitted by the host
    }
    else {
        const R m0 = m.m[0], m1 = m.m[1], m2 = m.m[2], m3 = m.m[3], m4 = m.m[4], m5 = m.m[5], m6 = m.m[6], m7 = m.m[7];
        Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
            if constexpr (((r >> P) & 1) == 0 && (C == 0 || ((r >> (C - 1)) & 1) == 1)) {
                constexpr int r1 = r | (1 << P);
                const Cx<R> a = v[r];
                const Cx<R> b = v[r1];
                if constexpr (KIND == K_SWAP) {
                    v[r] = b;
                    v[r1] = a;
                }
                else if constexpr (KIND == K_REAL) {
                    v[r] = {a.re * m0 + b.re * m2, a.im * m0 + b.im * m2};
                    v[r1] = {a.re * m4 + b.re * m6, a.im * m4 + b.im * m6};
                }
                else if constexpr (KIND == K_AXIS) {
                    // m00 = m0, m01 = i m3, m10 = i m5, m11 = m6
                    v[r] = {a.re * m0 - b.im * m3, a.im * m0 + b.re * m3};
                    v[r1] = {b.re * m6 - a.im * m5, b.im * m6 + a.re * m5};
                }
                else if constexpr (KIND == K_DIAG1) {
                    v[r1] = cmul(b, m6, m7);
                }
                else if constexpr (KIND == K_DIAG) {
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

// Uncontrolled one-qubit gates with a uniform factor taken out (the pass multiplies the factor back in
// once, before its store), so that one entry of each row or column is 1 and its products disappear.
// Pairs (v[r], v[r | 1 << P]) over the slots r with bit P clear.
//   FORM 0: [[1, t], [-t, 1]]   FORM 1: [[1, t], [t, -1]]     (real rotation / reflection, pivot m00)
//   FORM 2: [[t, 1], [-1, t]]   FORM 3: [[t, 1], [1, -t]]     (pivot m01)
template <typename R, int P, int FORM>
__device__ __forceinline__ void orth(Cx<R> (&v)[16], R t)
{
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (((r >> P) & 1) == 0) {
            constexpr int r1 = r | (1 << P);
            const Cx<R> a = v[r];
            const Cx<R> b = v[r1];
            if constexpr (FORM == 0) {
                v[r] = {a.re + t * b.re, a.im + t * b.im};
                v[r1] = {b.re - t * a.re, b.im - t * a.im};
            }
            else if constexpr (FORM == 1) {
                v[r] = {a.re + t * b.re, a.im + t * b.im};
                v[r1] = {t * a.re - b.re, t * a.im - b.im};
            }
            else if constexpr (FORM == 2) {
                v[r] = {t * a.re + b.re, t * a.im + b.im};
                v[r1] = {t * b.re - a.re, t * b.im - a.im};
            }
            else {
                v[r] = {t * a.re + b.re, t * a.im + b.im};
                v[r1] = {a.re - t * b.re, a.im - t * b.im};
            }
        }
    });
}

//   FORM 0: [[1, i t], [i t, 1]]   FORM 1: [[t, i], [i, t]]   (axis rotation, pivot m00 / m01)
template <typename R, int P, int FORM>
__device__ __forceinline__ void axis(Cx<R> (&v)[16], R t)
{
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (((r >> P) & 1) == 0) {
            constexpr int r1 = r | (1 << P);
            const Cx<R> a = v[r];
            const Cx<R> b = v[r1];
            if constexpr (FORM == 0) {
                v[r] = {a.re - t * b.im, a.im + t * b.re};
                v[r1] = {b.re - t * a.im, b.im + t * a.re};
            }
            else {
                v[r] = {t * a.re - b.im, t * a.im + b.re};
                v[r1] = {t * b.re - a.im, t * b.im + a.re};
            }
        }
    });
}

//   FORM 0: [[1, x], [y, z]]   FORM 1: [[x, 1], [y, z]]   (complex, pivot m00 / m01)
template <typename R, int P, int FORM>
__device__ __forceinline__ void dense1(Cx<R> (&v)[16], R xr, R xi, R yr, R yi, R zr, R zi)
{
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (((r >> P) & 1) == 0) {
            constexpr int r1 = r | (1 << P);
            const Cx<R> a = v[r];
            const Cx<R> b = v[r1];
            if constexpr (FORM == 0) {
                v[r] = {a.re + xr * b.re - xi * b.im, a.im + xr * b.im + xi * b.re};
            }
            else {
                v[r] = {xr * a.re - xi * a.im + b.re, xr * a.im + xi * a.re + b.im};
            }
            v[r1] = {yr * a.re - yi * a.im + zr * b.re - zi * b.im, yr * a.im + yi * a.re + zr * b.im + zi * b.re};
        }
    });
}

// every slot times a real factor
template <typename R>
__device__ __forceinline__ void scale_real(Cx<R> (&v)[16], R f)
{
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) { v[r] = {v[r].re * f, v[r].im * f}; });
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

// <w| dG |v> restricted to the slots a derivative record touches, real part, FP64 accumulation.
// Register target P (any matrix, dense formula) with register control C.
template <typename R, int P, int C>
__device__ __forceinline__ double ddot(const Cx<R> (&v)[16], const Cx<R> (&w)[16], const Coef<R>& m)
{
    const R m0 = m.m[0], m1 = m.m[1], m2 = m.m[2], m3 = m.m[3], m4 = m.m[4], m5 = m.m[5], m6 = m.m[6], m7 = m.m[7];
    double acc = 0.0;
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (((r >> P) & 1) == 0 && (C == 0 || ((r >> (C - 1)) & 1) == 1) && C != P + 1) {
            constexpr int r1 = r | (1 << P);
            const Cx<R> a = v[r];
            const Cx<R> b = v[r1];
            const R d0r = a.re * m0 - a.im * m1 + b.re * m2 - b.im * m3;
            const R d0i = a.re * m1 + a.im * m0 + b.re * m3 + b.im * m2;
            const R d1r = a.re * m4 - a.im * m5 + b.re * m6 - b.im * m7;
            const R d1i = a.re * m5 + a.im * m4 + b.re * m7 + b.im * m6;
            acc += static_cast<double>(w[r].re * d0r + w[r].im * d0i) + static_cast<double>(w[r1].re * d1r + w[r1].im * d1i);
        }
    });
    return acc;
}

// Re <w| Q |v> over the pairs of register bit P (register control C) for the fixed generators of the
// angle gates' derivatives (dG/dtheta = Q G): GEN 0: (-i/2) X, 1: (-i/2) Y, 2: (-i/2) Z, 3: i |1><1|.
// Products per pair in R, summed in FP64.
template <typename R, int GEN, int P, int C>
__device__ __forceinline__ double ddot_gen(const Cx<R> (&v)[16], const Cx<R> (&w)[16])
{
    double acc = 0.0;
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (((r >> P) & 1) == 0 && (C == 0 || ((r >> (C - 1)) & 1) == 1) && C != P + 1) {
            constexpr int r1 = r | (1 << P);
            const Cx<R> a = v[r], b = v[r1], w0 = w[r], w1 = w[r1];
            R t;
            if constexpr (GEN == 0) {
                t = (w0.re * b.im - w0.im * b.re) + (w1.re * a.im - w1.im * a.re);
            }
            else if constexpr (GEN == 1) {
                t = (w1.re * a.re + w1.im * a.im) - (w0.re * b.re + w0.im * b.im);
            }
            else if constexpr (GEN == 2) {
                t = (w0.re * a.im - w0.im * a.re) - (w1.re * b.im - w1.im * b.re);
            }
            else {
                t = w1.im * b.re - w1.re * b.im;
            }
            acc += static_cast<double>(t);
        }
    });
    return GEN == 3 ? acc : 0.5 * acc;
}

// uniform diagonal derivative: sum over the slots (with register control C) of Re(conj(w) f v)
template <typename R, int C>
__device__ __forceinline__ double ddot_uniform(const Cx<R> (&v)[16], const Cx<R> (&w)[16], R fr, R fi)
{
    double acc = 0.0;
    Unroll<0, 16>::run([&]<int r>() __attribute__((always_inline)) {
        if constexpr (C == 0 || ((r >> (C - 1)) & 1) == 1) {
            const Cx<R> d = cmul(v[r], fr, fi);
            acc += static_cast<double>(w[r].re * d.re + w[r].im * d.im);
        }
    });
    return acc;
}

// Run-time generated passes read coefficients from a per-pass buffer of doubles: 8 per gate
// (complex128), or 4 doubles holding 8 packed floats (complex64).
template <typename R>
__device__ __forceinline__ Coef<R> coef_at(const double* __restrict__ cf)
{
    Coef<R> c;
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            c.m[i] = cf[i];
        }
    }
    else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float2 f = __builtin_bit_cast(float2, cf[i]);
            c.m[2 * i] = f.x;
            c.m[2 * i + 1] = f.y;
        }
    }
    return c;
}

// diagonal term (d0, d1) from the buffer: 4 doubles (complex128) or 2 doubles of packed floats
template <typename R>
__device__ __forceinline__ void term_at(const double* __restrict__ cf, R& d0r, R& d0i, R& d1r, R& d1i)
{
    if constexpr (sizeof(R) == 8) {
        d0r = cf[0];
        d0i = cf[1];
        d1r = cf[2];
        d1i = cf[3];
    }
    else {
        const float2 a = __builtin_bit_cast(float2, cf[0]);
        const float2 b = __builtin_bit_cast(float2, cf[1]);
        d0r = a.x;
        d0i = a.y;
        d1r = b.x;
        d1i = b.y;
    }
}

// f *= d where the condition holds
template <typename R>
__device__ __forceinline__ void mul_if(bool ok, R& fr, R& fi, R dr, R di)
{
    const R nr = fr * dr - fi * di;
    const R ni = fr * di + fi * dr;
    fr = ok ? nr : fr;
    fi = ok ? ni : fi;
}

}  // namespace dev
}  // namespace kq
