// nlse_common.h -- shared device-side definitions for the NLSE RK4 kernels.
//
// All kernels are compiled with -ffp-contract=off: every multiply and add
// rounds separately, in the operand order of the reference CPU integrators
// (src/NLSEmagic{1,2}D/NLSEmagic{1,2}D_CPU.c, src/NLSEmagic3D/NLSE3D_TAKE_STEPS_*.c),
// so the double-precision results are bit-identical to them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace nlse {

// complex value stored interleaved (real, imaginary) in HBM and LDS
template <typename T> struct C2;
template <> struct C2<double> { using type = double2; };
template <> struct C2<float> { using type = float2; };
template <typename T> using c2 = typename C2<T>::type;

// store with a streaming (nontemporal) hint: for a kernel's outputs that the
// next launch reads after the kernel boundary, from the Infinity Cache
template <typename T> __device__ __forceinline__ void st_stream(c2<T>* p, c2<T> v)
{
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

template <typename T> __host__ __device__ inline c2<T> mk(T x, T y) { c2<T> r; r.x = x; r.y = y; return r; }

// f(integral_constant<int, B>), ..., f(integral_constant<int, E-1>): a loop
// whose index is a compile-time constant in every copy of the body
template <int B, int E, class F> __device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Scalars of one integrator call, precomputed on the host exactly as the
// reference file of each dimension does (coefs1d, coefs2d, coefs3d in nlse_host.h).
template <typename T> struct Coefs {
    T s, a, k, k2, k6;
    T ah2;            // a/h^2 (CD)
    T lh2, l_a;       // 1/h^2, 1/a (2SHOC)
    T a76, a112;      // 1D 2SHOC
    T a1_6h2, a1_12;  // 2D 2SHOC
    T a_6h2, a_12;    // 3D 2SHOC
};

// boundary-condition codes of the reference (NLSE3D_TAKE_STEPS_CD.c:18,
// NLSEmagic1D_CPU.c:33-76); any other value behaves like a zero boundary
enum { BC_DIRICHLET = 1, BC_MSD = 2, BC_LAP0 = 3, BC_ONESIDED = 4 };

// U += t for a state kept as an unevaluated sum hi + lo (compensated final
// combine of the single-precision integrator): TwoSum(hi, t) [Knuth], the
// rounding error folded into lo, then FastTwoSum renormalisation (|s| >= |l|).
template <typename T> __device__ __forceinline__ void comp_add(T& hi, T& lo, T t)
{
    const T s = hi + t;
    const T bb = s - hi;
    const T e = (hi - (s - bb)) + (t - bb);
    const T l = lo + e;
    const T h = s + l;
    lo = l - (h - s);
    hi = h;
}

// RK4 stage index -> epilogue (NLSEmagic1D_CPU.c:405-418)
//  1: ktot = K;          Y = U + k/2 K
//  2: ktot = ktot + 2K;  Y = U + k/2 K
//  3: ktot = ktot + 2K;  Y = U + k K
//  4: ktot = ktot + 1K;  U = U + k/6 ktot
// `ulo` (single precision only, else null): the low part of a compensated
// state; stage 4 then adds k/6 ktot to hi + lo with comp_add.
template <typename T>
__device__ __forceinline__ void rk4_epilogue(int stage, const Coefs<T>& c, c2<T> f, c2<T> u,
                                             c2<T>* __restrict__ acc, c2<T>* __restrict__ yout,
                                             c2<T>* __restrict__ uout, long p, c2<T>* ulo = nullptr)
{
    if (stage == 1) {
        acc[p] = f;
        yout[p] = mk<T>(u.x + c.k2 * f.x, u.y + c.k2 * f.y);
    } else if (stage == 4) {
        c2<T> a = acc[p];
        a.x = a.x + (T)1.0 * f.x;
        a.y = a.y + (T)1.0 * f.y;
        if (ulo) {
            c2<T> lo = ulo[p];
            comp_add<T>(u.x, lo.x, c.k6 * a.x);
            comp_add<T>(u.y, lo.y, c.k6 * a.y);
            ulo[p] = lo;
            uout[p] = u;
        } else {
            uout[p] = mk<T>(u.x + c.k6 * a.x, u.y + c.k6 * a.y);
        }
    } else {
        c2<T> a = acc[p];
        a.x = a.x + (T)2.0 * f.x;
        a.y = a.y + (T)2.0 * f.y;
        acc[p] = a;
        const T kk = (stage == 2) ? c.k2 : c.k;
        yout[p] = mk<T>(u.x + kk * f.x, u.y + kk * f.y);
    }
}

// DPP lane shifts over the whole 64-lane wave (GFX9 wavefront shifts)
constexpr int DPP_NEXT = 0x130;  // wave_shl:1 -- lane l reads lane l+1 (column j+1)
constexpr int DPP_PREV = 0x138;  // wave_shr:1 -- lane l reads lane l-1 (column j-1)

template <int CTRL> __device__ __forceinline__ float lane_shift(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL> __device__ __forceinline__ double lane_shift(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL, typename T> __device__ __forceinline__ c2<T> lane_shift2(c2<T> v)
{
    return mk<T>(lane_shift<CTRL>(v.x), lane_shift<CTRL>(v.y));
}

// s|U|^2 - V: the bracket of "- (s*(Ur*Ur + Ui*Ui) - V)*Ui"
template <typename T> __device__ __forceinline__ T nlin(T s, c2<T> u, T v)
{
    return s * (u.x * u.x + u.y * u.y) - v;
}

// MSD: boundary F from the interior partner's F and U (NLSEmagic2D_CPU.c:64-68)
template <typename T> __device__ __forceinline__ c2<T> msd_f(c2<T> fq, c2<T> uq, c2<T> ub)
{
    T om = (fq.y * uq.x - fq.x * uq.y) / (uq.x * uq.x + uq.y * uq.y);
    return mk<T>(-om * ub.y, om * ub.x);
}

// MSD: boundary D from the partner's D (NLSEmagic2D_CPU.c:297-303)
template <typename T>
__device__ __forceinline__ c2<T> msd_d(c2<T> dq, c2<T> uq, T vq, c2<T> ub, T vb, const Coefs<T>& c)
{
    T A = (dq.x * uq.x + dq.y * uq.y) / (uq.x * uq.x + uq.y * uq.y);
    T nb = nlin<T>(c.s, ub, vb);
    T nb1 = nlin<T>(c.s, uq, vq);
    return mk<T>((A + c.l_a * (nb1 - nb)) * ub.x, (A + c.l_a * (nb1 - nb)) * ub.y);
}

// Lap0 boundary F (NLSEmagic2D_CPU.c:109-112)
template <typename T> __device__ __forceinline__ c2<T> lap0_f(c2<T> ub, T vb, T s)
{
    T n = nlin<T>(s, ub, vb);
    return mk<T>(-n * ub.y, n * ub.x);
}

// Dirichlet boundary D for 2SHOC (NLSEmagic2D_CPU.c:282-285)
template <typename T> __device__ __forceinline__ c2<T> dir_d(c2<T> ub, T vb, const Coefs<T>& c)
{
    T t = c.l_a * (vb - c.s * (ub.x * ub.x + ub.y * ub.y));
    return mk<T>(t * ub.x, t * ub.y);
}

__device__ __forceinline__ int clamp_in(int x, int n) { return x < 1 ? 1 : (x > n - 2 ? n - 2 : x); }

// Launch-time description of a tiled 2D or 3D sweep.  Tiles are numbered so
// that consecutive tiles in space land on the same XCD (workgroups are dealt
// round-robin to the 8 XCDs; each XCD has its own L2), which keeps the halo
// re-reads of neighbouring tiles inside one L2.
struct TileMap {
    int ntiles;   // tiles in the sweep
    int per_xcd;  // ceil(ntiles / 8)
    __device__ __forceinline__ int tile() const
    {
        int b = blockIdx.x;
        return (b & 7) * per_xcd + (b >> 3);
    }
};

}  // namespace nlse
