// nlse1d.hip -- 1D NLSE RK4 kernels on MI355X.
//
// Replaces compute_F_CD / compute_F_2SHOC and the vec_add passes of
// src/NLSEmagic1D/NLSEmagic1D_CPU.c:3-447 (single precision :449-523).
// Two forms:
//   - stage1d_kernel: one fused RK4 stage over 256-point tiles (any N), the
//     stage input staged in LDS with a halo of 3 (the one-sided BC=4 formula
//     reaches three points inward);
//   - chunk1d_resident_kernel: a whole chunk of steps in ONE workgroup for
//     N <= 4096; U and ktot stay in registers, the stage input and D live in
//     LDS, and the four stages of every step are separated by barriers only.
//     This is the launch-free form for the small 1D problems the reference
//     targets (N = 2048): a grid-wide launch per stage would be latency bound;
//   - block1d_kernel (the default): S whole RK4 steps per launch, one
//     workgroup per segment of 64, 128 or 256 points (the largest that still
//     gives 128 segments, or 64), each on its own CU.  A workgroup loads its
//     segment plus a halo of 4*R*S points (R = 1 CD, 2 2SHOC: the reach of
//     one stage), runs the S steps on chip like the resident kernel, and
//     writes back its own points; values near the window edge go stale by R
//     points per stage and never reach the owned points.  A wave whose
//     points all lie more than R * (stages left) from the segment skips the
//     stage: the owned points depend on nothing that far out, and no point
//     they depend on reads it.  The boundary
//     formulas reach up to three points inward in one stage (MSD through the
//     partner's stencil, the one-sided BC 4 difference), further than R, so
//     the segment origins shift back until the last segment owns at least
//     four points: a boundary point's inward reach then lands on owned
//     points, whose dependence on the window stays R per stage.  Launches
//     alternate between two state buffers.
// Arithmetic follows the reference operand order with FP contraction off.
#include <algorithm>

#include "nlse_common.h"
#include "nlse_launch.h"

namespace nlse {

// one-sided second difference at end b pointing inward by d (NLSEmagic1D_CPU.c:60-68)
template <typename T, typename Get>
__device__ __forceinline__ T os1d(Get get, int b, int d)
{
    return -get(b + 3 * d) + 4 * get(b + 2 * d) - 5 * get(b + d) + 2 * get(b);
}

// Evaluates D (2SHOC) and F at global points from an accessor of the stage
// input; shared by the tiled and the resident kernels.
template <typename T, int METHOD, typename GetY, typename GetV, typename GetD>
struct Rhs1d {
    GetY Y;
    GetV Vv;
    GetD Dd;
    const Coefs<T>& c;
    int N, BC;

    __device__ __forceinline__ c2<T> d_int(int i) const
    {
        c2<T> a = Y(i + 1), m = Y(i), b = Y(i - 1);
        return mk<T>((a.x - 2 * m.x + b.x) * c.lh2, (a.y - 2 * m.y + b.y) * c.lh2);
    }
    // D at any point (NLSEmagic1D_CPU.c:150-198)
    __device__ __forceinline__ c2<T> d_at(int i) const
    {
        if (i > 0 && i < N - 1) return d_int(i);
        const int q = (i == 0) ? 1 : N - 2, dir = (i == 0) ? 1 : -1;
        switch (BC) {
        case BC_DIRICHLET: return dir_d<T>(Y(i), Vv(i), c);
        case BC_MSD: return msd_d<T>(d_int(q), Y(q), Vv(q), Y(i), Vv(i), c);
        case BC_ONESIDED: {
            auto yr = [&](int j) { return Y(j).x; };
            auto yi = [&](int j) { return Y(j).y; };
            return mk<T>(os1d<T>(yr, i, dir) * c.lh2, os1d<T>(yi, i, dir) * c.lh2);
        }
        default: return mk<T>(0, 0);
        }
    }
    // F at an interior point: CD (NLSEmagic1D_CPU.c:27-30) or 2SHOC (:204-207)
    __device__ __forceinline__ c2<T> f_int(int i) const
    {
        const c2<T> y = Y(i);
        const T n = nlin<T>(c.s, y, Vv(i));
        if (METHOD == 1) {
            const c2<T> a = Y(i + 1), b = Y(i - 1);
            return mk<T>(-c.ah2 * (a.y - 2 * y.y + b.y) - n * y.y, c.ah2 * (a.x - 2 * y.x + b.x) + n * y.x);
        } else {
            const c2<T> dc = Dd(i), da = Dd(i + 1), db = Dd(i - 1);
            return mk<T>(c.a112 * (da.y + db.y) - c.a76 * dc.y - n * y.y,
                         c.a76 * dc.x - c.a112 * (da.x + db.x) + n * y.x);
        }
    }
    // F at any point, with the boundary rules (NLSEmagic1D_CPU.c:33-76, 210-253)
    __device__ __forceinline__ c2<T> f_at(int i) const
    {
        if (i > 0 && i < N - 1) return f_int(i);
        const int q = (i == 0) ? 1 : N - 2, dir = (i == 0) ? 1 : -1;
        switch (BC) {
        case BC_MSD: return msd_f<T>(f_int(q), Y(q), Y(i));
        case BC_LAP0: return lap0_f<T>(Y(i), Vv(i), c.s);
        case BC_ONESIDED: {
            const c2<T> y = Y(i);
            const T n = nlin<T>(c.s, y, Vv(i));
            auto yr = [&](int j) { return Y(j).x; };
            auto yi = [&](int j) { return Y(j).y; };
            if (METHOD == 1)
                return mk<T>(-c.ah2 * os1d<T>(yi, i, dir) - n * y.y, c.ah2 * os1d<T>(yr, i, dir) + n * y.x);
            return mk<T>(-c.a * os1d<T>(yi, i, dir) * c.lh2 - n * y.y,
                         c.a * os1d<T>(yr, i, dir) * c.lh2 + n * y.x);
        }
        default: return mk<T>(0, 0);
        }
    }
};

template <typename T, int METHOD, typename GY, typename GV, typename GD>
__device__ __forceinline__ Rhs1d<T, METHOD, GY, GV, GD> make_rhs1d(GY y, GV v, GD d, const Coefs<T>& c,
                                                                   int N, int BC)
{
    return Rhs1d<T, METHOD, GY, GV, GD>{y, v, d, c, N, BC};
}

// ------------------------------------------------------------- tiled stage

template <typename T, int METHOD>
__global__ __launch_bounds__(256) void stage1d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                      c2<T>* __restrict__ acc, c2<T>* yout,
                                                      const T* __restrict__ V, Coefs<T> c, int N,
                                                      int toff, int BC, int stage, c2<T>* ulo)
{
    constexpr int TN = 256, H = 3;
    __shared__ c2<T> sY[TN + 2 * H];
    __shared__ c2<T> sD[(METHOD == 2) ? TN + 2 : 1];
    const int i0 = blockIdx.x * TN - toff;
    const int tid = threadIdx.x;
    for (int e = tid; e < TN + 2 * H; e += TN) {
        const int gi = i0 - H + e;
        sY[e] = (gi >= 0 && gi < N) ? yin[gi] : mk<T>(0, 0);
    }
    __syncthreads();
    auto Y = [&](int gi) { return sY[gi - i0 + H]; };
    auto Vv = [&](int gi) { return V[gi]; };
    auto Dd = [&](int gi) { return sD[gi - i0 + 1]; };
    auto rhs = make_rhs1d<T, METHOD>(Y, Vv, Dd, c, N, BC);
    if (METHOD == 2) {
        for (int e = tid; e < TN + 2; e += TN) {
            const int gi = i0 - 1 + e;
            if (gi >= 0 && gi < N) sD[e] = rhs.d_at(gi);
        }
        __syncthreads();
    }
    const int gi = i0 + tid;
    if (gi < 0 || gi >= N) return;
    const c2<T> f = rhs.f_at(gi);
    const c2<T> uu = (stage == 1) ? Y(gi) : u[gi];
    rk4_epilogue<T>(stage, c, f, uu, acc, yout, yout, gi, ulo);
}

template <typename T>
hipError_t launch_stage1d(const StageArgs<T>& a, hipStream_t st)
{
    const int toff = (a.N % 256 == 1) ? 1 : 0;
    const int nb = (a.N + toff + 255) / 256;
    if (a.method == 2)
        hipLaunchKernelGGL((stage1d_kernel<T, 2>), dim3(nb), dim3(256), 0, st, a.yin, a.u, a.acc, a.yout, a.V,
                           a.c, a.N, toff, a.BC, a.stage, a.ulo);
    else
        hipLaunchKernelGGL((stage1d_kernel<T, 1>), dim3(nb), dim3(256), 0, st, a.yin, a.u, a.acc, a.yout, a.V,
                           a.c, a.N, toff, a.BC, a.stage, a.ulo);
    return hipGetLastError();
}

// ------------------------------------------------ whole chunk, one workgroup

// 1024 threads: four waves on each SIMD of the CU hide the LDS and barrier
// latency of the stage-by-stage exchange
constexpr int RES_THREADS = 1024;
constexpr int RES_PPT = 4;  // points per thread -> N <= 4096

int chunk1d_resident_max_points(int elem_bytes) { (void)elem_bytes; return RES_THREADS * RES_PPT; }

template <typename T, int METHOD, int PPT>
__global__ __launch_bounds__(RES_THREADS) void chunk1d_resident_kernel(c2<T>* u_io, c2<T>* ulo_io,
                                                                       const T* __restrict__ Vg, Coefs<T> c, int N,
                                                                       int BC, int steps)
{
    constexpr int NMAX = RES_THREADS * PPT;
    __shared__ c2<T> sY[NMAX];
    __shared__ c2<T> sD[(METHOD == 2) ? NMAX : 1];
    const int tid = threadIdx.x;
    c2<T> U[PPT], A[PPT], Y[PPT], Lo[PPT];
    const bool comp = ulo_io != nullptr;
#pragma unroll
    for (int r = 0; r < PPT; ++r) {
        const int i = tid + RES_THREADS * r;
        U[r] = (i < N) ? u_io[i] : mk<T>(0, 0);
        Lo[r] = (comp && i < N) ? ulo_io[i] : mk<T>(0, 0);
        Y[r] = U[r];
        A[r] = mk<T>(0, 0);
    }
    auto Yl = [&](int i) { return sY[i]; };
    auto Vl = [&](int i) { return Vg[i]; };
    auto Dl = [&](int i) { return sD[i]; };
    auto rhs = make_rhs1d<T, METHOD>(Yl, Vl, Dl, c, N, BC);
    for (int step = 0; step < steps; ++step) {
#pragma unroll 1
        for (int stage = 1; stage <= 4; ++stage) {
#pragma unroll
            for (int r = 0; r < PPT; ++r) {
                const int i = tid + RES_THREADS * r;
                if (i < N) sY[i] = Y[r];
            }
            __syncthreads();
            if (METHOD == 2) {
#pragma unroll
                for (int r = 0; r < PPT; ++r) {
                    const int i = tid + RES_THREADS * r;
                    if (i < N) sD[i] = rhs.d_at(i);
                }
                __syncthreads();
            }
#pragma unroll
            for (int r = 0; r < PPT; ++r) {
                const int i = tid + RES_THREADS * r;
                if (i >= N) continue;
                const c2<T> f = rhs.f_at(i);
                if (stage == 1) {
                    A[r] = f;
                    Y[r] = mk<T>(U[r].x + c.k2 * f.x, U[r].y + c.k2 * f.y);
                } else if (stage == 4) {
                    A[r].x = A[r].x + (T)1.0 * f.x;
                    A[r].y = A[r].y + (T)1.0 * f.y;
                    if (comp) {
                        comp_add<T>(U[r].x, Lo[r].x, c.k6 * A[r].x);
                        comp_add<T>(U[r].y, Lo[r].y, c.k6 * A[r].y);
                    } else {
                        U[r] = mk<T>(U[r].x + c.k6 * A[r].x, U[r].y + c.k6 * A[r].y);
                    }
                    Y[r] = U[r];
                } else {
                    A[r].x = A[r].x + (T)2.0 * f.x;
                    A[r].y = A[r].y + (T)2.0 * f.y;
                    const T kk = (stage == 2) ? c.k2 : c.k;
                    Y[r] = mk<T>(U[r].x + kk * f.x, U[r].y + kk * f.y);
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int r = 0; r < PPT; ++r) {
        const int i = tid + RES_THREADS * r;
        if (i < N) u_io[i] = U[r];
        if (comp && i < N) ulo_io[i] = Lo[r];
    }
}

template <typename T, int PPT>
static hipError_t launch_res(c2<T>* u, c2<T>* ulo, const T* V, Coefs<T> c, int N, int BC, int method, int steps,
                             hipStream_t st)
{
    if (method == 2)
        hipLaunchKernelGGL((chunk1d_resident_kernel<T, 2, PPT>), dim3(1), dim3(RES_THREADS), 0, st, u, ulo, V, c,
                           N, BC, steps);
    else
        hipLaunchKernelGGL((chunk1d_resident_kernel<T, 1, PPT>), dim3(1), dim3(RES_THREADS), 0, st, u, ulo, V, c,
                           N, BC, steps);
    return hipGetLastError();
}

// ------------------------------------------ time-blocked windows, many CUs

constexpr int BLK_SEG = 256;  // the largest segment: owned points per workgroup

// One wave per 64 consecutive points: a core of CW = 64 - 2 HW points and a
// halo of HW = 4 R points (one RK4 step's reach) each side, copies of the
// neighbouring waves' core points.  A wave runs the four stages of a step on
// its own: the stage input (and D) of its 64 points goes to a wave-private
// LDS slice and the stencils (with the reference's boundary rules) read it
// there, so no workgroup barrier is needed within a step (neighbours by DPP
// instead ran 6 % slower: the lane shifts of FP64 values cost more than the
// LDS round trip).  The halo lanes go stale by R points per stage and
// are refreshed from the neighbours' cores (U and U_lo through LDS, one
// workgroup barrier) at the end of every step.  The window's outer halo of
// 4 R S points covers the S steps of the launch as in the resident kernel.
template <typename T, int METHOD, bool COMP>
__global__ __launch_bounds__(1024) void block1d_kernel(const c2<T>* __restrict__ uin, c2<T>* __restrict__ uout,
                                                       const c2<T>* __restrict__ loin, c2<T>* __restrict__ loout,
                                                       const T* __restrict__ Vg, Coefs<T> c, int N, int BC,
                                                       int steps, int halo, int seg, int soff)
{
    constexpr int R = (METHOD == 2) ? 2 : 1;
    constexpr int HW = 4 * R, CW = 64 - 2 * HW;
    constexpr int PAD = 4;  // the boundary formulas reach three points inward
    // wave-private stage input and D (padding lanes hold values no core
    // point depends on); window-indexed U and U_lo of the cores, by step parity
    __shared__ c2<T> sW[16][64 + 2 * PAD];
    __shared__ c2<T> sWD[(METHOD == 2) ? 16 : 1][(METHOD == 2) ? 64 + 2 * PAD : 1];
    __shared__ c2<T> sX[2][16 * CW];
    __shared__ c2<T> sXL[COMP ? 2 : 1][COMP ? 16 * CW : 1];
    const int o0 = max((int)blockIdx.x * seg - soff, 0), o1 = min((int)(blockIdx.x + 1) * seg - soff, N);
    const int w0 = max(o0 - halo, 0), w1 = min(o1 + halo, N);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = (w1 - w0 + CW - 1) / CW;  // waves with a core in the window
    const int ib = w0 + w * CW - HW;         // point of lane 0
    const int i = ib + lane;
    const bool core = lane >= HW && lane < 64 - HW && i < w1;
    const int ic = i < 0 ? 0 : (i >= N ? N - 1 : i);
    c2<T> U = uin[ic], Lo = COMP ? loin[ic] : mk<T>(0, 0), A = mk<T>(0, 0), Y = U;
    const T Vr = Vg[ic];
    auto Vl = [&](int gi) { return gi == i ? Vr : Vg[gi]; };
    auto Yl = [&](int gi) { return sW[w][gi - ib + PAD]; };
    auto Dl = [&](int gi) { return sWD[METHOD == 2 ? w : 0][gi - ib + PAD]; };
    auto rhs = make_rhs1d<T, METHOD>(Yl, Vl, Dl, c, N, BC);
    const bool ingrid = i >= 0 && i < N;
    // distance of this wave's nearest point from the segment (uniform)
    const int whi = ib + 63;
    const int wdist = whi < o0 ? o0 - whi : (ib > o1 - 1 ? ib - (o1 - 1) : 0);
    const int G = 4 * steps;  // stages in this launch
    for (int step = 0; step < steps; ++step) {
        static_for<0, 4>([&](auto qc) {
            constexpr int stage = decltype(qc)::value + 1;
            // this stage's output is needed within R * (G - 1 - g) of the
            // segment, its input within R * (G - g); waves beyond that, and
            // waves with no core in the window, skip the stage
            const bool live = w < nw && wdist <= R * (G - (4 * step + stage - 1));
            if (!live) return;
            sW[w][PAD + lane] = Y;
            __builtin_amdgcn_wave_barrier();
            if constexpr (METHOD == 2) {
                const c2<T> d = ingrid ? rhs.d_at(i) : mk<T>(0, 0);
                sWD[w][PAD + lane] = d;
                __builtin_amdgcn_wave_barrier();
            }
            const c2<T> f = ingrid ? rhs.f_at(i) : mk<T>(0, 0);
            if constexpr (stage == 1) {
                A = f;
                Y = mk<T>(U.x + c.k2 * f.x, U.y + c.k2 * f.y);
            } else if constexpr (stage == 4) {
                A.x = A.x + (T)1.0 * f.x;
                A.y = A.y + (T)1.0 * f.y;
                if constexpr (COMP) {
                    comp_add<T>(U.x, Lo.x, c.k6 * A.x);
                    comp_add<T>(U.y, Lo.y, c.k6 * A.y);
                } else {
                    U = mk<T>(U.x + c.k6 * A.x, U.y + c.k6 * A.y);
                }
            } else {
                A.x = A.x + (T)2.0 * f.x;
                A.y = A.y + (T)2.0 * f.y;
                const T kk = (stage == 2) ? c.k2 : c.k;
                Y = mk<T>(U.x + kk * f.x, U.y + kk * f.y);
            }
            __builtin_amdgcn_wave_barrier();  // the slices are read before the next stage rewrites them
        });
        // refresh the halo lanes from the neighbouring waves' cores
        const int par = step & 1;
        if (core && w < nw) {
            sX[par][i - w0] = U;
            if constexpr (COMP) sXL[par][i - w0] = Lo;
        }
        __syncthreads();
        const int x = i - w0;
        if (!core && w < nw && x >= 0 && x < nw * CW && i < w1) {
            U = sX[par][x];
            if constexpr (COMP) Lo = sXL[par][x];
        }
        Y = U;
    }
    if (core && i >= o0 && i < o1) {
        uout[i] = U;
        if (COMP) loout[i] = Lo;
    }
}

// steps per launch: the window (a segment of up to BLK_SEG points and a halo
// of 4 R S each side) must fit the cores of 16 waves
int block1d_max_steps(int method)
{
    const int R = (method == 2) ? 2 : 1;
    return (16 * (64 - 8 * R) - BLK_SEG) / (2 * 4 * R);
}

// owned points per workgroup: the largest of 256, 128, 64 that gives at
// least 128 segments, else 64
static int block1d_segment(int N)
{
    int seg = BLK_SEG;
    while (seg > 64 && N / seg < 128) seg /= 2;
    return seg;
}

template <typename T>
hipError_t launch_block1d(const c2<T>* uin, c2<T>* uout, const c2<T>* loin, c2<T>* loout, const T* V, Coefs<T> c,
                          int N, int BC, int method, int steps, int halo_steps, hipStream_t st)
{
    const int R = (method == 2) ? 2 : 1;
    const int halo = 4 * R * halo_steps;
    const int seg = block1d_segment(N);
    const int nb = (N + seg - 1) / seg;
    // segment b owns [b*seg - soff, (b+1)*seg - soff) clipped to [0, N):
    // the last segment owns at least 4 points (see block1d_kernel)
    const int tail = N % seg;
    const int soff = (nb > 1 && tail > 0 && tail < 4) ? 4 - tail : 0;
    const int win = std::min(N, seg + 2 * halo);
    const int CW = 64 - 8 * R;  // core points per wave
    const dim3 blk(((win + CW - 1) / CW) * 64);
    const bool comp = loin != nullptr;
    if (method == 2) {
        if (comp)
            hipLaunchKernelGGL((block1d_kernel<T, 2, true>), dim3(nb), blk, 0, st, uin, uout, loin, loout, V, c, N,
                               BC, steps, halo, seg, soff);
        else
            hipLaunchKernelGGL((block1d_kernel<T, 2, false>), dim3(nb), blk, 0, st, uin, uout, loin, loout, V, c, N,
                               BC, steps, halo, seg, soff);
    } else {
        if (comp)
            hipLaunchKernelGGL((block1d_kernel<T, 1, true>), dim3(nb), blk, 0, st, uin, uout, loin, loout, V, c, N,
                               BC, steps, halo, seg, soff);
        else
            hipLaunchKernelGGL((block1d_kernel<T, 1, false>), dim3(nb), blk, 0, st, uin, uout, loin, loout, V, c, N,
                               BC, steps, halo, seg, soff);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_chunk1d_resident(c2<T>* u, c2<T>* ulo, const T* V, Coefs<T> c, int N, int BC, int method,
                                   int steps, hipStream_t st)
{
    if (N <= RES_THREADS) return launch_res<T, 1>(u, ulo, V, c, N, BC, method, steps, st);
    if (N <= RES_THREADS * 2) return launch_res<T, 2>(u, ulo, V, c, N, BC, method, steps, st);
    return launch_res<T, RES_PPT>(u, ulo, V, c, N, BC, method, steps, st);
}

template hipError_t launch_stage1d<double>(const StageArgs<double>&, hipStream_t);
template hipError_t launch_stage1d<float>(const StageArgs<float>&, hipStream_t);
template hipError_t launch_block1d<double>(const double2*, double2*, const double2*, double2*, const double*,
                                           Coefs<double>, int, int, int, int, int, hipStream_t);
template hipError_t launch_block1d<float>(const float2*, float2*, const float2*, float2*, const float*, Coefs<float>,
                                          int, int, int, int, int, hipStream_t);
template hipError_t launch_chunk1d_resident<double>(double2*, double2*, const double*, Coefs<double>, int, int,
                                                    int, int, hipStream_t);
template hipError_t launch_chunk1d_resident<float>(float2*, float2*, const float*, Coefs<float>, int, int, int,
                                                   int, hipStream_t);

}  // namespace nlse
