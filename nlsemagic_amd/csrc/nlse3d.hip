// nlse3d.hip -- 3D NLSE RK4 stage kernels on MI355X.
//
// Replaces NLSE3D_STEP and add_cube of src/NLSEmagic3D/NLSE3D_TAKE_STEPS_CD.c
// (:28-223) and NLSE3D_TAKE_STEPS_2SHOC.c (:28-412), with the chunk loop of
// :308-322 / :507-521 on the host.  Element (i, j, k) is at (i*N + j)*M + k:
// i < L slowest, k < M contiguous.
//   CD:    one fused kernel per RK4 stage (7-point stencil + BC + update).
//   2SHOC: one plane-marching kernel per RK4 stage that keeps D on chip
//          (nlse3d_march.h); for MSD boundaries, whose partner can lie in
//          another plane and row, a D pass (D = Lap_CD(Y)/h^2 with its
//          boundary values, written to a scratch field) and a fused F +
//          update pass that reads D's six neighbours and Y's twelve edge
//          neighbours.
// Neighbour reads of the per-point kernels go through the L1/L2 caches: a
// wavefront covers 64 consecutive k, so every plane and row neighbour is a
// coalesced line read.
#include "nlse3d_cdpair.h"
#include "nlse3d_march.h"
#include "nlse3d_shocpair.h"
#include "nlse_common.h"
#include "nlse_launch.h"

namespace nlse {

struct Grid3 {
    int N, M, L;
    long si, sj;  // strides of i and j
};

template <typename T>
__device__ __forceinline__ c2<T> d3_int(const c2<T>* __restrict__ y, long p, const Grid3& g, const Coefs<T>& c)
{
    // NLSE3D_TAKE_STEPS_2SHOC.c:62-69
    const c2<T> a = y[p + g.si], b = y[p - g.si], d = y[p + g.sj], e = y[p - g.sj], f = y[p + 1], h = y[p - 1],
                m = y[p];
    return mk<T>(c.lh2 * (a.x + b.x + d.x + e.x + f.x + h.x - 6 * m.x),
                 c.lh2 * (a.y + b.y + d.y + e.y + f.y + h.y - 6 * m.y));
}

template <typename T, int METHOD>
__device__ __forceinline__ c2<T> f3_int(const c2<T>* __restrict__ y, const c2<T>* __restrict__ D,
                                        const T* __restrict__ V, long p, const Grid3& g, const Coefs<T>& c)
{
    const c2<T> m = y[p];
    const T w = V[p] - c.s * (m.x * m.x + m.y * m.y);
    if (METHOD == 1) {
        // NLSE3D_TAKE_STEPS_CD.c:59-71
        const c2<T> a = y[p + g.si], b = y[p - g.si], d = y[p + g.sj], e = y[p - g.sj], f = y[p + 1],
                    h = y[p - 1];
        return mk<T>(-c.ah2 * (a.y + b.y + d.y + e.y + f.y + h.y - 6 * m.y) + w * m.y,
                     c.ah2 * (a.x + b.x + d.x + e.x + f.x + h.x - 6 * m.x) - w * m.x);
    } else {
        // NLSE3D_TAKE_STEPS_2SHOC.c:239-259; edge neighbours in source order
        const long si = g.si, sj = g.sj;
        const long e12[12] = {si + sj, sj + 1, si + 1, -si - sj, -sj - 1, -si - 1,
                              si - sj, sj - 1, si - 1, -si + sj, -sj + 1, -si + 1};
        c2<T> es = y[p + e12[0]];
#pragma unroll
        for (int t = 1; t < 12; ++t) {
            const c2<T> q = y[p + e12[t]];
            es.x = es.x + q.x;
            es.y = es.y + q.y;
        }
        const c2<T> da = D[p + si], db = D[p - si], dd = D[p + sj], de = D[p - sj], df = D[p + 1], dh = D[p - 1],
                    dm = D[p];
        return mk<T>(c.a_12 * (da.y + db.y + dd.y + de.y + df.y + dh.y - 10 * dm.y) - c.a_6h2 * (es.y - 12 * m.y)
                         + w * m.y,
                     -c.a_12 * (da.x + db.x + dd.x + de.x + df.x + dh.x - 10 * dm.x)
                         + c.a_6h2 * (es.x - 12 * m.x) - w * m.x);
    }
}

__device__ __forceinline__ bool interior3(int i, int j, int k, const Grid3& g)
{
    return i > 0 && i < g.L - 1 && j > 0 && j < g.N - 1 && k > 0 && k < g.M - 1;
}

__device__ __forceinline__ long partner3(int i, int j, int k, const Grid3& g)
{
    return ((long)clamp_in(i, g.L) * g.N + clamp_in(j, g.N)) * g.M + clamp_in(k, g.M);
}

// D with its boundary values (NLSE3D_TAKE_STEPS_2SHOC.c:56-230)
template <typename T>
__global__ __launch_bounds__(256) void d3_kernel(const c2<T>* __restrict__ y, c2<T>* __restrict__ D,
                                                 const T* __restrict__ V, Coefs<T> c, Grid3 g, int BC)
{
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= g.M || j >= g.N) return;
    const long p = ((long)i * g.N + j) * g.M + k;
    c2<T> d;
    if (interior3(i, j, k, g)) {
        d = d3_int<T>(y, p, g, c);
    } else if (BC == BC_DIRICHLET) {
        d = dir_d<T>(y[p], V[p], c);
    } else if (BC == BC_MSD) {
        const long q = partner3(i, j, k, g);
        d = msd_d<T>(d3_int<T>(y, q, g, c), y[q], V[q], y[p], V[p], c);
    } else {
        d = mk<T>(0, 0);
    }
    D[p] = d;
}

// F with its boundary values, then the RK4 stage update
template <typename T, int METHOD>
__global__ __launch_bounds__(256) void f3_kernel(const c2<T>* __restrict__ y, const c2<T>* u, c2<T>* __restrict__ acc,
                                                 c2<T>* yout, const c2<T>* __restrict__ D, const T* __restrict__ V,
                                                 Coefs<T> c, Grid3 g, int BC, int stage, c2<T>* ulo)
{
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= g.M || j >= g.N) return;
    const long p = ((long)i * g.N + j) * g.M + k;
    c2<T> f;
    if (interior3(i, j, k, g)) {
        f = f3_int<T, METHOD>(y, D, V, p, g, c);
    } else if (BC == BC_MSD) {
        const long q = partner3(i, j, k, g);
        f = msd_f<T>(f3_int<T, METHOD>(y, D, V, q, g, c), y[q], y[p]);
    } else if (BC == BC_LAP0) {
        f = lap0_f<T>(y[p], V[p], c.s);
    } else {
        f = mk<T>(0, 0);
    }
    const c2<T> uu = (stage == 1) ? y[p] : u[p];
    rk4_epilogue<T>(stage, c, f, uu, acc, yout, yout, p, ulo);
}

// 2SHOC stage with D on chip: one workgroup per (j tile, k tile, i chunk);
// consecutive tiles of a chunk go to the same XCD
template <typename T, int TJ, int NW>
__global__ __launch_bounds__(NW * 64) void march3d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                      c2<T>* __restrict__ acc, c2<T>* yout, const T* __restrict__ V,
                                                      Coefs<T> c, March3 g, int BC, int stage, c2<T>* ulo)
{
    __shared__ March3Lds<T, TJ> sm;
    const int t = (blockIdx.x & 7) * g.per_xcd + (blockIdx.x >> 3);
    if (t >= g.tj * g.tk * g.ti) return;
    const int kt = t % g.tk, rest = t / g.tk;
    const int jt = rest % g.tj, ci = rest / g.tj;
    march3d_tile<T, TJ, NW>(sm, ci * g.IC, jt * TJ, kt * March3Cfg<TJ>::W, g, yin, u, acc, yout, V, c, BC, stage,
                            ulo);
}

template <typename T, int TJ, int NW>
static hipError_t launch_march3d(const StageArgs<T>& a, int IC, hipStream_t st)
{
    March3 g;
    g.N = a.N;
    g.M = a.M;
    g.L = a.L;
    g.IC = IC;
    g.tj = (a.N + TJ - 1) / TJ;
    g.tk = (a.M + March3Cfg<TJ>::W - 1) / March3Cfg<TJ>::W;
    g.ti = (a.L + IC - 1) / IC;
    g.per_xcd = (g.tj * g.tk * g.ti + 7) / 8;
    hipLaunchKernelGGL((march3d_kernel<T, TJ, NW>), dim3(8 * g.per_xcd), dim3(NW * 64), 0, st, a.yin, a.u, a.acc,
                       a.yout, a.V, a.c, g, a.BC, a.stage, a.ulo);
    return hipGetLastError();
}

// CD, two stages per launch (stages 1-2 or 3-4) on plane-marching tiles
template <typename T, int FIRST, int TJ, int NW>
__global__ __launch_bounds__(NW * 64) void cdpair3d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                           c2<T>* __restrict__ acc, c2<T>* __restrict__ yout,
                                                           const T* __restrict__ V, Coefs<T> c, March3 g, int BC,
                                                           c2<T>* ulo)
{
    __shared__ CdPairLds<T, TJ> sm;
    const int t = (blockIdx.x & 7) * g.per_xcd + (blockIdx.x >> 3);
    if (t >= g.tj * g.tk * g.ti) return;
    const int kt = t % g.tk, rest = t / g.tk;
    const int jt = rest % g.tj, ci = rest / g.tj;
    cdpair_tile<T, FIRST, TJ, NW>(sm, ci * g.IC, jt * TJ, kt * 60, g, yin, u, acc, yout, V, c, BC, ulo);
}

// 2SHOC, two stages per launch on plane-marching tiles (one extended row per
// wave; D and the running ktot in LDS)
template <typename T, int FIRST, int TJ, int NW>
__global__ __launch_bounds__(NW * 64) void shocpair3d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                             c2<T>* __restrict__ acc, c2<T>* __restrict__ yout,
                                                             const T* __restrict__ V, Coefs<T> c, March3 g, int BC,
                                                             c2<T>* ulo)
{
    __shared__ ShocPairLds<T, TJ> sm;
    const int t = (blockIdx.x & 7) * g.per_xcd + (blockIdx.x >> 3);
    if (t >= g.tj * g.tk * g.ti) return;
    const int kt = t % g.tk, rest = t / g.tk;
    const int jt = rest % g.tj, ci = rest / g.tj;
    shocpair_tile<T, FIRST, TJ, NW>(sm, ci * g.IC, jt * TJ, kt * 56, g, yin, u, acc, yout, V, c, BC, ulo);
}

template <typename T>
static hipError_t launch_shocpair3d(const StageArgs<T>& a, bool first, hipStream_t st)
{
    // FP64: 4 owned rows on 12 waves (LDS rings of 160 KiB, one workgroup
    // per CU); FP32: 8 rows on 16 waves (120 KiB)
    constexpr int TJ = std::is_same<T, double>::value ? 4 : 8, NW = TJ + 8;
    if ((long)a.N * a.M > 0x7fffffffL) return hipErrorInvalidValue;  // 32-bit offsets within a plane
    March3 g;
    g.N = a.N;
    g.M = a.M;
    g.L = a.L;
    g.IC = 32;
    g.tj = (a.N + TJ - 1) / TJ;
    g.tk = (a.M + 55) / 56;
    g.ti = (a.L + g.IC - 1) / g.IC;
    g.per_xcd = (g.tj * g.tk * g.ti + 7) / 8;
    if (first)
        hipLaunchKernelGGL((shocpair3d_kernel<T, 1, TJ, NW>), dim3(8 * g.per_xcd), dim3(NW * 64), 0, st, a.yin, a.u,
                           a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
    else
        hipLaunchKernelGGL((shocpair3d_kernel<T, 3, TJ, NW>), dim3(8 * g.per_xcd), dim3(NW * 64), 0, st, a.yin, a.u,
                           a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pair3d(const StageArgs<T>& a, bool first, hipStream_t st)
{
    if (a.method == 2) return launch_shocpair3d<T>(a, first, st);
    // 12 owned rows on 16 waves, one extended row per wave: few registers
    // per wave (stages 3-4 in FP64: 117 VGPRs, against 208 with two rows per
    // wave) and j-halo rows re-read through L2 for 12 owned rows instead of
    // 4 (the halo re-reads dominate the HBM traffic beyond the 128 B per
    // point-step the pair needs)
    constexpr int TJ = 12, NW = TJ + 4;
    March3 g;
    g.N = a.N;
    g.M = a.M;
    g.L = a.L;
    g.IC = 32;
    g.tj = (a.N + TJ - 1) / TJ;
    g.tk = (a.M + 59) / 60;
    g.ti = (a.L + g.IC - 1) / g.IC;
    g.per_xcd = (g.tj * g.tk * g.ti + 7) / 8;
    if (first)
        hipLaunchKernelGGL((cdpair3d_kernel<T, 1, TJ, NW>), dim3(8 * g.per_xcd), dim3(NW * 64), 0, st, a.yin, a.u,
                           a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
    else
        hipLaunchKernelGGL((cdpair3d_kernel<T, 3, TJ, NW>), dim3(8 * g.per_xcd), dim3(NW * 64), 0, st, a.yin, a.u,
                           a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
    return hipGetLastError();
}

template hipError_t launch_pair3d<double>(const StageArgs<double>&, bool, hipStream_t);
template hipError_t launch_pair3d<float>(const StageArgs<float>&, bool, hipStream_t);

template <typename T>
hipError_t launch_stage3d(const StageArgs<T>& a, hipStream_t st)
{
    if (a.method == 2 && a.BC != BC_MSD && a.kernel != KERNEL_TILES) {
        // rows per tile: 4 or 8 (a.tile below 8 selects 4); planes per chunk 32
        if (a.tile < 8) return launch_march3d<T, 4, 4>(a, 32, st);
        return launch_march3d<T, 8, 4>(a, 32, st);
    }
    Grid3 g;
    g.N = a.N;
    g.M = a.M;
    g.L = a.L;
    g.sj = a.M;
    g.si = (long)a.N * a.M;
    dim3 grid((a.M + 63) / 64, (a.N + 3) / 4, a.L);
    if (a.method == 2) {
        hipLaunchKernelGGL((d3_kernel<T>), grid, dim3(256), 0, st, a.yin, a.dbuf, a.V, a.c, g, a.BC);
        hipLaunchKernelGGL((f3_kernel<T, 2>), grid, dim3(256), 0, st, a.yin, a.u, a.acc, a.yout, a.dbuf, a.V,
                           a.c, g, a.BC, a.stage, a.ulo);
    } else {
        hipLaunchKernelGGL((f3_kernel<T, 1>), grid, dim3(256), 0, st, a.yin, a.u, a.acc, a.yout,
                           (const c2<T>*)nullptr, a.V, a.c, g, a.BC, a.stage, a.ulo);
    }
    return hipGetLastError();
}

template hipError_t launch_stage3d<double>(const StageArgs<double>&, hipStream_t);
template hipError_t launch_stage3d<float>(const StageArgs<float>&, hipStream_t);

}  // namespace nlse
