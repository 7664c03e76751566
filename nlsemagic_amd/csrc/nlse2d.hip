// nlse2d.hip -- fused RK4 stages of the 2D NLSE integrator on MI355X.
//
// Replaces the reference's compute_F_CD / compute_F_2SHOC and the add_matrix
// passes around them (src/NLSEmagic2D/NLSEmagic2D_CPU.c:23-498 and the chunk
// loop :796-825).  One launch advances a 64 x TI tile through RK4 stages
// FIRST..LAST entirely on chip:
//   - the input of stage FIRST is staged in LDS with a halo of R per fused
//     stage (R = 1 for CD, 2 for 2SHOC), so stage s can be evaluated on the
//     tile plus a ring that shrinks by R per stage;
//   - for 2SHOC, D = Lap_CD(Y)/h^2 is computed into LDS on the stage's region
//     plus one, so D never touches HBM;
//   - each stage's new Y = U + c K is written back into the LDS tile in place
//     (between barriers), and ktot stays in registers for the owned points;
//   - boundary points follow the reference BC rules; an MSD point's partner
//     (each coordinate clamped into [1, n-2]) is re-evaluated in-thread,
//     which gives the same bits as the partner's own evaluation.
// Per point, every stage performs exactly the reference's operations in its
// operand order (FP contraction off): the results are bit-identical to the
// CPU integrator whatever the fusion depth.  Fusing trades redundant halo
// compute for HBM traffic: per RK4 step a point moves 304 B (4 launches),
// 128 B (two launches of two stages) or 40 B (one launch of all four).
//
// Layout: interleaved complex, element (i, j) at i*M + j; i < N is the
// reference's x index (its second MATLAB dimension), j < M is contiguous.
#include "nlse2d_rows.h"
#include "nlse2d_strip.h"
#include "nlse2d_tile.h"

namespace nlse {

template <typename T, int METHOD, int TI, int FIRST, int LAST>
__global__ __launch_bounds__(256) void fused2d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                      c2<T>* __restrict__ acc, c2<T>* out,
                                                      const T* __restrict__ V, Coefs<T> c, Grid2 g, int BC,
                                                      c2<T>* __restrict__ ulo)
{
    using L = TileLds<T, METHOD, TI, FIRST, LAST>;
    __shared__ c2<T> sY[L::NY];
    __shared__ c2<T> sD[L::ND];
    const int t = g.map.tile();
    if (t >= g.map.ntiles) return;
    const int ti = t / g.tiles_j, tj = t - ti * g.tiles_j;
    fused2d_tile<T, METHOD, TI, FIRST, LAST>(sY, sD, ti * TI - g.ioff, tj * TJ2 - g.joff, g.N, g.M, yin, u, acc,
                                             out, V, c, BC, ulo);
}

template <typename T, int METHOD, int TI, int FIRST, int LAST>
static hipError_t launch_fused_ti(const StageArgs<T>& a, hipStream_t st)
{
    Grid2 g;
    g.N = a.N;
    g.M = a.M;
    g.ioff = (a.N % TI == 1) ? 1 : 0;
    g.joff = (a.M % TJ2 == 1) ? 1 : 0;
    g.tiles_i = (a.N + g.ioff + TI - 1) / TI;
    g.tiles_j = (a.M + g.joff + TJ2 - 1) / TJ2;
    g.map.ntiles = g.tiles_i * g.tiles_j;
    g.map.per_xcd = (g.map.ntiles + 7) / 8;
    hipLaunchKernelGGL((fused2d_kernel<T, METHOD, TI, FIRST, LAST>), dim3(8 * g.map.per_xcd), dim3(256), 0, st,
                       a.yin, a.u, a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
    return hipGetLastError();
}

// Register-marching variant (nlse2d_strip.h): 4 independent waves per
// workgroup, one strip x segment unit each.
template <typename T, int METHOD, int FIRST, int LAST, bool COMP>
__global__ __launch_bounds__(256) void strip2d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                      c2<T>* __restrict__ acc, c2<T>* out,
                                                      const T* __restrict__ V, Coefs<T> c, Strip2 g, int BC,
                                                      c2<T>* __restrict__ ulo)
{
    constexpr int W = StripCfg<METHOD, FIRST, LAST>::W;
    const int unit = g.unit();
    if (unit >= g.nunit) return;
    const int seg = unit / g.nstrip, strip = unit - seg * g.nstrip;
    strip2d_wave<T, METHOD, FIRST, LAST, COMP>(seg * g.S - g.ioff, g.S, strip * W - g.joff, g.N, g.M, yin, u, acc, out,
                                         V, c, BC, ulo);
}

template <typename T, int METHOD, int FIRST, int LAST>
static hipError_t launch_strip(const StageArgs<T>& a, hipStream_t st)
{
    const Strip2 g = make_strip2(a.N, a.M, a.tile, StripCfg<METHOD, FIRST, LAST>::W);
    if constexpr (std::is_same<T, float>::value && LAST == 4) {  // compensated state: the low part too
        if (a.ulo) {
            hipLaunchKernelGGL((strip2d_kernel<T, METHOD, FIRST, LAST, true>), dim3(8 * g.per_xcd), dim3(256), 0,
                               st, a.yin, a.u, a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((strip2d_kernel<T, METHOD, FIRST, LAST, false>), dim3(8 * g.per_xcd), dim3(256), 0, st, a.yin,
                       a.u, a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
    return hipGetLastError();
}

// Row-sweep LDS tiles (nlse2d_rows.h): one workgroup per tile of TI rows.
// At most 128 VGPRs (four waves per SIMD): the FP64 stage-3-4 tile would take
// 132 and run three waves per SIMD, slower than with four waves and a few
// registers spilled.
template <typename T, int METHOD, int FIRST, int LAST, int TI, bool COMP, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(4))) void rows2d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                     c2<T>* __restrict__ acc, c2<T>* out, const T* __restrict__ V,
                                                     Coefs<T> c, Grid2 g, int BC, c2<T>* __restrict__ ulo)
{
    __shared__ RowsLds<T, METHOD, FIRST, LAST, TI> sm;
    constexpr int W = RowsCfg<METHOD, FIRST, LAST, TI>::W;
    int ti, tj;
    if (!g.edge_first(blockIdx.x, ti, tj)) return;
    const int i0 = ti * TI - g.ioff, j0 = tj * W - g.joff;
    if (rows_inner<METHOD, FIRST, LAST, TI>(i0, j0, g.N, g.M))
        rows2d_tile<T, METHOD, FIRST, LAST, TI, COMP, true, NW>(sm, i0, j0, g.N, g.M, yin, u, acc, out, V, c, BC, ulo);
    else
        rows2d_tile<T, METHOD, FIRST, LAST, TI, COMP, false, NW>(sm, i0, j0, g.N, g.M, yin, u, acc, out, V, c, BC,
                                                                 ulo);
}

template <typename T, int METHOD, int FIRST, int LAST, int TI>
static hipError_t launch_rows_ti(const StageArgs<T>& a, hipStream_t st)
{
    constexpr int W = RowsCfg<METHOD, FIRST, LAST, TI>::W;
    Grid2 g;
    g.N = a.N;
    g.M = a.M;
    g.ioff = (a.N % TI == 1) ? 1 : 0;
    g.joff = (a.M % W == 1) ? 1 : 0;
    g.tiles_i = (a.N + g.ioff + TI - 1) / TI;
    g.tiles_j = (a.M + g.joff + W - 1) / W;
    g.map.ntiles = g.tiles_i * g.tiles_j;
    g.map.per_xcd = (g.map.ntiles + 7) / 8;
    const int nwg = g.edge_first_setup();
    constexpr int NW = rows_waves<METHOD, FIRST, LAST, TI>();
    if constexpr (std::is_same<T, float>::value && LAST == 4) {
        if (a.ulo) {
            hipLaunchKernelGGL((rows2d_kernel<T, METHOD, FIRST, LAST, TI, true, NW>), dim3(nwg), dim3(NW * 64), 0,
                               st, a.yin, a.u, a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((rows2d_kernel<T, METHOD, FIRST, LAST, TI, false, NW>), dim3(nwg), dim3(NW * 64), 0, st,
                       a.yin, a.u, a.acc, a.yout, a.V, a.c, g, a.BC, a.ulo);
    return hipGetLastError();
}

template <typename T, int METHOD, int FIRST, int LAST>
static hipError_t launch_rows(const StageArgs<T>& a, hipStream_t st)
{
    switch (a.tile) {
    case 12: return launch_rows_ti<T, METHOD, FIRST, LAST, 12>(a, st);
    case 16: return launch_rows_ti<T, METHOD, FIRST, LAST, 16>(a, st);
    default: return launch_rows_ti<T, METHOD, FIRST, LAST, 8>(a, st);
    }
}

// the kernel family and its rows per tile (KERNEL_* in nlse_launch.h)
template <typename T, int METHOD, int FIRST, int LAST>
static hipError_t launch_fused(const StageArgs<T>& a, hipStream_t st)
{
    if (a.kernel == KERNEL_STRIPS) return launch_strip<T, METHOD, FIRST, LAST>(a, st);
    if (a.kernel == KERNEL_ROWS) return launch_rows<T, METHOD, FIRST, LAST>(a, st);
    switch (a.tile) {
    case 8: return launch_fused_ti<T, METHOD, 8, FIRST, LAST>(a, st);
    case 32: return launch_fused_ti<T, METHOD, 32, FIRST, LAST>(a, st);
    default: return launch_fused_ti<T, METHOD, 16, FIRST, LAST>(a, st);
    }
}

template <typename T, int METHOD>
static hipError_t launch_stage2d_m(const StageArgs<T>& a, hipStream_t st)
{
    switch (a.stage) {
    case 1: return launch_fused<T, METHOD, 1, 1>(a, st);
    case 2: return launch_fused<T, METHOD, 2, 2>(a, st);
    case 3: return launch_fused<T, METHOD, 3, 3>(a, st);
    default: return launch_fused<T, METHOD, 4, 4>(a, st);
    }
}

template <typename T>
hipError_t launch_stage2d(const StageArgs<T>& a, hipStream_t st)
{
    return a.method == 2 ? launch_stage2d_m<T, 2>(a, st) : launch_stage2d_m<T, 1>(a, st);
}

// stages 1-2 (first = true) or 3-4 of one RK4 step in one launch
template <typename T>
hipError_t launch_pair2d(const StageArgs<T>& a, bool first, hipStream_t st)
{
    if (a.method == 2) return first ? launch_fused<T, 2, 1, 2>(a, st) : launch_fused<T, 2, 3, 4>(a, st);
    return first ? launch_fused<T, 1, 1, 2>(a, st) : launch_fused<T, 1, 3, 4>(a, st);
}

// a whole RK4 step in one launch (yin == u, yout != u)
template <typename T>
hipError_t launch_step2d(const StageArgs<T>& a, hipStream_t st)
{
    if (a.method == 2) return launch_fused<T, 2, 1, 4>(a, st);
    return launch_fused<T, 1, 1, 4>(a, st);
}

template hipError_t launch_stage2d<double>(const StageArgs<double>&, hipStream_t);
template hipError_t launch_stage2d<float>(const StageArgs<float>&, hipStream_t);
template hipError_t launch_pair2d<double>(const StageArgs<double>&, bool, hipStream_t);
template hipError_t launch_pair2d<float>(const StageArgs<float>&, bool, hipStream_t);
template hipError_t launch_step2d<double>(const StageArgs<double>&, hipStream_t);
template hipError_t launch_step2d<float>(const StageArgs<float>&, hipStream_t);

}  // namespace nlse
