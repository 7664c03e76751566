// nlse2d_batch.hip -- many independent 2D problems (uneven sizes, own s, a,
// h, k and V each) advanced by ONE persistent launch per fused group of RK4
// stages.  Each workgroup walks the concatenated tile list of all problems;
// the tile code is the single-problem one (nlse2d_tile.h, nlse2d_rows.h,
// nlse2d_strip.h), so every problem's result is bit-identical to running it
// alone.
#include "nlse2d_rows.h"
#include "nlse2d_strip.h"
#include "nlse2d_tile.h"

namespace nlse {

// A persistent grid walks the concatenated tile list of all problems: block b
// works on XCD b mod 8, which owns a contiguous eighth of the list, so the
// tiles of one problem (and their shared halos) stay in one L2.
template <typename T, int METHOD, int TI, int FIRST, int LAST>
__global__ __launch_bounds__(256) void batch2d_kernel(const BatchProb2<T>* __restrict__ probs,
                                                      const int* __restrict__ tile_prob, int ntiles, int r_yin,
                                                      int r_u, int r_out)
{
    using L = TileLds<T, METHOD, TI, FIRST, LAST>;
    __shared__ c2<T> sY[L::NY];
    __shared__ c2<T> sD[L::ND];
    const int xcd = blockIdx.x & 7, nper = gridDim.x >> 3;
    const int per = (ntiles + 7) >> 3;
    for (int i = blockIdx.x >> 3; i < per; i += nper) {
        const int t = xcd * per + i;
        if (t >= ntiles) break;
        const BatchProb2<T>& P = probs[tile_prob[t]];
        const int lt = t - P.tile0;
        const int ti = lt / P.tiles_j, tj = lt - ti * P.tiles_j;
        fused2d_tile<T, METHOD, TI, FIRST, LAST>(sY, sD, ti * TI - P.ioff, tj * TJ2 - P.joff, P.N, P.M,
                                                 P.buf[r_yin], P.buf[r_u], P.acc, P.buf[r_out], P.V, P.c, P.BC,
                                                 P.ulo);
    }
}

// the same walk over row-sweep tiles (tile width 64 - 2H per problem)
template <typename T, int METHOD, int TI, int FIRST, int LAST, bool COMP, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(4))) void batch2d_rows_kernel(const BatchProb2<T>* __restrict__ probs,
                                                           const int* __restrict__ tile_prob, int ntiles,
                                                           int r_yin, int r_u, int r_out)
{
    __shared__ RowsLds<T, METHOD, FIRST, LAST, TI> sm;
    constexpr int W = RowsCfg<METHOD, FIRST, LAST, TI>::W;
    const int xcd = blockIdx.x & 7, nper = gridDim.x >> 3;
    const int per = (ntiles + 7) >> 3;
    for (int i = blockIdx.x >> 3; i < per; i += nper) {
        const int t = xcd * per + i;
        if (t >= ntiles) break;
        const BatchProb2<T>& P = probs[tile_prob[t]];
        const int lt = t - P.tile0;
        const int ti = lt / P.tiles_j, tj = lt - ti * P.tiles_j;
        const int i0 = ti * TI - P.ioff, j0 = tj * W - P.joff;
        c2<T>* lo = COMP ? P.ulo : nullptr;
        if (rows_inner<METHOD, FIRST, LAST, TI>(i0, j0, P.N, P.M))
            rows2d_tile<T, METHOD, FIRST, LAST, TI, COMP, true, NW>(sm, i0, j0, P.N, P.M, P.buf[r_yin], P.buf[r_u],
                                                                    P.acc, P.buf[r_out], P.V, P.c, P.BC, lo);
        else
            rows2d_tile<T, METHOD, FIRST, LAST, TI, COMP, false, NW>(sm, i0, j0, P.N, P.M, P.buf[r_yin],
                                                                     P.buf[r_u], P.acc, P.buf[r_out], P.V, P.c,
                                                                     P.BC, lo);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// the same walk over register-marching strips (nlse2d_strip.h): each wave
// of a workgroup takes its own unit (a segment of S rows x one strip of
// 64 - 2H columns of one problem); waves never synchronise
template <typename T, int METHOD, int FIRST, int LAST, bool COMP>
__global__ __launch_bounds__(256) void batch2d_strip_kernel(const BatchProb2<T>* __restrict__ probs,
                                                            const int* __restrict__ tile_prob, int ntiles, int S,
                                                            int r_yin, int r_u, int r_out)
{
    constexpr int W = StripCfg<METHOD, FIRST, LAST>::W;
    const int xcd = blockIdx.x & 7, nper = gridDim.x >> 3;
    const int per = (ntiles + 7) >> 3;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = (blockIdx.x >> 3) * 4 + w; i < per; i += nper * 4) {
        const int t = xcd * per + i;
        if (t >= ntiles) break;
        const BatchProb2<T>& P = probs[tile_prob[t]];
        const int lt = t - P.tile0;
        const int ti = lt / P.tiles_j, tj = lt - ti * P.tiles_j;
        strip2d_wave<T, METHOD, FIRST, LAST, COMP>(ti * S - P.ioff, S, tj * W - P.joff, P.N, P.M, P.buf[r_yin],
                                                   P.buf[r_u], P.acc, P.buf[r_out], P.V, P.c, P.BC,
                                                   COMP ? P.ulo : nullptr);
    }
}

template <typename T, int METHOD, int FIRST, int LAST>
static hipError_t launch_b_strips(const BatchArgs<T>& a, hipStream_t st)
{
    if constexpr (std::is_same<T, float>::value && LAST == 4) {
        if (a.comp) {
            hipLaunchKernelGGL((batch2d_strip_kernel<T, METHOD, FIRST, LAST, true>), dim3(a.grid), dim3(256), 0, st,
                               a.probs, a.tile_prob, a.ntiles, a.tile, a.r_yin, a.r_u, a.r_out);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((batch2d_strip_kernel<T, METHOD, FIRST, LAST, false>), dim3(a.grid), dim3(256), 0, st,
                       a.probs, a.tile_prob, a.ntiles, a.tile, a.r_yin, a.r_u, a.r_out);
    return hipGetLastError();
}

template <typename T, int METHOD, int TI, int FIRST, int LAST>
static hipError_t launch_b_rows(const BatchArgs<T>& a, hipStream_t st)
{
    constexpr int NW = rows_waves<METHOD, FIRST, LAST, TI>();
    if constexpr (std::is_same<T, float>::value && LAST == 4) {
        if (a.comp) {
            hipLaunchKernelGGL((batch2d_rows_kernel<T, METHOD, TI, FIRST, LAST, true, NW>), dim3(a.grid),
                               dim3(NW * 64), 0, st, a.probs, a.tile_prob, a.ntiles, a.r_yin, a.r_u, a.r_out);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((batch2d_rows_kernel<T, METHOD, TI, FIRST, LAST, false, NW>), dim3(a.grid), dim3(NW * 64), 0,
                       st, a.probs, a.tile_prob, a.ntiles, a.r_yin, a.r_u, a.r_out);
    return hipGetLastError();
}

template <typename T, int METHOD, int TI, int FIRST, int LAST>
static hipError_t launch_b(const BatchArgs<T>& a, hipStream_t st)
{
    hipLaunchKernelGGL((batch2d_kernel<T, METHOD, TI, FIRST, LAST>), dim3(a.grid), dim3(256), 0, st, a.probs,
                       a.tile_prob, a.ntiles, a.r_yin, a.r_u, a.r_out);
    return hipGetLastError();
}

template <typename T, int METHOD, int FIRST, int LAST>
static hipError_t launch_b_t(const BatchArgs<T>& a, hipStream_t st)
{
    if (a.kernel == KERNEL_STRIPS) return launch_b_strips<T, METHOD, FIRST, LAST>(a, st);
    if (a.kernel == KERNEL_ROWS) {
        switch (a.tile) {
        case 12: return launch_b_rows<T, METHOD, 12, FIRST, LAST>(a, st);
        case 16: return launch_b_rows<T, METHOD, 16, FIRST, LAST>(a, st);
        default: return launch_b_rows<T, METHOD, 8, FIRST, LAST>(a, st);
        }
    }
    switch (a.tile) {
    case 8: return launch_b<T, METHOD, 8, FIRST, LAST>(a, st);
    case 32: return launch_b<T, METHOD, 32, FIRST, LAST>(a, st);
    default: return launch_b<T, METHOD, 16, FIRST, LAST>(a, st);
    }
}

template <typename T, int METHOD>
static hipError_t launch_b_m(const BatchArgs<T>& a, hipStream_t st)
{
    if (a.first == 1 && a.last == 4) return launch_b_t<T, METHOD, 1, 4>(a, st);
    if (a.first == 1 && a.last == 2) return launch_b_t<T, METHOD, 1, 2>(a, st);
    if (a.first == 3 && a.last == 4) return launch_b_t<T, METHOD, 3, 4>(a, st);
    switch (a.first) {
    case 1: return launch_b_t<T, METHOD, 1, 1>(a, st);
    case 2: return launch_b_t<T, METHOD, 2, 2>(a, st);
    case 3: return launch_b_t<T, METHOD, 3, 3>(a, st);
    default: return launch_b_t<T, METHOD, 4, 4>(a, st);
    }
}

template <typename T>
hipError_t launch_batch2d(const BatchArgs<T>& a, hipStream_t st)
{
    return a.method == 2 ? launch_b_m<T, 2>(a, st) : launch_b_m<T, 1>(a, st);
}

template hipError_t launch_batch2d<double>(const BatchArgs<double>&, hipStream_t);
template hipError_t launch_batch2d<float>(const BatchArgs<float>&, hipStream_t);

}  // namespace nlse
