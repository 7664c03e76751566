// nlse2d.hip -- one fused RK4 stage of the 2D NLSE integrator on MI355X.
//
// Replaces, per stage, the reference's compute_F_CD / compute_F_2SHOC plus the
// add_matrix passes around it (src/NLSEmagic2D/NLSEmagic2D_CPU.c:23-498 and
// the chunk loop :796-825).  One launch evaluates F on a 64 x TI tile (with
// its boundary values) and applies the stage's RK4 update:
//   - the stage input Y is staged in LDS with a halo of 1 (CD) or 2 (2SHOC);
//   - for 2SHOC, D = Lap_CD(Y) is computed into LDS on the tile plus a ring
//     of one, so D never touches HBM;
//   - boundary points follow the reference BC rules; an MSD point's partner
//     (each coordinate clamped into [1, n-2]) is re-evaluated in-thread,
//     which gives the same bits as the partner's own thread;
//   - HBM traffic per point and stage: read Y, V, U, ktot; write ktot, Y'.
// Arithmetic follows the reference operand order with FP contraction off.
//
// Layout: interleaved complex, element (i, j) at i*M + j; i < N is the
// reference's x index (its second MATLAB dimension), j < M is contiguous.
#include "nlse_common.h"
#include "nlse_launch.h"

namespace nlse {

struct Grid2 {
    int N, M;
    int tiles_i, tiles_j;
    int ioff, joff;  // tile origins shift by one when n mod tile == 1
    TileMap map;
};

template <typename T, int METHOD, int TI>
__global__ __launch_bounds__(256) void stage2d_kernel(const c2<T>* __restrict__ yin,
                                                      const c2<T>* u, c2<T>* __restrict__ acc,
                                                      c2<T>* yout, const T* __restrict__ V,
                                                      Coefs<T> c, Grid2 g, int BC, int stage)
{
    constexpr int TJ = 64;
    constexpr int R = (METHOD == 2) ? 2 : 1;
    constexpr int LW = TJ + 2 * R, LH = TI + 2 * R;
    constexpr int DW = TJ + 2, DH = TI + 2;
    __shared__ c2<T> sY[LH * LW];
    __shared__ c2<T> sD[(METHOD == 2) ? DH * DW : 1];

    const int N = g.N, M = g.M;
    const int t = g.map.tile();
    if (t >= g.map.ntiles) return;
    const int ti = t / g.tiles_j, tj = t - ti * g.tiles_j;
    const int i0 = ti * TI - g.ioff, j0 = tj * TJ - g.joff;
    const int tid = threadIdx.x;

    for (int e = tid; e < LH * LW; e += 256) {
        const int r = e / LW, q = e - r * LW;
        const int gi = i0 - R + r, gj = j0 - R + q;
        c2<T> v = mk<T>(0, 0);
        if (gi >= 0 && gi < N && gj >= 0 && gj < M) v = yin[(long)gi * M + gj];
        sY[e] = v;
    }
    __syncthreads();

    auto Y = [&](int gi, int gj) -> c2<T> { return sY[(gi - i0 + R) * LW + (gj - j0 + R)]; };
    auto D = [&](int gi, int gj) -> c2<T> { return sD[(gi - i0 + 1) * DW + (gj - j0 + 1)]; };

    // D = Lap_CD(Y)/h^2 at an interior point (NLSEmagic2D_CPU.c:272-275)
    auto d_int = [&](int gi, int gj) -> c2<T> {
        c2<T> a = Y(gi + 1, gj), b = Y(gi - 1, gj), m = Y(gi, gj), l = Y(gi, gj + 1), r = Y(gi, gj - 1);
        return mk<T>((a.x + b.x - 4 * m.x + l.x + r.x) * c.lh2, (a.y + b.y - 4 * m.y + l.y + r.y) * c.lh2);
    };

    if (METHOD == 2) {
        for (int e = tid; e < DH * DW; e += 256) {
            const int r = e / DW, q = e - r * DW;
            const int gi = i0 - 1 + r, gj = j0 - 1 + q;
            if (gi < 0 || gi >= N || gj < 0 || gj >= M) continue;
            c2<T> d;
            if (gi > 0 && gi < N - 1 && gj > 0 && gj < M - 1) {
                d = d_int(gi, gj);
            } else if (BC == BC_DIRICHLET) {
                d = dir_d<T>(Y(gi, gj), V[(long)gi * M + gj], c);
            } else if (BC == BC_MSD) {
                const int qi = clamp_in(gi, N), qj = clamp_in(gj, M);
                d = msd_d<T>(d_int(qi, qj), Y(qi, qj), V[(long)qi * M + qj], Y(gi, gj),
                             V[(long)gi * M + gj], c);
            } else {
                d = mk<T>(0, 0);
            }
            sD[e] = d;
        }
        __syncthreads();
    }

    // F at an interior point: CD (NLSEmagic2D_CPU.c:37-41) or 2SHOC (:383-395)
    auto f_int = [&](int gi, int gj) -> c2<T> {
        const c2<T> y = Y(gi, gj);
        const T n = nlin<T>(c.s, y, V[(long)gi * M + gj]);
        if (METHOD == 1) {
            const c2<T> a = Y(gi + 1, gj), b = Y(gi - 1, gj), l = Y(gi, gj + 1), r = Y(gi, gj - 1);
            return mk<T>(-c.ah2 * (a.y - 4 * y.y + b.y + l.y + r.y) - n * y.y,
                         c.ah2 * (a.x - 4 * y.x + b.x + l.x + r.x) + n * y.x);
        } else {
            const c2<T> dc = D(gi, gj), da = D(gi + 1, gj), db = D(gi - 1, gj), dl = D(gi, gj + 1),
                        dr = D(gi, gj - 1);
            const c2<T> aa = Y(gi + 1, gj + 1), ab = Y(gi + 1, gj - 1), ba = Y(gi - 1, gj + 1),
                        bb = Y(gi - 1, gj - 1);
            return mk<T>(-c.a * dc.y + c.a1_12 * (da.y + db.y + dl.y + dr.y)
                             - c.a1_6h2 * (aa.y + ab.y - 4 * y.y + ba.y + bb.y) - n * y.y,
                         c.a * dc.x - c.a1_12 * (da.x + db.x + dl.x + dr.x)
                             + c.a1_6h2 * (aa.x + ab.x - 4 * y.x + ba.x + bb.x) + n * y.x);
        }
    };

    const int lane = tid & 63, w = tid >> 6;
    const int gj = j0 + lane;
    if (gj < 0 || gj >= M) return;
#pragma unroll
    for (int rr = 0; rr < TI / 4; ++rr) {
        const int gi = i0 + w + 4 * rr;
        if (gi < 0 || gi >= N) continue;
        c2<T> f;
        if (gi > 0 && gi < N - 1 && gj > 0 && gj < M - 1) {
            f = f_int(gi, gj);
        } else if (BC == BC_MSD) {
            const int qi = clamp_in(gi, N), qj = clamp_in(gj, M);
            f = msd_f<T>(f_int(qi, qj), Y(qi, qj), Y(gi, gj));
        } else if (BC == BC_LAP0) {
            f = lap0_f<T>(Y(gi, gj), V[(long)gi * M + gj], c.s);
        } else {
            f = mk<T>(0, 0);
        }
        const long p = (long)gi * M + gj;
        const c2<T> uu = (stage == 1) ? Y(gi, gj) : u[p];
        rk4_epilogue<T>(stage, c, f, uu, acc, yout, yout, p);
    }
}

template <typename T, int METHOD>
hipError_t launch_stage2d_t(const StageArgs<T>& a, hipStream_t st)
{
    constexpr int TI = 16;
    Grid2 g;
    g.N = a.N;
    g.M = a.M;
    g.ioff = (a.N % TI == 1) ? 1 : 0;
    g.joff = (a.M % 64 == 1) ? 1 : 0;
    g.tiles_i = (a.N + g.ioff + TI - 1) / TI;
    g.tiles_j = (a.M + g.joff + 63) / 64;
    g.map.ntiles = g.tiles_i * g.tiles_j;
    g.map.per_xcd = (g.map.ntiles + 7) / 8;
    hipLaunchKernelGGL((stage2d_kernel<T, METHOD, TI>), dim3(8 * g.map.per_xcd), dim3(256), 0, st,
                       a.yin, a.u, a.acc, a.yout, a.V, a.c, g, a.BC, a.stage);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_stage2d(const StageArgs<T>& a, hipStream_t st)
{
    return a.method == 2 ? launch_stage2d_t<T, 2>(a, st) : launch_stage2d_t<T, 1>(a, st);
}

template hipError_t launch_stage2d<double>(const StageArgs<double>&, hipStream_t);
template hipError_t launch_stage2d<float>(const StageArgs<float>&, hipStream_t);

}  // namespace nlse
