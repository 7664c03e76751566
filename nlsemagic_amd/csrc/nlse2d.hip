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
#include "nlse_common.h"
#include "nlse_launch.h"

namespace nlse {

struct Grid2 {
    int N, M;
    int tiles_i, tiles_j;
    int ioff, joff;  // tile origins shift by one when n mod tile == 1
    TileMap map;
};

constexpr int TJ2 = 64;

template <int TI, int H, int R>
struct FusedGeom {
    static constexpr int YW = TJ2 + 2 * H, YH = TI + 2 * H;  // LDS tile of the stage input
    static constexpr int DE = H - R + 1;                     // halo of the widest D region
    static constexpr int DW = TJ2 + 2 * DE, DH = TI + 2 * DE;
    static constexpr int E1 = H - R;                         // F halo of the first fused stage
    static constexpr int RING1 = (TJ2 + 2 * E1) * (TI + 2 * E1) - TI * TJ2;
    static constexpr int RMAX = RING1 > 0 ? (RING1 + 255) / 256 : 1;
};

// (gi, gj) of ring element x of the region "tile +- e" minus the tile
__device__ __forceinline__ void ring_coord(int x, int e, int TI, int i0, int j0, int& gi, int& gj)
{
    const int W = TJ2 + 2 * e;
    const int band = e * W;
    if (x < band) {
        gi = i0 - e + x / W;
        gj = j0 - e + x % W;
    } else if (x < 2 * band) {
        x -= band;
        gi = i0 + TI + x / W;
        gj = j0 - e + x % W;
    } else if (x < 2 * band + e * TI) {
        x -= 2 * band;
        gi = i0 + x / e;
        gj = j0 - e + x % e;
    } else {
        x -= 2 * band + e * TI;
        gi = i0 + x / e;
        gj = j0 + TJ2 + x % e;
    }
}

template <typename T, int METHOD, int TI, int FIRST, int LAST>
__global__ __launch_bounds__(256) void fused2d_kernel(const c2<T>* __restrict__ yin, const c2<T>* u,
                                                      c2<T>* __restrict__ acc, c2<T>* out,
                                                      const T* __restrict__ V, Coefs<T> c, Grid2 g, int BC)
{
    constexpr int R = (METHOD == 2) ? 2 : 1;
    constexpr int NF = LAST - FIRST + 1;
    constexpr int H = NF * R;
    using G = FusedGeom<TI, H, R>;
    constexpr int YW = G::YW, YH = G::YH, DE = G::DE, DW = G::DW, DH = G::DH;
    constexpr int OPT = TI / 4;  // owned points per thread
    __shared__ c2<T> sY[YH * YW];
    __shared__ c2<T> sD[(METHOD == 2) ? DH * DW : 1];

    const int N = g.N, M = g.M;
    const int t = g.map.tile();
    if (t >= g.map.ntiles) return;
    const int ti = t / g.tiles_j, tj = t - ti * g.tiles_j;
    const int i0 = ti * TI - g.ioff, j0 = tj * TJ2 - g.joff;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;

    auto inside = [&](int gi, int gj) { return gi >= 0 && gi < N && gj >= 0 && gj < M; };
    auto Y = [&](int gi, int gj) -> c2<T> { return sY[(gi - i0 + H) * YW + (gj - j0 + H)]; };
    auto Ys = [&](int gi, int gj) -> c2<T>& { return sY[(gi - i0 + H) * YW + (gj - j0 + H)]; };
    auto D = [&](int gi, int gj) -> c2<T> { return sD[(gi - i0 + DE) * DW + (gj - j0 + DE)]; };
    auto Vat = [&](int gi, int gj) -> T { return V[(long)gi * M + gj]; };

    // owned points: V (and ktot / U when needed) are fetched first so their
    // latency overlaps the tile load and the D phase
    const int gj_own = j0 + lane;
    T Vo[OPT];
    c2<T> A[OPT], Uo[OPT];
#pragma unroll
    for (int r = 0; r < OPT; ++r) {
        const int gi = i0 + wv + 4 * r;
        const bool in = inside(gi, gj_own);
        const long p = (long)gi * M + gj_own;
        Vo[r] = in ? V[p] : (T)0;
        A[r] = (FIRST > 1 && in) ? acc[p] : mk<T>(0, 0);
        Uo[r] = (FIRST > 1 && in) ? u[p] : mk<T>(0, 0);
    }
    for (int e = tid; e < YH * YW; e += 256) {
        const int r = e / YW, q = e - r * YW;
        const int gi = i0 - H + r, gj = j0 - H + q;
        sY[e] = inside(gi, gj) ? yin[(long)gi * M + gj] : mk<T>(0, 0);
    }

    // D = Lap_CD(Y)/h^2 at an interior point (NLSEmagic2D_CPU.c:272-275)
    auto d_int = [&](int gi, int gj) -> c2<T> {
        const c2<T> a = Y(gi + 1, gj), b = Y(gi - 1, gj), m = Y(gi, gj), l = Y(gi, gj + 1), r = Y(gi, gj - 1);
        return mk<T>((a.x + b.x - 4 * m.x + l.x + r.x) * c.lh2, (a.y + b.y - 4 * m.y + l.y + r.y) * c.lh2);
    };
    // D with the boundary rules (NLSEmagic2D_CPU.c:279-375)
    auto d_at = [&](int gi, int gj) -> c2<T> {
        if (gi > 0 && gi < N - 1 && gj > 0 && gj < M - 1) return d_int(gi, gj);
        if (BC == BC_DIRICHLET) return dir_d<T>(Y(gi, gj), Vat(gi, gj), c);
        if (BC == BC_MSD) {
            const int qi = clamp_in(gi, N), qj = clamp_in(gj, M);
            return msd_d<T>(d_int(qi, qj), Y(qi, qj), Vat(qi, qj), Y(gi, gj), Vat(gi, gj), c);
        }
        return mk<T>(0, 0);
    };
    // F at an interior point: CD (NLSEmagic2D_CPU.c:37-41) or 2SHOC (:383-395)
    auto f_int = [&](int gi, int gj, T v) -> c2<T> {
        const c2<T> y = Y(gi, gj);
        const T n = nlin<T>(c.s, y, v);
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
    // F with the boundary rules (NLSEmagic2D_CPU.c:399-497)
    auto f_at = [&](int gi, int gj, T v) -> c2<T> {
        if (gi > 0 && gi < N - 1 && gj > 0 && gj < M - 1) return f_int(gi, gj, v);
        if (BC == BC_MSD) {
            const int qi = clamp_in(gi, N), qj = clamp_in(gj, M);
            return msd_f<T>(f_int(qi, qj, Vat(qi, qj)), Y(qi, qj), Y(gi, gj));
        }
        if (BC == BC_LAP0) return lap0_f<T>(Y(gi, gj), v, c.s);
        return mk<T>(0, 0);
    };

    __syncthreads();
    if (FIRST == 1) {  // stage 1 reads U = its own input from the tile
#pragma unroll
        for (int r = 0; r < OPT; ++r) {
            const int gi = i0 + wv + 4 * r;
            if (inside(gi, gj_own)) Uo[r] = Y(gi, gj_own);
        }
    }

#pragma unroll
    for (int s = FIRST; s <= LAST; ++s) {
        const int e = H - (s - FIRST + 1) * R;  // F halo of this stage
        constexpr int RMAX = G::RMAX;
        const int ring = (TJ2 + 2 * e) * (TI + 2 * e) - TI * TJ2;
        // ring points of this stage: fetch V and U before the D phase
        T Vr[RMAX];
        c2<T> Ur[RMAX];
        if (s < LAST) {
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                const int x = tid + 256 * r;
                Vr[r] = (T)0;
                Ur[r] = mk<T>(0, 0);
                if (x < ring) {
                    int gi, gj;
                    ring_coord(x, e, TI, i0, j0, gi, gj);
                    if (inside(gi, gj)) {
                        Vr[r] = Vat(gi, gj);
                        Ur[r] = (s == 1) ? Y(gi, gj) : u[(long)gi * M + gj];
                    }
                }
            }
        }
        if (METHOD == 2) {
            const int ed = e + 1, Wd = TJ2 + 2 * ed, Hd = TI + 2 * ed;
            for (int x = tid; x < Wd * Hd; x += 256) {
                const int r = x / Wd, q = x - r * Wd;
                const int gi = i0 - ed + r, gj = j0 - ed + q;
                if (inside(gi, gj)) sD[(gi - i0 + DE) * DW + (gj - j0 + DE)] = d_at(gi, gj);
            }
            __syncthreads();
        }
        // evaluate F on the tile (fixed owner per point) and on the ring
        c2<T> Fo[OPT];
#pragma unroll
        for (int r = 0; r < OPT; ++r) {
            const int gi = i0 + wv + 4 * r;
            Fo[r] = inside(gi, gj_own) ? f_at(gi, gj_own, Vo[r]) : mk<T>(0, 0);
        }
        c2<T> Fr[RMAX];
        if (s < LAST) {
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                const int x = tid + 256 * r;
                Fr[r] = mk<T>(0, 0);
                if (x < ring) {
                    int gi, gj;
                    ring_coord(x, e, TI, i0, j0, gi, gj);
                    if (inside(gi, gj)) Fr[r] = f_at(gi, gj, Vr[r]);
                }
            }
        }
        const T cs = (s == 3) ? c.k : c.k2;  // Y_s = U + cs K_s (stages 1-3)
        __syncthreads();  // every read of this stage's input is done
#pragma unroll
        for (int r = 0; r < OPT; ++r) {
            const int gi = i0 + wv + 4 * r;
            if (!inside(gi, gj_own)) continue;
            const c2<T> f = Fo[r];
            if (s == 1) {
                A[r] = f;
            } else if (s == 4) {
                A[r].x = A[r].x + (T)1.0 * f.x;
                A[r].y = A[r].y + (T)1.0 * f.y;
            } else {
                A[r].x = A[r].x + (T)2.0 * f.x;
                A[r].y = A[r].y + (T)2.0 * f.y;
            }
            const long p = (long)gi * M + gj_own;
            if (s < LAST) {
                Ys(gi, gj_own) = mk<T>(Uo[r].x + cs * f.x, Uo[r].y + cs * f.y);
            } else if (LAST == 4) {
                out[p] = mk<T>(Uo[r].x + c.k6 * A[r].x, Uo[r].y + c.k6 * A[r].y);
            } else {
                out[p] = mk<T>(Uo[r].x + cs * f.x, Uo[r].y + cs * f.y);
                acc[p] = A[r];
            }
        }
        if (s < LAST) {
#pragma unroll
            for (int r = 0; r < RMAX; ++r) {
                const int x = tid + 256 * r;
                if (x < ring) {
                    int gi, gj;
                    ring_coord(x, e, TI, i0, j0, gi, gj);
                    if (inside(gi, gj))
                        Ys(gi, gj) = mk<T>(Ur[r].x + cs * Fr[r].x, Ur[r].y + cs * Fr[r].y);
                }
            }
            __syncthreads();  // the next stage's input is in place
        }
    }
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
                       a.yin, a.u, a.acc, a.yout, a.V, a.c, g, a.BC);
    return hipGetLastError();
}

// tile heights offered to the tuner: 8, 16 (default) and 32 rows of 64 columns
template <typename T, int METHOD, int FIRST, int LAST>
static hipError_t launch_fused(const StageArgs<T>& a, hipStream_t st)
{
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
