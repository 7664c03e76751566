// nlse2d_tile.h -- the 2D fused RK4 tile (device code shared by the single-
// problem kernels in nlse2d.hip and the batched kernels in nlse2d_batch.hip).
#pragma once

#include "nlse_common.h"
#include "nlse_launch.h"

namespace nlse {


struct Grid2 {
    int N, M;
    int tiles_i, tiles_j;
    int ioff, joff;  // tile origins shift by one when n mod tile == 1
    TileMap map;
    // edge-first order (row-sweep tiles): the tiles of the first and last
    // tile row and column (which hold the boundary conditions and take
    // about twice as long) are dealt out first, so that the launch does not
    // end on them; then the interior tiles, each XCD a contiguous block
    int nb, nbp;          // edge tiles; nb rounded up to a multiple of 8
    int ninner, per_xcd;  // interior tiles; per XCD
    __device__ __forceinline__ bool edge_first(int b, int& ti, int& tj) const
    {
        if (b < nbp) {
            if (b >= nb) return false;
            if (ninner == 0) {
                ti = b / tiles_j;
                tj = b - ti * tiles_j;
            } else if (b < 2 * tiles_j) {
                ti = b < tiles_j ? 0 : tiles_i - 1;
                tj = b < tiles_j ? b : b - tiles_j;
            } else {
                const int r = b - 2 * tiles_j;
                ti = 1 + (r >> 1);
                tj = (r & 1) ? tiles_j - 1 : 0;
            }
            return true;
        }
        const int bb = b - nbp, l = bb >> 3;
        const int u = (bb & 7) * per_xcd + l;  // block b runs on XCD b mod 8 (nbp is a multiple of 8)
        if (l >= per_xcd || u >= ninner) return false;
        const int wi = tiles_j - 2, r = u / wi;
        ti = 1 + r;
        tj = 1 + (u - r * wi);
        return true;
    }
    // workgroups of an edge-first launch
    int edge_first_setup()
    {
        if (tiles_i <= 2 || tiles_j <= 2) {
            nb = tiles_i * tiles_j;
            ninner = 0;
        } else {
            nb = 2 * tiles_j + 2 * (tiles_i - 2);
            ninner = (tiles_i - 2) * (tiles_j - 2);
        }
        nbp = (nb + 7) / 8 * 8;
        per_xcd = (ninner + 7) / 8;
        return nbp + 8 * per_xcd;
    }
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

// LDS footprint of one tile
template <typename T, int METHOD, int TI, int FIRST, int LAST>
struct TileLds {
    static constexpr int R = (METHOD == 2) ? 2 : 1;
    static constexpr int H = (LAST - FIRST + 1) * R;
    using G = FusedGeom<TI, H, R>;
    static constexpr int NY = G::YH * G::YW;
    static constexpr int ND = (METHOD == 2) ? G::DH * G::DW : 1;
};

// One 64 x TI tile with origin (i0, j0) of an N x M problem through RK4
// stages FIRST..LAST.  `yin` is the input of stage FIRST (U when FIRST == 1),
// `u` the state at the start of the step, `acc` ktot (read when FIRST > 1,
// written when LAST < 4), `out` Y_LAST (LAST < 4) or the new U (LAST == 4).
// `ulo` (single precision, or null): low part of a compensated state, used by
// stage 4.  All threads of the block must call this (it synchronises).
template <typename T, int METHOD, int TI, int FIRST, int LAST>
__device__ __forceinline__ void fused2d_tile(c2<T>* __restrict__ sY, c2<T>* __restrict__ sD, int i0, int j0, int N,
                                             int M, const c2<T>* __restrict__ yin, const c2<T>* u,
                                             c2<T>* __restrict__ acc, c2<T>* out, const T* __restrict__ V,
                                             const Coefs<T>& c, int BC, c2<T>* __restrict__ ulo)
{
    constexpr int R = (METHOD == 2) ? 2 : 1;
    constexpr int NF = LAST - FIRST + 1;
    constexpr int H = NF * R;
    using G = FusedGeom<TI, H, R>;
    constexpr int YW = G::YW, YH = G::YH, DE = G::DE, DW = G::DW;
    constexpr int OPT = TI / 4;  // owned points per thread
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
                if (ulo) {
                    c2<T> hi = Uo[r], lo = ulo[p];
                    comp_add<T>(hi.x, lo.x, c.k6 * A[r].x);
                    comp_add<T>(hi.y, lo.y, c.k6 * A[r].y);
                    ulo[p] = lo;
                    out[p] = hi;
                } else {
                    out[p] = mk<T>(Uo[r].x + c.k6 * A[r].x, Uo[r].y + c.k6 * A[r].y);
                }
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

}  // namespace nlse
