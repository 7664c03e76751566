// nlse2d_rows.h -- row-sweep LDS tile for fused 2D RK4 stages (device code).
//
// A workgroup of NW (4 or 8) waves advances a tile of TI rows x W = 64 - 2H owned
// columns through RK4 stages FIRST..LAST (H = R per fused stage):
//   - lane l is column j0 - H + l, so one row of the extended tile (the owned
//     block plus H rows/columns of halo) is exactly one wave-wide access:
//     a coalesced 1 KB (FP64) load from HBM and conflict-free 16 B per lane
//     in LDS;
//   - extended row e belongs to wave e mod NW in every phase, which keeps the
//     row's Y, U, V and ktot in that wave's registers; the rows are loaded
//     all at once at the start, stage input first (every load in flight
//     together; rows outside the grid load a clamped copy, so every wave
//     issues the same loads);
//   - the rows i +- 1, the diagonal corners of the 2SHOC stencil and a row's
//     own j +- 1 neighbours come from the LDS copy of the stage input (lane
//     +- 1 of a row's copy costs no VALU moves, unlike DPP shifts of FP64
//     values), and for 2SHOC D = Lap_CD(Y)/h^2 of rows i +- 1 and columns
//     j +- 1 from an LDS copy of D;
//   - per stage: D phase, barrier, F phase (each row's F and RK4 update at
//     once, the new Y in the row's registers), barrier, new Y rows written
//     to LDS, barrier; the last stage writes its rows to HBM.
// MSD boundary rows take their partner row's D/F from LDS after the barrier
// that completes it (D: one extra barrier, only in tiles holding row 0 or
// N-1).  MSD boundary columns take the partner lane's F by DPP; their D is
// evaluated once per D phase rather than once per row: after the phase, one
// wave holds (row, side) pairs in its lanes and computes every boundary
// column point's D from the partner column's values in LDS (in-lane, each
// row would spend a full wave of instructions, a division among them, on two
// useful lanes); the MSD column's F partner comes by DPP from the adjacent
// lane.  Every owned point gets the reference's operations in its operand
// order.
#pragma once

#include "nlse2d_strip.h"

namespace nlse {

// keeps the compiler from interleaving the rows of a phase (which would hold
// every row's operands live at once and cut the waves per SIMD); latency is
// hidden by the other waves instead
#ifndef ROWS_SCHED_FENCE
#define ROWS_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#endif

// Materialise a value here: an empty asm that takes and returns it in its
// registers, so the optimiser cannot sink the arithmetic that produces it
// past a later barrier (which would keep every row's LDS operands live
// across the barrier).  No instruction is emitted.
template <typename T> __device__ __forceinline__ void pin_here(c2<T>& v)
{
    asm volatile("" : "+v"(v.x), "+v"(v.y));
}

template <int METHOD, int FIRST, int LAST, int TI, int NW = 4>
struct RowsCfg;

// waves per workgroup: tiles of more than 16 extended rows use 8 (half the
// rows per wave, twice the waves to hide latency), smaller ones 4
template <int METHOD, int FIRST, int LAST, int TI>
constexpr int rows_waves()
{
    return RowsCfg<METHOD, FIRST, LAST, TI>::TE > 16 ? 8 : 4;
}

template <int METHOD, int FIRST, int LAST, int TI, int NW>
struct RowsCfg {
    static constexpr int R = (METHOD == 2) ? 2 : 1;
    static constexpr int NF = LAST - FIRST + 1;
    static constexpr int H = NF * R;        // halo rows and columns
    static constexpr int W = 64 - 2 * H;    // owned columns
    static constexpr int TE = TI + 2 * H;   // extended rows
    static constexpr int RW = (TE + NW - 1) / NW;  // extended rows per wave
    static constexpr int ND = (METHOD == 2) ? TE - 2 : 1;  // D rows held in LDS (rows 1..TE-2)
};

template <typename T, int METHOD, int FIRST, int LAST, int TI>
struct RowsLds {
    using Cfg = RowsCfg<METHOD, FIRST, LAST, TI>;
    // one padding element before and after each array: lanes 0 and 63 read
    // their out-of-row neighbour (a value no owned point uses) in bounds
    // MSD: F of rows 1 and N-2 for the boundary rows goes to extended rows 0
    // and TE-1 of y, which no F phase reads
    c2<T> y[Cfg::TE * 64 + 2];
    c2<T> d[Cfg::ND * 64 + 2];
    // tiles holding column 0 or M-1 with MSD: V at columns 0, 1, M-2, M-1 of
    // every extended row, for the D column pass
    T bv[Cfg::TE][4];
};

// Tile with owned rows [i0, i0 + TI) and columns [j0, j0 + W) of an N x M
// problem (the block may overhang the grid by one row/column at the start
// and by any amount at the end).  Buffers as in fused2d_tile.  INNER: the
// extended tile lies inside the grid and holds no boundary row or column
// (rows_inner()), so every range below is a compile-time constant.
template <int METHOD, int FIRST, int LAST, int TI>
__device__ __forceinline__ bool rows_inner(int i0, int j0, int N, int M)
{
    using Cfg = RowsCfg<METHOD, FIRST, LAST, TI>;
    return i0 - Cfg::H >= 0 && i0 - Cfg::H + Cfg::TE <= N && j0 - Cfg::H >= 1 && j0 - Cfg::H + 64 <= M - 1;
}

template <typename T, int METHOD, int FIRST, int LAST, int TI, bool COMP, bool INNER, int NW = 4>
__device__ __forceinline__ void rows2d_tile(RowsLds<T, METHOD, FIRST, LAST, TI>& sm, int i0, int j0, int N, int M,
                                            const c2<T>* __restrict__ yin, const c2<T>* u, c2<T>* __restrict__ acc,
                                            c2<T>* out, const T* __restrict__ V, const Coefs<T>& c, int BC,
                                            c2<T>* __restrict__ ulo)
{
    using Cfg = RowsCfg<METHOD, FIRST, LAST, TI, NW>;
    constexpr int R = Cfg::R, NF = Cfg::NF, H = Cfg::H, W = Cfg::W, TE = Cfg::TE, RW = Cfg::RW;
    static_assert(TE <= 32, "the MSD column pass gives each extended row one lane of a half wave");
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __builtin_assume(w >= 0 && w < NW);
    const int j = j0 - H + lane;
    const bool jin = INNER || (j >= 0 && j < M);
    const bool own = jin && lane >= H && lane < H + W;
    const int jc = INNER ? j : (j < 0 ? 0 : (j >= M ? M - 1 : j));
    const bool colL = (j == 0), colR = (j == M - 1);
    const int lL = H - j0, lR = M - 1 - j0 + H;
    const bool colbc = !INNER && ((lL >= 0 && lL < 64) || (lR >= 0 && lR < 64));
    const int g0 = i0 - H;  // grid row of extended row 0
    // extended rows inside the grid: [eg0, eg1)
    const int eg0 = INNER ? 0 : -g0, eg1 = INNER ? TE : N - g0;
    const bool msd = (BC == BC_MSD);
    const bool top = !INNER && msd && eg0 >= 0 && eg0 < TE, bot = !INNER && msd && eg1 - 1 >= 0 && eg1 - 1 < TE;
    // MSD columns 0 / M-1 in this tile (uniform over the workgroup): their D
    // comes from the column pass after each D phase
    const bool colmsd = !INNER && msd && colbc;
    // a D/F row at grid row g is interior (the reference's full stencil)
    auto interior = [&](int g) { return INNER || (g > 0 && g < N - 1); };
    const c2<T> z = mk<T>(0, 0);

    auto sy = [&](int e, int dl) -> c2<T> { return sm.y[1 + e * 64 + lane + dl]; };
    auto sd = [&](int e, int dl) -> c2<T> { return sm.d[1 + (e - 1) * 64 + lane + dl]; };
    auto colq = [&](auto v) {
        const auto p = lane_shift2<DPP_NEXT, T>(v), m = lane_shift2<DPP_PREV, T>(v);
        return colL ? p : (colR ? m : v);
    };
    auto colq1 = [&](T v) {
        const T p = lane_shift<DPP_NEXT>(v), m = lane_shift<DPP_PREV>(v);
        return colL ? p : (colR ? m : v);
    };

    // rows of this wave: e = 4k + w; everything about row e stays in slot k
    c2<T> Yc[RW], Uc[RW], Ac[RW], Lc[RW];
    T Vc[RW];

    // phase 0: every load of the tile in flight at once, the stage input
    // first (the D phase needs nothing else, so it can start while V, U and
    // ktot are still arriving); no branches around the loads: a row outside
    // the range a field is used in reads a row inside it (out-of-grid rows
    // read clamped rows: values no owned point uses)
    auto rowp = [&](int e) -> long {
        const int g = g0 + e;
        const int gc = INNER ? g : (g < 0 ? 0 : (g >= N ? N - 1 : g));
        return (long)gc * M + jc;
    };
#pragma unroll
    for (int k = 0; k < RW; ++k) Yc[k] = yin[rowp(min(NW * k + w, TE - 1))];
#pragma unroll
    for (int k = 0; k < RW; ++k) Vc[k] = V[rowp(min(NW * k + w, TE - 1))];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const int e = NW * k + w;
        if constexpr (FIRST > 1) {  // the rows the first fused stage updates: R .. TE-R-1
            const long pf = rowp(e < R ? R : (e >= TE - R ? TE - R - 1 : e));
            Uc[k] = u[pf];
            Ac[k] = acc[pf];
        } else {
            Ac[k] = z;
        }
        if constexpr (COMP && LAST == 4) Lc[k] = ulo[rowp(e < H ? H : (e >= H + TI ? H + TI - 1 : e))];
    }
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const int e = NW * k + w;
        if constexpr (FIRST == 1) Uc[k] = Yc[k];
        if (e < TE) sm.y[1 + e * 64 + lane] = Yc[k];
        if (colmsd && e < TE) {
            if (lane == lL) sm.bv[e][0] = Vc[k];
            if (lane == lL + 1) sm.bv[e][1] = Vc[k];
            if (lane == lR - 1) sm.bv[e][2] = Vc[k];
            if (lane == lR) sm.bv[e][3] = Vc[k];
        }
    }
    __syncthreads();

#pragma unroll
    for (int q = 0; q < NF; ++q) {
        const int QS = FIRST + q;
        // F rows and D rows of this stage
        const int flo = INNER ? (q + 1) * R : max((q + 1) * R, eg0);
        const int fhi = INNER ? TE - (q + 1) * R : min(TE - (q + 1) * R, eg1);
        const int dlo = INNER ? (q + 1) * R - 1 : max((q + 1) * R - 1, eg0);
        const int dhi = INNER ? TE - (q + 1) * R + 1 : min(TE - (q + 1) * R + 1, eg1);
        c2<T> ypart[RW];  // MSD boundary rows: the partner row's Y, read before the Y rows change
#pragma unroll
        for (int k = 0; k < RW; ++k) ypart[k] = z;
        if constexpr (METHOD == 2) {
#pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int e = NW * k + w;
                if (e < dlo || e >= dhi) continue;
                const int g = g0 + e;
                const c2<T> yx = Yc[k];
                const T vx = Vc[k];
                c2<T> dd = z;
                if (interior(g)) {
                    const c2<T> a = sy(e + 1, 0), b = sy(e - 1, 0);
                    const c2<T> yp = sy(e, 1), ym = sy(e, -1);  // the row's k +- 1 neighbours from its LDS copy
                    dd = mk<T>((a.x + b.x - 4 * yx.x + yp.x + ym.x) * c.lh2, (a.y + b.y - 4 * yx.y + yp.y + ym.y) * c.lh2);
                    if (colbc && !msd) {  // MSD columns: the column pass below
                        c2<T> db = z;
                        if (BC == BC_DIRICHLET) db = dir_d<T>(yx, vx, c);
                        if (colL || colR) dd = db;
                    }
                } else if (BC == BC_DIRICHLET) {
                    dd = dir_d<T>(yx, vx, c);
                }
                sm.d[1 + (e - 1) * 64 + lane] = dd;
                ROWS_SCHED_FENCE();
            }
            __syncthreads();
            if (colmsd) {
                // the D column pass: lane l of wave 0 takes extended row l & 31
                // on side l >> 5 (0: column 0 at lane lL, partner column 1;
                // 1: column M-1 at lane lR, partner column M-2) of the interior
                // grid rows (rows 0 and N-1, corners included, belong to the
                // row pass)
                const int e = lane & 31, side = lane >> 5;
                const int lb = side ? lR : lL, lq = side ? lR - 1 : lL + 1;
                if (w == 0 && lb >= 0 && lb < 64 && e < TE && e >= dlo && e < dhi && interior(g0 + e))
                    sm.d[1 + (e - 1) * 64 + lb] = msd_d<T>(sd(e, lq - lane), sy(e, lq - lane), sm.bv[e][side ? 2 : 1],
                                                           sy(e, lb - lane), sm.bv[e][side ? 3 : 0], c);
                __syncthreads();
            }
            if (top || bot) {  // MSD rows 0 / N-1 from rows 1 / N-2 (uniform over the workgroup)
#pragma unroll
                for (int k = 0; k < RW; ++k) {
                    const int e = NW * k + w;
                    if (e < dlo || e >= dhi) continue;
                    const int g = g0 + e;
                    if (g != 0 && g != N - 1) continue;
                    const int ep = (g == 0) ? e + 1 : e - 1;  // partner row
                    const T vp = V[(long)(g == 0 ? 1 : N - 2) * M + jc];
                    sm.d[1 + (e - 1) * 64 + lane] = msd_d<T>(colq(sd(ep, 0)), colq(sy(ep, 0)), colq1(vp), Yc[k], Vc[k], c);
                }
                __syncthreads();
            }
        }
        // RK4 epilogue of row e (slot k): ktot and the next stage's input row
        // go to registers (their LDS copy follows the barrier), or the last
        // stage's result to HBM
        const T cs = (QS == 3) ? c.k : c.k2;
        auto epilogue = [&](int k, int e, c2<T> f) {
            c2<T> a;
            if (QS == 1) {
                a = f;
            } else {
                const T wt = (QS == 4) ? (T)1.0 : (T)2.0;
                a = mk<T>(Ac[k].x + wt * f.x, Ac[k].y + wt * f.y);
            }
            const c2<T> uo = Uc[k];
            if (q < NF - 1) {
                Ac[k] = a;
                Yc[k] = mk<T>(uo.x + cs * f.x, uo.y + cs * f.y);
                pin_here<T>(Ac[k]);
                pin_here<T>(Yc[k]);
            } else if (own) {  // the last stage's rows are the owned rows
                const long p = (long)(g0 + e) * M + j;
                if (QS == 4) {
                    if constexpr (COMP) {
                        c2<T> hi = uo, lo = Lc[k];
                        comp_add<T>(hi.x, lo.x, c.k6 * a.x);
                        comp_add<T>(hi.y, lo.y, c.k6 * a.y);
                        ulo[p] = lo;
                        out[p] = hi;
                    } else if constexpr (std::is_same<T, float>::value) {
                        // streaming stores: the next step reads the state
                        // from the Infinity Cache after the kernel boundary
                        // in any case (single precision CD 7-8 % faster on
                        // 1024^2; FP64 and the compensated state gain
                        // nothing, and keep plain stores)
                        st_stream<T>(&out[p], mk<T>(uo.x + c.k6 * a.x, uo.y + c.k6 * a.y));
                    } else {
                        out[p] = mk<T>(uo.x + c.k6 * a.x, uo.y + c.k6 * a.y);
                    }
                } else {
                    out[p] = mk<T>(uo.x + cs * f.x, uo.y + cs * f.y);
                    acc[p] = a;
                }
            }
        };
        // F phase: each row's F and epilogue at once (a row's own Y is only
        // read by its own wave); MSD boundary rows wait for the partner row's
        // F, which goes to LDS
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int e = NW * k + w;
            if (e < flo || e >= fhi) continue;
            const int g = g0 + e;
            const c2<T> yx = Yc[k];
            const T vx = Vc[k];
            c2<T> ff = z;
            if (interior(g)) {
                const T n = nlin<T>(c.s, yx, vx);
                if constexpr (METHOD == 2) {
                    const c2<T> dc = sd(e, 0), da = sd(e + 1, 0), db = sd(e - 1, 0);
                    const c2<T> dl = sd(e, 1), dr = sd(e, -1);
                    const c2<T> aa = sy(e + 1, 1), ab = sy(e + 1, -1), ba = sy(e - 1, 1), bb = sy(e - 1, -1);
                    ff = mk<T>(-c.a * dc.y + c.a1_12 * (da.y + db.y + dl.y + dr.y)
                                   - c.a1_6h2 * (aa.y + ab.y - 4 * yx.y + ba.y + bb.y) - n * yx.y,
                               c.a * dc.x - c.a1_12 * (da.x + db.x + dl.x + dr.x)
                                   + c.a1_6h2 * (aa.x + ab.x - 4 * yx.x + ba.x + bb.x) + n * yx.x);
                } else {
                    const c2<T> a = sy(e + 1, 0), b = sy(e - 1, 0);
                    const c2<T> l = sy(e, 1), rt = sy(e, -1);
                    ff = mk<T>(-c.ah2 * (a.y - 4 * yx.y + b.y + l.y + rt.y) - n * yx.y,
                               c.ah2 * (a.x - 4 * yx.x + b.x + l.x + rt.x) + n * yx.x);
                }
                if (colbc) {
                    c2<T> fb = z;
                    if (msd) fb = msd_f<T>(colq(ff), colq(yx), yx);
                    else if (BC == BC_LAP0) fb = lap0_f<T>(yx, vx, c.s);
                    if (colL || colR) ff = fb;
                }
                if (top && g == 1) sm.y[1 + lane] = ff;
                if (bot && g == N - 2) sm.y[1 + (TE - 1) * 64 + lane] = ff;
            } else if (BC == BC_LAP0) {
                ff = lap0_f<T>(yx, vx, c.s);
            } else if (msd) {  // row 0 or N-1: F after the barrier
                ypart[k] = sy(g == 0 ? e + 1 : e - 1, 0);
                continue;
            }
            epilogue(k, e, ff);
            ROWS_SCHED_FENCE();
        }
        // every read of this stage's Y and D rows is done (after the last
        // stage only the MSD boundary rows below still read LDS)
        if (q < NF - 1 || top || bot) __syncthreads();
        if (top || bot) {
#pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int e = NW * k + w;
                if (e < flo || e >= fhi) continue;
                const int g = g0 + e;
                if (g == 0) epilogue(k, e, msd_f<T>(colq(sm.y[1 + lane]), colq(ypart[k]), Yc[k]));
                else if (g == N - 1) epilogue(k, e, msd_f<T>(colq(sm.y[1 + (TE - 1) * 64 + lane]), colq(ypart[k]), Yc[k]));
            }
        }
        if (q < NF - 1) {
#pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int e = NW * k + w;
                if (e >= flo && e < fhi) sm.y[1 + e * 64 + lane] = Yc[k];
            }
        }
        if (q < NF - 1) __syncthreads();  // the next stage's input rows are in LDS
    }
}

}  // namespace nlse
