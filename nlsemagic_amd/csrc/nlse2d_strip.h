// nlse2d_strip.h -- register-marching 2D RK4 stages (device code).
//
// One 64-lane wavefront owns a strip of 64 - 2*NF*R columns (lane = column,
// j contiguous in HBM) and a segment of S rows, and marches down the rows:
//   - each row of the stage input arrives once (a coalesced 1 KB load for
//     FP64) and lives in a 4-row register window;
//   - j +- 1 neighbours come from the adjacent lanes through DPP wave shifts
//     (wave_shl:1 / wave_shr:1), i +- 1 neighbours from the window, so the
//     kernel uses no LDS and no barrier;
//   - for 2SHOC, D = Lap_CD(Y)/h^2 of row r-1 is formed when row r arrives,
//     F of row r-2 one step later, and the RK4 epilogue of row r-3 after
//     that (the extra row lets the top MSD row take its partner's F);
//   - stage q+1 of a fused group (1, 2 or 4 stages) consumes the rows stage
//     q emits in the same sweep, three (2SHOC) or two (CD) rows behind, so
//     the intermediate Y never leaves the registers; each stage's rings are
//     written three sweep steps before their last read and have four slots,
//     whatever the number of stages.
// Rows and columns outside the owned block are computed as halo (R per fused
// stage); lanes whose inputs reach beyond the wave hold garbage that no owned
// value depends on.  Boundary rows/columns follow the reference BC rules
// with the MSD partner taken from the neighbouring row (pipeline) or lane
// (DPP); the partner values are the interior values, so every owned point
// gets exactly the reference's operations in its operand order.
#pragma once

#include <type_traits>

#include "nlse_common.h"
#include "nlse_launch.h"

namespace nlse {


template <int METHOD, int FIRST, int LAST>
struct StripCfg {
    static constexpr int R = (METHOD == 2) ? 2 : 1;  // stencil radius of one stage in rows/columns
    static constexpr int NF = LAST - FIRST + 1;       // fused stages
    static constexpr int HJ = NF * R;                 // input halo (columns and rows)
    static constexpr int W = 64 - 2 * HJ;             // owned columns per strip
    static constexpr int DLY = R + 1;                 // rows between a stage's input row and output row
    static_assert(NF == 1 || NF == 2 || NF == 4, "one, two or four fused stages");
};

// One wave: RK4 stages FIRST..LAST on rows [i0, i0+S) x columns [j0, j0+W)
// of an N x M problem (i0 may be -1 and the block may overhang the grid).
// Buffers as in fused2d_tile: `yin` is the input of stage FIRST (== u when
// FIRST == 1), `acc` ktot, `out` Y_LAST or the new U, `ulo` the low part of a
// compensated state (COMP; unused otherwise).
template <typename T, int METHOD, int FIRST, int LAST, bool COMP>
__device__ __forceinline__ void strip2d_wave(int i0, int S, int j0, int N, int M, const c2<T>* __restrict__ yin,
                                             const c2<T>* u, c2<T>* __restrict__ acc, c2<T>* out,
                                             const T* __restrict__ V, const Coefs<T>& c, int BC,
                                             c2<T>* __restrict__ ulo)
{
    using Cfg = StripCfg<METHOD, FIRST, LAST>;
    constexpr int R = Cfg::R, NF = Cfg::NF, HJ = Cfg::HJ, W = Cfg::W, DLY = Cfg::DLY;
    const int lane = threadIdx.x & 63;
    const int j = j0 - HJ + lane;
    const bool jin = j >= 0 && j < M;
    const bool own = jin && lane >= HJ && lane < HJ + W;
    // lanes outside the grid load the nearest column (always in bounds); no
    // owned value depends on what they hold
    const int jc = j < 0 ? 0 : (j >= M ? M - 1 : j);
    const bool colL = (j == 0), colR = (j == M - 1);
    const int lL = HJ - j0, lR = M - 1 - j0 + HJ;  // lanes of columns 0 and M-1
    const bool colbc = (lL >= 0 && lL < 64) || (lR >= 0 && lR < 64);
    const int ibeg = i0 > 0 ? i0 : 0;
    const int iend = (i0 + S < N) ? i0 + S : N;
    const int base = ibeg - HJ;  // stage-1 input row of sweep step 0

    // rows of F (the stage's output), D and input of each fused stage
    int Flo[NF], Fhi[NF], Dlo[NF], Dhi[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) {
        Flo[q] = max(0, ibeg - R * (NF - 1 - q));
        Fhi[q] = min(N, iend + R * (NF - 1 - q));
        Dlo[q] = max(0, Flo[q] - 1);
        Dhi[q] = min(N, Fhi[q] + 1);
    }
    const int Ilo = max(0, Flo[0] - R), Ihi = min(N, Fhi[0] + R);
    const int nsweep = iend - base + NF * DLY;

    const c2<T> z = mk<T>(0, 0);
    // HBM rows of the stage-1 input and V: row base + t in slot t & 7, loaded
    // PF = 4 sweep steps ahead; stage 1 reads rows t .. t-3 of them
    constexpr int PF = 4;
    c2<T> gy[8];
    T gv[8];
    // U and ktot (FIRST > 1) of row base + t, loaded at step t for stage 1's
    // epilogue at step t + DLY; compensated low part of the last stage's row
    c2<T> gu[4], ga[4], gl[4];
    // per fused stage: D and F rows (slot row & 3); between stage q and q+1
    // the rows stage q emits: Y (the next stage's input), ktot, U and V
    c2<T> sd[NF][4], sf[NF][4], sy[NF][4], sa[NF][4], su[NF][4];
    T sv[NF][4];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        gy[k] = z;
        gv[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gu[k] = z; ga[k] = z; gl[k] = z;
    }
#pragma unroll
    for (int q = 0; q < NF; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            sd[q][k] = z; sf[q][k] = z; sy[q][k] = z; sa[q][k] = z; su[q][k] = z; sv[q][k] = 0;
        }

    auto colq = [&](auto v) {  // the value at the MSD partner column of a boundary lane
        const auto p = lane_shift2<DPP_NEXT, T>(v), m = lane_shift2<DPP_PREV, T>(v);
        return colL ? p : (colR ? m : v);
    };
    auto colq1 = [&](T v) {
        const T p = lane_shift<DPP_NEXT>(v), m = lane_shift<DPP_PREV>(v);
        return colL ? p : (colR ? m : v);
    };

    // Every ring slot is loaded each time its row comes round, unconditionally:
    // rows outside the needed range read the nearest needed row instead (a
    // cache hit, never used), so each load targets its ring register directly
    // and stays in flight until the step that consumes it.
    auto rowc = [&](int rn, int lo, int hi) { return rn < lo ? lo : (rn >= hi ? hi - 1 : rn); };
    auto load_y = [&](auto sc, int rn) {
        constexpr int sl = decltype(sc)::value & 7;
        const long p = (long)rowc(rn, Ilo, Ihi) * M + jc;
        gy[sl] = yin[p];
        gv[sl] = V[p];
    };

    static_for<0, PF>([&](auto kc) { load_y(kc, base + decltype(kc)::value); });
    for (int t0 = 0; t0 < nsweep; t0 += 8) {
        static_for<0, 8>([&](auto uc) {
            constexpr int uu = decltype(uc)::value;
            const int t = t0 + uu;
            if (t >= nsweep) return;
            load_y(std::integral_constant<int, uu + PF>{}, base + t + PF);
            if constexpr (FIRST > 1) {  // U and ktot of stage 1's output row DLY steps later
                const long p = (long)rowc(base + t, Flo[0], Fhi[0]) * M + jc;
                gu[uu & 3] = u[p];
                ga[uu & 3] = acc[p];
            }
            if constexpr (COMP && LAST == 4) {  // low part of the last stage's output row 3 steps later
                constexpr int dl = 3 - NF * DLY;
                gl[(uu + dl) & 3] = ulo[(long)rowc(base + t + dl, ibeg, iend) * M + jc];
            }
            static_for<0, NF>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int QS = FIRST + q;     // RK4 stage number
                constexpr int rr = uu - q * DLY;  // slot index of this stage's input row
                const int r = base + t - q * DLY;
                // input row r - k and its V
                auto Y = [&](int k) -> c2<T> {
                    if constexpr (q == 0) return gy[(rr - k) & 7];
                    else return sy[q - 1][(rr - k) & 3];
                };
                auto Vr = [&](int k) -> T {
                    if constexpr (q == 0) return gv[(rr - k) & 7];
                    else return sv[q - 1][(rr - k) & 3];
                };
                // (b) 2SHOC: D at row r-1 (its j+-1 neighbours are kept for F)
                c2<T> yp1 = z, ym1 = z;
                if constexpr (METHOD == 2) {
                    const int x = r - 1;
                    c2<T> dd = z;
                    if (x >= Dlo[q] && x < Dhi[q]) {
                        const c2<T> yx = Y(1);
                        const T vx = Vr(1);
                        yp1 = lane_shift2<DPP_NEXT, T>(yx);
                        ym1 = lane_shift2<DPP_PREV, T>(yx);
                        if (x > 0 && x < N - 1) {
                            const c2<T> a = Y(0), b = Y(2);
                            dd = mk<T>((a.x + b.x - 4 * yx.x + yp1.x + ym1.x) * c.lh2,
                                       (a.y + b.y - 4 * yx.y + yp1.y + ym1.y) * c.lh2);
                            if (colbc) {
                                c2<T> db = z;
                                if (BC == BC_DIRICHLET) db = dir_d<T>(yx, vx, c);
                                else if (BC == BC_MSD) db = msd_d<T>(colq(dd), colq(yx), colq1(vx), yx, vx, c);
                                if (colL || colR) dd = db;
                            }
                        } else if (x == N - 1) {
                            if (BC == BC_DIRICHLET) dd = dir_d<T>(yx, vx, c);
                            else if (BC == BC_MSD)
                                dd = msd_d<T>(colq(sd[q][(rr - 2) & 3]), colq(Y(2)), colq1(Vr(2)), yx, vx, c);
                        } else {  // x == 0; MSD takes row 1 below
                            if (BC == BC_DIRICHLET) dd = dir_d<T>(yx, vx, c);
                        }
                        if (x == 1 && BC == BC_MSD && Dlo[q] == 0)
                            sd[q][(rr - 2) & 3] = msd_d<T>(colq(dd), colq(yx), colq1(vx), Y(2), Vr(2), c);
                    }
                    sd[q][(rr - 1) & 3] = dd;
                }
                // (c) F at row r - R
                {
                    constexpr int KF = R;  // window offset of the F row
                    const int x = r - KF;
                    c2<T> ff = z;
                    if (x >= Flo[q] && x < Fhi[q]) {
                        const c2<T> yx = Y(KF);
                        const T vx = Vr(KF);
                        if (x > 0 && x < N - 1) {
                            const T n = nlin<T>(c.s, yx, vx);
                            if constexpr (METHOD == 2) {
                                const c2<T> dc = sd[q][(rr - 2) & 3], da = sd[q][(rr - 1) & 3],
                                            db = sd[q][(rr - 3) & 3];
                                const c2<T> dl = lane_shift2<DPP_NEXT, T>(dc), dr = lane_shift2<DPP_PREV, T>(dc);
                                const c2<T> y3 = Y(3);
                                const c2<T> ba = lane_shift2<DPP_NEXT, T>(y3), bb = lane_shift2<DPP_PREV, T>(y3);
                                const c2<T> aa = yp1, ab = ym1;
                                ff = mk<T>(-c.a * dc.y + c.a1_12 * (da.y + db.y + dl.y + dr.y)
                                               - c.a1_6h2 * (aa.y + ab.y - 4 * yx.y + ba.y + bb.y) - n * yx.y,
                                           c.a * dc.x - c.a1_12 * (da.x + db.x + dl.x + dr.x)
                                               + c.a1_6h2 * (aa.x + ab.x - 4 * yx.x + ba.x + bb.x) + n * yx.x);
                            } else {
                                const c2<T> a = Y(0), b = Y(2);
                                const c2<T> l = lane_shift2<DPP_NEXT, T>(yx), rt = lane_shift2<DPP_PREV, T>(yx);
                                ff = mk<T>(-c.ah2 * (a.y - 4 * yx.y + b.y + l.y + rt.y) - n * yx.y,
                                           c.ah2 * (a.x - 4 * yx.x + b.x + l.x + rt.x) + n * yx.x);
                            }
                            if (colbc) {
                                c2<T> fb = z;
                                if (BC == BC_MSD) fb = msd_f<T>(colq(ff), colq(yx), yx);
                                else if (BC == BC_LAP0) fb = lap0_f<T>(yx, vx, c.s);
                                if (colL || colR) ff = fb;
                            }
                        } else if (x == N - 1) {
                            if (BC == BC_MSD) ff = msd_f<T>(colq(sf[q][(rr - KF - 1) & 3]), colq(Y(KF + 1)), yx);
                            else if (BC == BC_LAP0) ff = lap0_f<T>(yx, vx, c.s);
                        } else {  // x == 0; MSD takes row 1 below
                            if (BC == BC_LAP0) ff = lap0_f<T>(yx, vx, c.s);
                        }
                        if (x == 1 && BC == BC_MSD && Flo[q] == 0)
                            sf[q][(rr - KF - 1) & 3] = msd_f<T>(colq(ff), colq(yx), Y(KF + 1));
                    }
                    sf[q][(rr - KF) & 3] = ff;
                }
                // (d) RK4 epilogue at row o = r - DLY
                {
                    const int o = r - DLY;
                    constexpr int ro = uu - (q + 1) * DLY;  // slot index of row o
                    c2<T> ynext = z, anext = z, unext = z;
                    T vnext = 0;
                    if (o >= Flo[q] && o < Fhi[q]) {
                        const c2<T> f = sf[q][(rr - DLY) & 3];
                        c2<T> uo;
                        if constexpr (q > 0) uo = su[q - 1][ro & 3];
                        else if constexpr (FIRST == 1) uo = Y(DLY);
                        else uo = gu[ro & 3];
                        c2<T> a;
                        if constexpr (QS == 1) {
                            a = f;
                        } else {
                            c2<T> pa;
                            if constexpr (q == 0) pa = ga[ro & 3];
                            else pa = sa[q - 1][ro & 3];
                            const T w = (QS == 4) ? (T)1.0 : (T)2.0;
                            a = mk<T>(pa.x + w * f.x, pa.y + w * f.y);
                        }
                        const T cs = (QS == 3) ? c.k : c.k2;
                        if constexpr (q < NF - 1) {
                            anext = a;
                            ynext = mk<T>(uo.x + cs * f.x, uo.y + cs * f.y);
                            unext = uo;
                            vnext = Vr(DLY);
                        } else {
                            if (o >= ibeg && o < iend && own) {
                                const long p = (long)o * M + jc;
                                if constexpr (QS == 4) {
                                    if constexpr (COMP) {
                                        c2<T> hi = uo, lo = gl[ro & 3];
                                        comp_add<T>(hi.x, lo.x, c.k6 * a.x);
                                        comp_add<T>(hi.y, lo.y, c.k6 * a.y);
                                        ulo[p] = lo;
                                        out[p] = hi;
                                    } else {
                                        out[p] = mk<T>(uo.x + c.k6 * a.x, uo.y + c.k6 * a.y);
                                    }
                                } else {
                                    out[p] = mk<T>(uo.x + cs * f.x, uo.y + cs * f.y);
                                    acc[p] = a;
                                }
                            }
                        }
                    }
                    if constexpr (q < NF - 1) {
                        sy[q][ro & 3] = ynext;
                        sa[q][ro & 3] = anext;
                        su[q][ro & 3] = unext;
                        sv[q][ro & 3] = vnext;
                    }
                }
            });
        });
    }
}

// Decomposition of an N x M grid into strips of W columns and segments of S
// rows (S >= 3).  Origins shift by one when n mod size == 1 so that every
// block owns at least two rows and two columns (an MSD boundary point's
// partner is then computed in the same wave).  Units are numbered segment-major; consecutive
// units go to the same XCD (workgroup b runs on XCD b mod 8).
struct Strip2 {
    int N, M, S;
    int ioff, joff;
    int nstrip, nunit;
    int per_xcd;  // workgroups (4 waves each) per XCD
    __device__ __forceinline__ int unit() const
    {
        const int b = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
        // the wave index is uniform across the wave: keep it (and every row
        // index derived from it) in scalar registers
        return b * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    }
};

inline Strip2 make_strip2(int N, int M, int S, int W)
{
    Strip2 g;
    g.N = N;
    g.M = M;
    g.S = S;
    g.ioff = (N % S == 1) ? 1 : 0;
    g.joff = (M % W == 1) ? 1 : 0;
    g.nstrip = (M + g.joff + W - 1) / W;
    const int nseg = (N + g.ioff + S - 1) / S;
    g.nunit = nseg * g.nstrip;
    const int nblk = (g.nunit + 3) / 4;
    g.per_xcd = (nblk + 7) / 8;
    return g;
}

}  // namespace nlse
