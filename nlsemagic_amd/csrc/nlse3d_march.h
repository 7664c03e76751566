// nlse3d_march.h -- 3D 2SHOC RK4 stage with D kept on chip (device code).
//
// A workgroup of NW waves owns a tile of TJ rows (j) x 60 columns (k) and
// marches along i over a chunk of IC planes:
//   - lane l is column k0 - 2 + l (two halo columns each side), extended row
//     e = 0 .. TJ+3 is row j0 - 2 + e and belongs to wave e mod NW;
//   - at step p the plane p of the stage input arrives (one coalesced row
//     load per row), D = Lap_CD(Y)/h^2 of plane p-1 is formed, and F plus
//     the RK4 update of plane p-2 (the output plane) follow;
//   - each wave keeps its rows' Y of planes p..p-3 and D of planes p-1..p-3
//     in registers (the i +- 1 neighbours), reads the j +- 1 neighbours (and
//     the j,k edge neighbours) from LDS rings of 4 Y planes and 2 D planes,
//     and takes the k +- 1 neighbours from adjacent lanes (DPP);
//   - one barrier per plane: every ring slot written at step p is read at
//     steps p+1 .. p+3 only.
// D therefore never touches HBM: per stage a point moves the bytes of one
// fused stage (56/88/88/72 B in FP64).  Boundary points follow the local
// rules (Dirichlet D, Lap0 F, zero otherwise); MSD needs a partner that may
// sit in another plane and another wave, and runs on the two-pass kernels.
// The floating-point operations and their order are those of
// NLSE3D_TAKE_STEPS_2SHOC.c:62-69,239-259, so results are bit-identical to
// the two-pass kernels and to the reference.
#pragma once

#include <type_traits>

#include "nlse2d_strip.h"

namespace nlse {

struct March3 {
    int N, M, L;     // grid: L planes (i) x N rows (j) x M columns (k)
    int IC;          // planes per chunk
    int tj, tk, ti;  // tiles along j, k and chunks along i
    int per_xcd;     // workgroups per XCD
};

template <int TJ, int NW = 4>
struct March3Cfg {
    static constexpr int TE = TJ + 4;               // extended rows
    static constexpr int RW = (TE + NW - 1) / NW;   // extended rows per wave
    static constexpr int W = 60;             // owned columns
};

template <typename T, int TJ>
struct March3Lds {
    static constexpr int TE = March3Cfg<TJ>::TE;
    c2<T> y[4][TE * 64 + 2];  // one padding element before and after each plane
    c2<T> d[2][TE * 64];
};

// the RK4 stage update with the ktot and U_lo values already loaded
// (the arithmetic of rk4_epilogue, operand for operand)
template <typename T>
__device__ __forceinline__ void rk4_update(int stage, const Coefs<T>& c, c2<T> f, c2<T> u, c2<T> a, c2<T> lo,
                                           c2<T>* __restrict__ acc, c2<T>* yout, long p, c2<T>* ulo)
{
    if (stage == 1) {
        acc[p] = f;
        yout[p] = mk<T>(u.x + c.k2 * f.x, u.y + c.k2 * f.y);
    } else if (stage == 4) {
        a.x = a.x + (T)1.0 * f.x;
        a.y = a.y + (T)1.0 * f.y;
        if (ulo) {
            comp_add<T>(u.x, lo.x, c.k6 * a.x);
            comp_add<T>(u.y, lo.y, c.k6 * a.y);
            ulo[p] = lo;
            yout[p] = u;
        } else {
            yout[p] = mk<T>(u.x + c.k6 * a.x, u.y + c.k6 * a.y);
        }
    } else {
        a.x = a.x + (T)2.0 * f.x;
        a.y = a.y + (T)2.0 * f.y;
        acc[p] = a;
        const T kk = (stage == 2) ? c.k2 : c.k;
        yout[p] = mk<T>(u.x + kk * f.x, u.y + kk * f.y);
    }
}

template <typename T, int TJ, int NW>
__device__ __forceinline__ void march3d_tile(March3Lds<T, TJ>& sm, int i0, int j0, int k0, const March3& g,
                                             const c2<T>* __restrict__ yin, const c2<T>* u, c2<T>* __restrict__ acc,
                                             c2<T>* yout, const T* __restrict__ V, const Coefs<T>& c, int BC,
                                             int stage, c2<T>* ulo)
{
    using Cfg = March3Cfg<TJ, NW>;
    constexpr int TE = Cfg::TE, RW = Cfg::RW;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __builtin_assume(w >= 0 && w < NW);
    const int N = g.N, M = g.M, L = g.L;
    const int k = k0 - 2 + lane;
    const int kc = k < 0 ? 0 : (k >= M ? M - 1 : k);
    const bool own_k = lane >= 2 && lane < 62 && k < M;
    const bool kb = (k == 0 || k == M - 1);
    const long si = (long)N * M;
    const c2<T> z = mk<T>(0, 0);

    int jrow[RW];
    long rowoff[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int e = NW * r + w;
        const int j = j0 - 2 + e;
        jrow[r] = j;
        rowoff[r] = (long)(j < 0 ? 0 : (j >= N ? N - 1 : j)) * M + kc;
    }
    // own rows: Y of planes p .. p-3 and D of planes p-1 .. p-3 (ring slot
    // = (plane - p_start) & 3), V of planes p .. p-2
    c2<T> Yr[4][RW], Dr[4][RW], Yn[RW];
    T Vr[4][RW], Vn[RW];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Yr[s][r] = z;
            Dr[s][r] = z;
            Vr[s][r] = 0;
        }
    // RK4 update operands (U, ktot and, for a compensated single-precision
    // state, U_lo) of the plane updated at the next step, loaded at the end
    // of the step before
    constexpr bool HAS_LO = std::is_same<T, float>::value;
    c2<T> Up[RW], Ap[RW], Lp[HAS_LO ? RW : 1];
#pragma unroll
    for (int r = 0; r < RW; ++r) Up[r] = Ap[r] = z;
#pragma unroll
    for (int r = 0; r < (HAS_LO ? RW : 1); ++r) Lp[r] = z;
    auto sy = [&](int slot, int e, int dl) -> c2<T> { return sm.y[slot][1 + e * 64 + lane + dl]; };
    auto sdr = [&](int slot, int e) -> c2<T> { return sm.d[slot][e * 64 + lane]; };

    const int pstart = i0 - 2;
    const int nstep = g.IC + 4;
    const int iend = min(i0 + g.IC, L);
    // the next plane's rows are loaded one step ahead (planes outside the
    // grid read a clamped plane, never used)
    auto load_plane = [&](int pl) {
        const long pb = (long)(pl < 0 ? 0 : (pl >= L ? L - 1 : pl)) * si;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Yn[r] = yin[pb + rowoff[r]];
            Vn[r] = V[pb + rowoff[r]];
        }
    };
    load_plane(pstart);
    for (int q0 = 0; q0 < nstep; q0 += 4) {
        static_for<0, 4>([&](auto uc) {
            constexpr int u4 = decltype(uc)::value;
            const int step = q0 + u4;
            if (step >= nstep) return;
            const int p = pstart + step;
            constexpr int s0 = u4 & 3, s1 = (u4 + 3) & 3, s2 = (u4 + 2) & 3, s3 = (u4 + 1) & 3;  // planes p .. p-3
            // (a) plane p of the stage input, then the loads of plane p+1
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const int e = NW * r + w;
                Yr[s0][r] = Yn[r];
                Vr[s0][r] = Vn[r];
                if (e < TE) sm.y[s0][1 + e * 64 + lane] = Yr[s0][r];
            }
            load_plane(p + 1);
            // (b) D of plane p-1, extended rows 1 .. TE-2
            const int pd = p - 1;
            if (pd >= 0 && pd < L) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 1 || e > TE - 2) continue;
                    const int j = jrow[r];
                    const c2<T> m = Yr[s1][r];
                    c2<T> dd = z;
                    const bool inner_row = pd > 0 && pd < L - 1 && j > 0 && j < N - 1;  // uniform
                    if (inner_row) {
                        // NLSE3D_TAKE_STEPS_2SHOC.c:62-69: i+1, i-1, j+1, j-1, k+1, k-1, centre
                        const c2<T> a = Yr[s0][r], b = Yr[s2][r], dj = sy(s1, e + 1, 0), ej = sy(s1, e - 1, 0);
                        const c2<T> f = lane_shift2<DPP_NEXT, T>(m), h = lane_shift2<DPP_PREV, T>(m);
                        dd = mk<T>(c.lh2 * (a.x + b.x + dj.x + ej.x + f.x + h.x - 6 * m.x),
                                   c.lh2 * (a.y + b.y + dj.y + ej.y + f.y + h.y - 6 * m.y));
                    }
                    if (!inner_row || kb) dd = (BC == BC_DIRICHLET) ? dir_d<T>(m, Vr[s1][r], c) : z;
                    Dr[s1][r] = dd;
                    sm.d[(pd - pstart) & 1][e * 64 + lane] = dd;
                }
            }
            // (c) F and the RK4 update of plane p-2 (owned rows 2 .. TE-3)
            const int pf = p - 2;
            if (pf >= i0 && pf < iend) {
                const int sd = (pf - pstart) & 1;
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 2 || e > TE - 3) continue;
                    const int j = jrow[r];
                    if (j >= N) continue;
                    const c2<T> m = Yr[s2][r];
                    const T vx = Vr[s2][r];
                    c2<T> f = z;
                    const bool inner = pf > 0 && pf < L - 1 && j > 0 && j < N - 1;  // uniform; k below
                    if (inner) {
                        const T wv = vx - c.s * (m.x * m.x + m.y * m.y);
                        const c2<T> yp1 = Yr[s1][r], ym1 = Yr[s3][r];  // planes i+1, i-1
                        // edge neighbours in NLSE3D_TAKE_STEPS_2SHOC.c source order (i, j, k offsets)
                        const c2<T> E[12] = {
                            sy(s1, e + 1, 0),                    // (+1,+1, 0)
                            sy(s2, e + 1, 1),                    // ( 0,+1,+1)
                            lane_shift2<DPP_NEXT, T>(yp1),       // (+1, 0,+1)
                            sy(s3, e - 1, 0),                    // (-1,-1, 0)
                            sy(s2, e - 1, -1),                   // ( 0,-1,-1)
                            lane_shift2<DPP_PREV, T>(ym1),       // (-1, 0,-1)
                            sy(s1, e - 1, 0),                    // (+1,-1, 0)
                            sy(s2, e + 1, -1),                   // ( 0,+1,-1)
                            lane_shift2<DPP_PREV, T>(yp1),       // (+1, 0,-1)
                            sy(s3, e + 1, 0),                    // (-1,+1, 0)
                            sy(s2, e - 1, 1),                    // ( 0,-1,+1)
                            lane_shift2<DPP_NEXT, T>(ym1)};      // (-1, 0,+1)
                        c2<T> es = E[0];
#pragma unroll
                        for (int t = 1; t < 12; ++t) {
                            es.x = es.x + E[t].x;
                            es.y = es.y + E[t].y;
                        }
                        const c2<T> dm = Dr[s2][r];
                        const c2<T> da = Dr[s1][r], db = Dr[s3][r], dd = sdr(sd, e + 1), de = sdr(sd, e - 1);
                        const c2<T> df = lane_shift2<DPP_NEXT, T>(dm), dh = lane_shift2<DPP_PREV, T>(dm);
                        f = mk<T>(c.a_12 * (da.y + db.y + dd.y + de.y + df.y + dh.y - 10 * dm.y) - c.a_6h2 * (es.y - 12 * m.y)
                                      + wv * m.y,
                                  -c.a_12 * (da.x + db.x + dd.x + de.x + df.x + dh.x - 10 * dm.x)
                                      + c.a_6h2 * (es.x - 12 * m.x) - wv * m.x);
                    }
                    if (!inner || kb) f = (BC == BC_LAP0) ? lap0_f<T>(m, vx, c.s) : z;
                    if (own_k) {
                        const long pp = (long)pf * si + (long)j * M + k;
                        const c2<T> uu = (stage == 1) ? m : Up[r];
                        rk4_update<T>(stage, c, f, uu, Ap[r], HAS_LO ? Lp[HAS_LO ? r : 0] : z, acc, yout, pp,
                                      HAS_LO ? ulo : nullptr);
                    }
                }
            }
            // (d) operands of plane p-1, updated at the next step (own rows;
            // clamped addresses: values of columns this tile does not own are
            // never used)
            const int pn = p - 1;
            if (stage != 1 && pn >= i0 && pn < iend) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 2 || e > TE - 3) continue;
                    const long pp = (long)pn * si + rowoff[r];
                    Up[r] = u[pp];
                    Ap[r] = acc[pp];
                    if constexpr (HAS_LO)
                        if (stage == 4 && ulo) Lp[r] = ulo[pp];
                }
            }
            __syncthreads();  // the slots written at this step are complete; the oldest may be reused
        });
    }
}

}  // namespace nlse
