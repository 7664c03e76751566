// nlse3d_cdpair.h -- two CD RK4 stages per launch in 3D, marching along i
// (device code).
//
// A workgroup of NW waves owns TJ rows (j) x 60 columns (k) and marches
// along i over a chunk of IC planes.  Lane l is column k0 - 2 + l; extended
// row e = 0 .. TJ+3 is row j0 - 2 + e and belongs to wave e mod NW in every
// phase.  At step p:
//   (a) plane p of the pair's input Y arrives (with V, and for stages 3-4
//       the base U and ktot), loaded one step ahead;
//   (b) the first stage of the pair at plane p-1, extended rows 1 .. TJ+2:
//       its i-neighbours are the wave's own rows of planes p and p-2
//       (registers), its j-neighbours come from the LDS copy of plane p-1
//       written at the previous step, its k-neighbours from adjacent lanes;
//   (c) the second stage at plane p-2, owned rows, from the first stage's
//       output the same way, then the update is written;
//   (d) one barrier.
// The first stage's output and its ktot contribution never leave the chip:
// per point a pair moves the bytes of one fused stage (stages 1-2: Y, V in;
// Y2, ktot out; stages 3-4: Y2, V, U, ktot in; U out).  Boundary points
// follow the local rules (Lap0 F, zero otherwise); MSD needs a partner in
// another plane and runs on the per-stage kernel.  The floating-point
// operations and their order are those of NLSE3D_TAKE_STEPS_CD.c:59-71 and
// the RK4 stage algebra (:300-321), so results are bit-identical to the
// per-stage kernel and to the reference.
#pragma once

#include <type_traits>

#include "nlse2d_strip.h"
#include "nlse3d_march.h"

namespace nlse {

template <typename T, int TJ>
struct CdPairLds {
    static constexpr int TE = TJ + 4;
    c2<T> y[2][TE][64];   // the pair's input, last two planes (slot = step parity)
    c2<T> y1[2][TE][64];  // the first stage's output, last two planes
};

template <typename T, int FIRST, int TJ, int NW>
__device__ __forceinline__ void cdpair_tile(CdPairLds<T, TJ>& sm, int i0, int j0, int k0, const March3& g,
                                            const c2<T>* __restrict__ yin, const c2<T>* u, c2<T>* __restrict__ acc,
                                            c2<T>* __restrict__ yout, const T* __restrict__ V, const Coefs<T>& c,
                                            int BC, c2<T>* ulo)
{
    constexpr int TE = TJ + 4, RW = (TE + NW - 1) / NW;
    constexpr bool SECOND = (FIRST == 3);  // stages 3-4: base U and ktot come in, U goes out
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __builtin_assume(w >= 0 && w < NW);
    const int N = g.N, M = g.M, L = g.L;
    const int k = k0 - 2 + lane;
    const int kc = k < 0 ? 0 : (k >= M ? M - 1 : k);
    const bool kin = k >= 0 && k < M;
    const bool own_k = lane >= 2 && lane < 62 && k < M;
    const long si = (long)N * M;
    const c2<T> z = mk<T>(0, 0);

    constexpr bool HAS_LO = SECOND && std::is_same<T, float>::value;
    // row offsets (j * M + k, clamped into the grid) of each extended row and
    // of the owned row whose U_lo it loads (rows outside 2 .. TE-3 load a copy
    // of the nearest owned row, so every wave issues the same loads)
    int jrow[RW];
    long rowoff[RW], rob[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int e = NW * r + w;
        const int j = j0 - 2 + e;
        jrow[r] = j;
        auto off = [&](int ee) -> long {
            const int jj = j0 - 2 + ee;
            return (long)(jj < 0 ? 0 : (jj >= N ? N - 1 : jj)) * M + kc;
        };
        rowoff[r] = off(e);
        rob[r] = off(e < 2 ? 2 : (e > TE - 3 ? TE - 3 : e));
    }
    // register rings (slot = (plane - pstart) mod 3): the input Y and V of
    // planes p, p-1, p-2; the first stage's output and its ktot part of
    // planes p-1, p-2, p-3; for stages 3-4 the base U and ktot of planes
    // p, p-1, p-2
    c2<T> Yr[3][RW], Y1r[3][RW], Kr[3][RW], Br[SECOND ? 3 : 1][RW], Ar[SECOND ? 3 : 1][RW];
    T Vr[3][RW];
    c2<T> Yn[RW], Bn[RW], An[RW], Ln[RW];
    T Vn[RW];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Yr[s][r] = Y1r[s][r] = Kr[s][r] = z;
            Vr[s][r] = 0;
            if constexpr (SECOND) Br[s][r] = Ar[s][r] = z;
        }
#pragma unroll
    for (int r = 0; r < RW; ++r) Bn[r] = An[r] = z;

    const int pstart = i0 - 2;
    const int iend = min(i0 + g.IC, L);
    const int nstep = (iend - i0 + 4 + 2) / 3 * 3;  // whole groups of three steps
    // Every global load is issued on every step in the same order (clamped
    // planes and rows; values of planes and rows a step does not use are never
    // read), and the loads before the loop are those of a step, so the
    // compiler waits for each load by count instead of draining all of them
    // at each step.
    auto plane = [&](int pl, int lo, int hi) -> long { return (long)(pl < lo ? lo : (pl > hi ? hi : pl)) * si; };
    auto load_plane = [&](int pl) {
        const long pb = plane(pl, 0, L - 1);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Yn[r] = yin[pb + rowoff[r]];
            Vn[r] = V[pb + rowoff[r]];
            if constexpr (SECOND) {
                Bn[r] = u[pb + rowoff[r]];
                An[r] = acc[pb + rowoff[r]];
            }
        }
    };
    // CD F at a grid point from its value, V and the six neighbours in the
    // reference's order (+i, -i, +j, -j, +k, -k), with the boundary rules
    auto fcd = [&](bool interior, c2<T> m, T v, c2<T> a, c2<T> b, c2<T> d, c2<T> e, c2<T> f, c2<T> h) {
        if (interior) {
            const T wv = v - c.s * (m.x * m.x + m.y * m.y);
            return mk<T>(-c.ah2 * (a.y + b.y + d.y + e.y + f.y + h.y - 6 * m.y) + wv * m.y,
                         c.ah2 * (a.x + b.x + d.x + e.x + f.x + h.x - 6 * m.x) - wv * m.x);
        }
        return (BC == BC_LAP0) ? lap0_f<T>(m, v, c.s) : z;
    };
    // U_lo of the next step's update plane p-1 (compensated stages 3-4; read
    // from U and unused when there is no low part)
    auto load_lo = [&](int p) {
        if constexpr (HAS_LO) {
            const c2<T>* const src = ulo ? ulo : u;
            const long pb = plane(p - 1, i0, iend - 1);
#pragma unroll
            for (int r = 0; r < RW; ++r) Ln[r] = src[pb + rob[r]];
        }
    };
    load_plane(pstart);
    load_lo(pstart - 1);
    for (int q0 = 0; q0 < nstep; q0 += 3) {
        static_for<0, 3>([&](auto uc) {
            constexpr int u3 = decltype(uc)::value;
            const int st = q0 + u3;
            const int p = pstart + st;
            constexpr int a0 = u3, a1 = (u3 + 2) % 3, a2 = (u3 + 1) % 3;  // planes p, p-1, p-2 (p-3 -> a0)
            const int lnow = st & 1, lprev = lnow ^ 1;
            // (a) plane p of the input, then the loads of plane p+1
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const int e = NW * r + w;
                Yr[a0][r] = Yn[r];
                Vr[a0][r] = Vn[r];
                if constexpr (SECOND) {
                    Br[a0][r] = Bn[r];
                    Ar[a0][r] = An[r];
                }
                if (e < TE) sm.y[lnow][e][lane] = Yn[r];
            }
            load_plane(p + 1);
            // (b) the first stage at plane q = p-1 (i0-1 .. iend), extended rows 1 .. TE-2
            const int q = p - 1;
            if (q >= 0 && q < L && q >= i0 - 1 && q <= iend) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 1 || e > TE - 2) continue;
                    const c2<T> m = Yr[a1][r];
                    const c2<T> fp = lane_shift2<DPP_NEXT, T>(m), fm = lane_shift2<DPP_PREV, T>(m);
                    const int j = jrow[r];
                    c2<T> y1 = z, kk = z;
                    if (kin && j >= 0 && j < N) {
                        const bool in = q > 0 && q < L - 1 && j > 0 && j < N - 1 && k > 0 && k < M - 1;
                        const c2<T> f = fcd(in, m, Vr[a1][r], Yr[a0][r], Yr[a2][r], sm.y[lprev][e + 1][lane],
                                            sm.y[lprev][e - 1][lane], fp, fm);
                        if constexpr (!SECOND) {
                            // stage 1: ktot = K1, Y1 = U + k/2 K1 (the input is U)
                            kk = f;
                            y1 = mk<T>(m.x + c.k2 * f.x, m.y + c.k2 * f.y);
                        } else {
                            // stage 3: ktot = ktot + 2 K3, Y3 = U + k K3
                            const c2<T> b = Br[a1][r];
                            c2<T> a = Ar[a1][r];
                            a.x = a.x + (T)2.0 * f.x;
                            a.y = a.y + (T)2.0 * f.y;
                            kk = a;
                            y1 = mk<T>(b.x + c.k * f.x, b.y + c.k * f.y);
                        }
                    }
                    Y1r[a1][r] = y1;
                    Kr[a1][r] = kk;
                    sm.y1[lnow][e][lane] = y1;
                }
            }
            // (c) the second stage at plane p-2, owned rows 2 .. TE-3, and the update
            const int q2 = p - 2;
            if (q2 >= i0 && q2 < iend) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 2 || e > TE - 3) continue;
                    const c2<T> m = Y1r[a2][r];
                    const c2<T> fp = lane_shift2<DPP_NEXT, T>(m), fm = lane_shift2<DPP_PREV, T>(m);
                    const int j = jrow[r];
                    if (!own_k || j >= N) continue;
                    const bool in = q2 > 0 && q2 < L - 1 && j > 0 && j < N - 1 && k > 0 && k < M - 1;
                    const c2<T> f = fcd(in, m, Vr[a2][r], Y1r[a1][r], Y1r[a0][r], sm.y1[lprev][e + 1][lane],
                                        sm.y1[lprev][e - 1][lane], fp, fm);
                    const long pp = (long)q2 * si + (long)j * M + k;
                    c2<T> a = Kr[a2][r];
                    if constexpr (!SECOND) {
                        // stage 2: ktot = K1 + 2 K2, Y2 = U + k/2 K2
                        const c2<T> b = Yr[a2][r];
                        a.x = a.x + (T)2.0 * f.x;
                        a.y = a.y + (T)2.0 * f.y;
                        st_stream<T>(&acc[pp], a);
                        st_stream<T>(&yout[pp], mk<T>(b.x + c.k2 * f.x, b.y + c.k2 * f.y));
                    } else {
                        // stage 4: U + k/6 (ktot + K4), compensated when a low part is kept
                        c2<T> b = Br[a2][r];
                        a.x = a.x + (T)1.0 * f.x;
                        a.y = a.y + (T)1.0 * f.y;
                        if (HAS_LO && ulo) {
                            c2<T> lo = Ln[r];
                            comp_add<T>(b.x, lo.x, c.k6 * a.x);
                            comp_add<T>(b.y, lo.y, c.k6 * a.y);
                            st_stream<T>(&ulo[pp], lo);
                            st_stream<T>(&yout[pp], b);
                        } else {
                            st_stream<T>(&yout[pp], mk<T>(b.x + c.k6 * a.x, b.y + c.k6 * a.y));
                        }
                    }
                }
            }
            load_lo(p);
            __syncthreads();  // this step's LDS planes are complete; the previous step's may be overwritten
        });
    }
}

}  // namespace nlse
