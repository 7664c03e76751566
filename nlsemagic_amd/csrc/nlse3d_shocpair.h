// nlse3d_shocpair.h -- two 2SHOC RK4 stages per launch in 3D, marching along
// i (device code).
//
// A workgroup of NW = TJ + 8 waves owns TJ rows (j) x 56 columns (k) and
// marches along i over a chunk of IC planes.  Lane l is column k0 - 4 + l
// (four halo columns each side); extended row e = 0 .. TJ+7 is row j0 - 4 + e
// and belongs to wave e in every phase.  At step p:
//   (a) plane p of the pair's input Y arrives (loaded one step ahead) and goes
//       to the LDS ring;
//   (b) D = Lap_CD(Y)/h^2 of the first stage at plane p-1, rows 1 .. TJ+6;
//   (c) F of the first stage at plane p-2, rows 2 .. TJ+5: its output Y1
//       (to a register ring and the LDS ring) and, for owned rows, the
//       running ktot (to a per-lane LDS ring);
//   (d) D of the second stage, from Y1, at plane p-3, rows 3 .. TJ+4;
//   (e) F of the second stage at plane p-4, owned rows 4 .. TJ+3, and the
//       update is written;
//   (f) one barrier.
// Where the values come from:
//   - i+-1 neighbours of the stage input: the wave's own rows of neighbouring
//     planes (register rings, slot = (plane - first plane) mod 4), except
//     the (i-1, k+-1) edge neighbours, read from the LDS ring of plane i-1;
//   - j+-1 neighbours and the j,k edge neighbours: LDS rings of four planes;
//   - k+-1 neighbours: adjacent lanes (DPP); the D phase's k+-1 lane shifts
//     and j+-1 reads of a plane are also the (+1,0,+-1) and (+1,+-1,0) edge
//     neighbours of the F phase of the same step;
//   - D of both stages: four-plane LDS rings (centre, j+-1 and i-1 values;
//     the D just formed, the i+1 value, stays in a register).
// Every LDS slot written at step p is read at steps p+1 .. p+3 only, so one
// barrier per plane suffices.  Lanes and rows in the halo hold values no
// owned point depends on once their inputs reach beyond the tile (each
// stencil level loses one lane and one row on each side).  TJ = 4 for FP64
// (the LDS rings then take exactly 160 KiB), 8 for FP32.
//
// Per point a pair moves the bytes of one fused stage (stages 1-2: Y, V in;
// Y2, ktot out; stages 3-4: Y2, V, U, ktot in; U out); D, the first stage's
// output and its ktot never leave the chip.  Boundary points follow the
// local rules (Dirichlet D, Lap0 F, zero otherwise); MSD needs a partner in
// another plane and runs on the per-stage kernels.  The floating-point
// operations and their order are those of NLSE3D_TAKE_STEPS_2SHOC.c:62-69,
// 239-259 and the RK4 stage algebra (:490-521), so results are
// bit-identical to one stage per launch and to the reference.
//
// Every global load is issued on every step in the same order (clamped
// planes and rows; the values of planes and rows a step does not use are
// never read), and the load sequence before the loop is the one at the end
// of a step, so the compiler waits for each load by count instead of
// draining all of them at each step.  The LDS rings need no padding: lane
// +- 1 reads touch rows 1 .. TE-2 of y and 1 .. TE-6 of y1 only.  Row offsets
// within a plane are 32-bit (N * M < 2^31).
#pragma once

#include <type_traits>

#include "nlse2d_strip.h"
#include "nlse3d_march.h"

namespace nlse {

template <typename T, int TJ>
struct ShocPairLds {
    static constexpr int TE = TJ + 8;
    c2<T> y[4][TE * 64];         // the pair's input, rows 0 .. TE-1
    c2<T> y1[4][(TE - 4) * 64];  // the first stage's output, rows 2 .. TE-3
    c2<T> da[4][(TE - 2) * 64];  // the first stage's D, rows 1 .. TE-2
    c2<T> db[4][(TE - 6) * 64];  // the second stage's D, rows 3 .. TE-4
    c2<T> kk[4][TJ * 64];        // the running ktot of owned rows 4 .. TE-5
};

template <typename T, int FIRST, int TJ, int NW>
__device__ __forceinline__ void shocpair_tile(ShocPairLds<T, TJ>& sm, int i0, int j0, int k0, const March3& g,
                                              const c2<T>* __restrict__ yin, const c2<T>* u,
                                              c2<T>* __restrict__ acc, c2<T>* __restrict__ yout,
                                              const T* __restrict__ V, const Coefs<T>& c, int BC, c2<T>* ulo)
{
    constexpr int TE = TJ + 8, RW = (TE + NW - 1) / NW;
    constexpr bool SECOND = (FIRST == 3);
    constexpr bool HAS_LO = SECOND && std::is_same<T, float>::value;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __builtin_assume(w >= 0 && w < NW);
    const int N = g.N, M = g.M, L = g.L;
    const int k = k0 - 4 + lane;
    const int kc = k < 0 ? 0 : (k >= M ? M - 1 : k);
    const bool kb = (k == 0 || k == M - 1);
    const bool own_k = lane >= 4 && lane < 60 && k < M;
    const long si = (long)N * M;
    const c2<T> z = mk<T>(0, 0);

    // row offsets (j * M + k, clamped into the grid) of each extended row and
    // of the rows whose update operands it loads: rows 2 .. TE-3 for the first
    // stage, owned rows 4 .. TE-5 for the second (other rows load a copy of
    // the nearest such row, so every wave issues the same loads)
    int jrow[RW], ro[RW], roa[RW], rob[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int e = NW * r + w;
        const int j = j0 - 4 + e;
        jrow[r] = j;
        auto off = [&](int ee) {
            const int jj = j0 - 4 + ee;
            return (jj < 0 ? 0 : (jj >= N ? N - 1 : jj)) * M + kc;
        };
        ro[r] = off(e);
        roa[r] = off(e < 2 ? 2 : (e > TE - 3 ? TE - 3 : e));
        rob[r] = off(e < 4 ? 4 : (e > TE - 5 ? TE - 5 : e));
    }
    // register rings (slot = (plane - pstart) & 3): Y and V of planes
    // p .. p-3 (Y of p-3 unused), Y1 of p-2 .. p-4
    c2<T> Yr[4][RW], Y1r[4][RW];
    T Vr[4][RW];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Yr[s][r] = Y1r[s][r] = z;
            Vr[s][r] = 0;
        }
    // loaded one step ahead: the next input plane; the operands of the next
    // step's first-stage update (stages 3-4: U and ktot of plane p-1) and of
    // its second-stage update (U, V and U_lo of plane p-3)
    c2<T> Yn[RW], Ua[RW], Aa[RW], Ub[RW], Lb[RW];
    T Vn[RW], Vb[RW];
    auto ly = [&](int slot, int e, int dl) -> c2<T> { return sm.y[slot][e * 64 + lane + dl]; };
    auto ly1 = [&](int slot, int e, int dl) -> c2<T> { return sm.y1[slot][(e - 2) * 64 + lane + dl]; };
    auto lda = [&](int slot, int e) -> c2<T> { return sm.da[slot][(e - 1) * 64 + lane]; };
    auto ldb = [&](int slot, int e) -> c2<T> { return sm.db[slot][(e - 3) * 64 + lane]; };

    const int pstart = i0 - 4;
    const int iend = min(i0 + g.IC, L);
    const int nstep = (iend - i0 + 8 + 3) & ~3;  // whole groups of four steps
    auto plane = [&](int pl, int lo, int hi) -> long { return (long)(pl < lo ? lo : (pl > hi ? hi : pl)) * si; };
    auto load_plane = [&](int pl) {
        const long pb = plane(pl, 0, L - 1);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Yn[r] = yin[pb + ro[r]];
            Vn[r] = V[pb + ro[r]];
        }
    };
    auto load_operands = [&](int p) {
        if constexpr (SECOND) {
            const long pa = plane(p - 1, 0, L - 1);  // the next step's first-stage plane
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                Ua[r] = u[pa + roa[r]];
                Aa[r] = acc[pa + roa[r]];
            }
        }
        const long pb = plane(p - 3, i0, iend - 1);  // the next step's second-stage plane
        const c2<T>* const lo_src = ulo ? ulo : u;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            Ub[r] = u[pb + rob[r]];
            Vb[r] = V[pb + rob[r]];
            if constexpr (HAS_LO) Lb[r] = lo_src[pb + rob[r]];  // unused without a low part
        }
    };
    // D of one row (NLSE3D_TAKE_STEPS_2SHOC.c:62-69: i+1, i-1, j+1, j-1, k+1,
    // k-1, centre) with its boundary value; kn, kp are the k+-1 neighbours
    auto dcalc = [&](bool inner_row, c2<T> m, T v, c2<T> a, c2<T> b, c2<T> dj, c2<T> ej, c2<T> kn, c2<T> kp) {
        c2<T> dd;  // set on every path below
        if (inner_row)
            dd = mk<T>(c.lh2 * (a.x + b.x + dj.x + ej.x + kn.x + kp.x - 6 * m.x),
                       c.lh2 * (a.y + b.y + dj.y + ej.y + kn.y + kp.y - 6 * m.y));
        if (!inner_row || kb) dd = (BC == BC_DIRICHLET) ? dir_d<T>(m, v, c) : z;
        return dd;
    };
    // F of one row (NLSE3D_TAKE_STEPS_2SHOC.c:239-259) from its centre m, V,
    // the twelve edge neighbours E(0 .. 11) (source order) and the D values
    // (i+1, i-1, j+1, j-1, centre); the scheduler may not hoist the loads of
    // the second six edge neighbours above the first six's sum, which bounds
    // the registers held by loads in flight
    auto fcalc = [&](bool inner, c2<T> m, T vx, auto E, c2<T> da, c2<T> db, c2<T> dd, c2<T> de, c2<T> dm) {
        c2<T> f;  // set on every path below
        if (inner) {
            const T wv = vx - c.s * (m.x * m.x + m.y * m.y);
            c2<T> es = E(0);
            static_for<1, 12>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                if constexpr (t == 6) __builtin_amdgcn_sched_barrier(0);
                const c2<T> v = E(t);
                es.x = es.x + v.x;
                es.y = es.y + v.y;
            });
            const c2<T> df = lane_shift2<DPP_NEXT, T>(dm), dh = lane_shift2<DPP_PREV, T>(dm);
            f = mk<T>(c.a_12 * (da.y + db.y + dd.y + de.y + df.y + dh.y - 10 * dm.y) - c.a_6h2 * (es.y - 12 * m.y)
                          + wv * m.y,
                      -c.a_12 * (da.x + db.x + dd.x + de.x + df.x + dh.x - 10 * dm.x)
                          + c.a_6h2 * (es.x - 12 * m.x) - wv * m.x);
        }
        if (!inner || kb) f = (BC == BC_LAP0) ? lap0_f<T>(m, vx, c.s) : z;
        return f;
    };
    // planes each phase works on: D of the first stage on i0-3 .. iend+2, its
    // F on i0-2 .. iend+1, D of the second on i0-1 .. iend, its F on the
    // owned planes (each clipped to the grid)
    auto in = [&](int pl, int lo, int hi) { return pl >= lo && pl <= hi && pl >= 0 && pl < L; };

    load_plane(pstart);
    load_operands(pstart - 1);
    for (int q0 = 0; q0 < nstep; q0 += 4) {
        static_for<0, 4>([&](auto uc) {
            constexpr int u4 = decltype(uc)::value;
            const int p = pstart + q0 + u4;
            // slots of planes p, p-1, p-2, p-3 (p-4 shares p's, p-5 p-1's)
            constexpr int s0 = u4 & 3, s1 = (u4 + 3) & 3, s2 = (u4 + 2) & 3, s3 = (u4 + 1) & 3;
            // (a) plane p of the input, then the loads of plane p+1
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const int e = NW * r + w;
                Yr[s0][r] = Yn[r];
                Vr[s0][r] = Vn[r];
                if (e < TE) sm.y[s0][e * 64 + lane] = Yn[r];
            }
            load_plane(p + 1);
            // (b) the first stage's D at plane p-1, rows 1 .. TE-2; its
            // neighbours of Y(p-1) at j+-1 and k+-1 are also F's at (c).
            // (Left unset on rows and planes without D: (c) reads them only
            // for rows 2 .. TE-3 and, D being used only at interior planes,
            // planes whose i+1 plane (b) has formed.)
            c2<T> Dn[RW], Xj[RW], Xjm[RW], Xk[RW], Xkm[RW];
            {
                const int pd = p - 1;
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 1 || e > TE - 2) continue;
                    const c2<T> m = Yr[s1][r];
                    Xk[r] = lane_shift2<DPP_NEXT, T>(m);
                    Xkm[r] = lane_shift2<DPP_PREV, T>(m);
                    Xj[r] = ly(s1, e + 1, 0);
                    Xjm[r] = ly(s1, e - 1, 0);
                    if (in(pd, i0 - 3, iend + 2)) {
                        const int j = jrow[r];
                        const bool inner_row = pd > 0 && pd < L - 1 && j > 0 && j < N - 1;  // uniform
                        const c2<T> dd = dcalc(inner_row, m, Vr[s1][r], Yr[s0][r], Yr[s2][r], Xj[r], Xjm[r],
                                               Xk[r], Xkm[r]);
                        Dn[r] = dd;
                        sm.da[s1][(e - 1) * 64 + lane] = dd;
                    }
                }
            }
            // (c) the first stage's F at plane p-2, rows 2 .. TE-3, and Y1
            {
                const int q = p - 2;
                if (in(q, i0 - 2, iend + 1)) {
#pragma unroll
                    for (int r = 0; r < RW; ++r) {
                        const int e = NW * r + w;
                        if (e < 2 || e > TE - 3) continue;
                        const int j = jrow[r];
                        const c2<T> m = Yr[s2][r];
                        auto E = [&](int t) -> c2<T> {
                            switch (t) {
                            case 0: return Xj[r];              // (+1,+1, 0)
                            case 1: return ly(s2, e + 1, 1);   // ( 0,+1,+1)
                            case 2: return Xk[r];              // (+1, 0,+1)
                            case 3: return ly(s3, e - 1, 0);   // (-1,-1, 0)
                            case 4: return ly(s2, e - 1, -1);  // ( 0,-1,-1)
                            case 5: return ly(s3, e, -1);      // (-1, 0,-1)
                            case 6: return Xjm[r];             // (+1,-1, 0)
                            case 7: return ly(s2, e + 1, -1);  // ( 0,+1,-1)
                            case 8: return Xkm[r];             // (+1, 0,-1)
                            case 9: return ly(s3, e + 1, 0);   // (-1,+1, 0)
                            case 10: return ly(s2, e - 1, 1);  // ( 0,-1,+1)
                            default: return ly(s3, e, 1);      // (-1, 0,+1)
                            }
                        };
                        const bool inner = q > 0 && q < L - 1 && j > 0 && j < N - 1;  // uniform; k in fcalc
                        const c2<T> f = fcalc(inner, m, Vr[s2][r], E, Dn[r], lda(s3, e), lda(s2, e + 1),
                                              lda(s2, e - 1), lda(s2, e));
                        c2<T> y1, kk;
                        if constexpr (!SECOND) {
                            // stage 1: ktot = K1, Y1 = U + k/2 K1 (the input is U)
                            kk = f;
                            y1 = mk<T>(m.x + c.k2 * f.x, m.y + c.k2 * f.y);
                        } else {
                            // stage 3: ktot = ktot + 2 K3, Y3 = U + k K3
                            c2<T> a = Aa[r];
                            a.x = a.x + (T)2.0 * f.x;
                            a.y = a.y + (T)2.0 * f.y;
                            kk = a;
                            y1 = mk<T>(Ua[r].x + c.k * f.x, Ua[r].y + c.k * f.y);
                        }
                        Y1r[s2][r] = y1;
                        sm.y1[s2][(e - 2) * 64 + lane] = y1;
                        if (e >= 4 && e <= TE - 5) sm.kk[s2][(e - 4) * 64 + lane] = kk;
                    }
                }
            }
            // (d) the second stage's D at plane p-3, rows 3 .. TE-4, from Y1;
            // its neighbours of Y1(p-3) at j+-1 and k+-1 are also F's at (e)
            // (left unset where (e) does not read them, as at (b))
            c2<T> Bn[RW], Bj[RW], Bjm[RW], Bk[RW], Bkm[RW];
            {
                const int pd = p - 3;
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int e = NW * r + w;
                    if (e < 3 || e > TE - 4) continue;
                    const c2<T> m = Y1r[s3][r];
                    Bk[r] = lane_shift2<DPP_NEXT, T>(m);
                    Bkm[r] = lane_shift2<DPP_PREV, T>(m);
                    Bj[r] = ly1(s3, e + 1, 0);
                    Bjm[r] = ly1(s3, e - 1, 0);
                    if (in(pd, i0 - 1, iend)) {
                        const int j = jrow[r];
                        const bool inner_row = pd > 0 && pd < L - 1 && j > 0 && j < N - 1;  // uniform
                        const c2<T> dd = dcalc(inner_row, m, Vr[s3][r], Y1r[s2][r], Y1r[s0][r], Bj[r], Bjm[r],
                                               Bk[r], Bkm[r]);
                        Bn[r] = dd;
                        sm.db[s3][(e - 3) * 64 + lane] = dd;
                    }
                }
            }
            // (e) the second stage's F at plane p-4, owned rows 4 .. TE-5, and the update
            {
                const int q = p - 4;
                if (q >= i0 && q < iend) {
#pragma unroll
                    for (int r = 0; r < RW; ++r) {
                        const int e = NW * r + w;
                        if (e < 4 || e > TE - 5) continue;
                        const int j = jrow[r];
                        if (j >= N) continue;
                        // Y1 planes: i+1 = p-3 (s3), i = p-4 (s0), i-1 = p-5 (s1)
                        const c2<T> m = Y1r[s0][r];
                        auto E = [&](int t) -> c2<T> {
                            switch (t) {
                            case 0: return Bj[r];               // (+1,+1, 0)
                            case 1: return ly1(s0, e + 1, 1);   // ( 0,+1,+1)
                            case 2: return Bk[r];               // (+1, 0,+1)
                            case 3: return ly1(s1, e - 1, 0);   // (-1,-1, 0)
                            case 4: return ly1(s0, e - 1, -1);  // ( 0,-1,-1)
                            case 5: return ly1(s1, e, -1);      // (-1, 0,-1)
                            case 6: return Bjm[r];              // (+1,-1, 0)
                            case 7: return ly1(s0, e + 1, -1);  // ( 0,+1,-1)
                            case 8: return Bkm[r];              // (+1, 0,-1)
                            case 9: return ly1(s1, e + 1, 0);   // (-1,+1, 0)
                            case 10: return ly1(s0, e - 1, 1);  // ( 0,-1,+1)
                            default: return ly1(s1, e, 1);      // (-1, 0,+1)
                            }
                        };
                        const bool inner = q > 0 && q < L - 1 && j > 0 && j < N - 1;  // uniform; k in fcalc
                        const c2<T> f = fcalc(inner, m, Vb[r], E, Bn[r], ldb(s1, e), ldb(s0, e + 1), ldb(s0, e - 1),
                                              ldb(s0, e));
                        if (own_k) {
                            const long pp = (long)q * si + (long)j * M + k;
                            c2<T> a = sm.kk[s0][(e - 4) * 64 + lane];  // written at (c) two steps ago
                            c2<T> b = Ub[r];
                            if constexpr (!SECOND) {
                                // stage 2: ktot = K1 + 2 K2, Y2 = U + k/2 K2
                                a.x = a.x + (T)2.0 * f.x;
                                a.y = a.y + (T)2.0 * f.y;
                                st_stream<T>(&acc[pp], a);
                                st_stream<T>(&yout[pp], mk<T>(b.x + c.k2 * f.x, b.y + c.k2 * f.y));
                            } else {
                                // stage 4: U + k/6 (ktot + K4), compensated when a low part is kept
                                a.x = a.x + (T)1.0 * f.x;
                                a.y = a.y + (T)1.0 * f.y;
                                if (HAS_LO && ulo) {
                                    c2<T> lo = Lb[r];
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
                }
            }
            load_operands(p);
            __syncthreads();  // this step's LDS slots are complete; none is read at the step that writes it
        });
    }
}

}  // namespace nlse
