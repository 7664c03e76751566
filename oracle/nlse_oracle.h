/*
 * nlse_oracle.h -- CPU restatement of NLSEmagic's RK4 chunk integrators.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * integrators in nlsemagic_amd: only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it.  The product path never calls it.
 *
 * It restates, expression for expression (same operand order, no FP
 * contraction), the reference CPU integrators:
 *   1D: src/NLSEmagic1D/NLSEmagic1D_CPU.c:3-523
 *   2D: src/NLSEmagic2D/NLSEmagic2D_CPU.c:3-960
 *   3D: src/NLSEmagic3D/NLSE3D_TAKE_STEPS_CD.c:28-322,
 *       src/NLSEmagic3D/NLSE3D_TAKE_STEPS_2SHOC.c:28-530 (+ the _F twins)
 * Differences from the reference, all intended:
 *   - inputs are not modified (the reference writes the evolving state into
 *     the caller's U, e.g. NLSEmagic1D_CPU.c:417, NLSEmagic2D_CPU.c:781-824);
 *   - 3D single precision implements BC=3 (Lap0) as the double version does;
 *     the reference's NLSE3D_TAKE_STEPS_2SHOC_F.c:357-383 lacks a `break`
 *     and falls into the zero default.
 *
 * Layout: split real/imaginary planes, MATLAB column-major order.
 *   1D: U[i], i < N.
 *   2D: U[i*M + j], i < N (x, slow), j < M (y, contiguous).
 *   3D: U[(i*N + j)*M + k], i < L (slow), j < N, k < M (contiguous).
 * method: 1 = CD (2nd order), 2 = 2SHOC (4th order compact).
 * BC: 1 Dirichlet, 2 MSD, 3 Lap0, 4 one-sided (1D only), other -> zero.
 */
#ifndef NLSE_ORACLE_H
#define NLSE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* chunk_size RK4 steps; argument meaning of NLSE1D_TAKE_STEPS
 * (src/NLSEmagic1D/NLSEmagic.h:7-8) */
void oracle_nlse1d(const double *Ur, const double *Ui, const double *V,
                   double *Unr, double *Uni, double s, double a, double h2,
                   int BC, int chunk_size, double k, int N, int method);
/* NLSE1D_TAKE_STEPS_F (NLSEmagic.h:10-11): float in, float out */
void oracle_nlse1d_f(const float *Ur, const float *Ui, const float *V,
                     float *Unr, float *Uni, float s, float a, float h2,
                     int BC, int chunk_size, float k, int N, int method);
/* NLSE2D_TAKE_STEPS (NLSEmagic.h:19-20); N columns (x), M rows (y) */
void oracle_nlse2d(const double *Ur, const double *Ui, const double *V,
                   double *Unr, double *Uni, double s, double a, double h2,
                   int BC, int chunk_size, double k, int N, int M, int method);
/* NLSE2D_TAKE_STEPS_F (NLSEmagic.h:22-23): float in, double out */
void oracle_nlse2d_f(const float *Ur, const float *Ui, const float *V,
                     double *Unr, double *Uni, float s, float a, float h2,
                     int BC, int chunk_size, float k, int N, int M, int method);
/* NLSE3D_TAKE_STEPS (NLSEmagic.h:31-32); dims M x N x L */
void oracle_nlse3d(const double *Ur, const double *Ui, const double *V,
                   double *Unr, double *Uni, double s, double a, double h2,
                   int BC, int chunk_size, double k, int N, int M, int L,
                   int method);
/* NLSE3D_TAKE_STEPS_F (NLSEmagic.h:34-35): float in, double out */
void oracle_nlse3d_f(const float *Ur, const float *Ui, const float *V,
                     double *Unr, double *Uni, float s, float a, float h2,
                     int BC, int chunk_size, float k, int N, int M, int L,
                     int method);

/* One evaluation of the right-hand side F(U) = U_t, including the boundary
 * values, in double precision (used by the boundary-condition checks). */
void oracle_rhs1d(const double *Ur, const double *Ui, const double *V,
                  double *Fr, double *Fi, double s, double a, double h2,
                  int BC, int N, int method);
void oracle_rhs2d(const double *Ur, const double *Ui, const double *V,
                  double *Fr, double *Fi, double s, double a, double h2,
                  int BC, int N, int M, int method);
void oracle_rhs3d(const double *Ur, const double *Ui, const double *V,
                  double *Fr, double *Fi, double s, double a, double h2,
                  int BC, int N, int M, int L, int method);

/* Diagnostics on a state: sum |U|^2 in double (mass without the h^d factor) */
double oracle_sum_abs2(const double *Ur, const double *Ui, long n);

/* 3D vorticity of a state (src/NLSEmagic3D/rmc_vorticity3D.c:24-354):
 * wx, wy, wz = curl of V = Im(conj(U) grad U) / (2|U|^2 + veps), W = |w|;
 * lh2 = 1/(2h); dims M x N x L as oracle_nlse3d (nlse_vorticity.c). */
void oracle_vorticity3d(const double *Ur, const double *Ui, double lh2, double veps, int M, int N, int L,
                        double *wx, double *wy, double *wz, double *W);
/* the same with the unset edge/corner gradient given the value edge_g */
void oracle_vorticity3d_fill(const double *Ur, const double *Ui, double lh2, double veps, int M, int N, int L,
                             double edge_g, double *wx, double *wy, double *wz, double *W);

#ifdef __cplusplus
}
#endif
#endif
