/*
 * nlse_oracle.c -- CPU restatement of NLSEmagic's RK4 chunk integrators.
 * TEST INFRASTRUCTURE ONLY (see nlse_oracle.h).  Build with
 * -O2 -ffp-contract=off (oracle/Makefile).
 */
#include <stdlib.h>
#include <string.h>
#include "nlse_oracle.h"

#define R double
#define FN(x) x##_d
#include "nlse_oracle_body.inc"
#undef R
#undef FN

#define R float
#define FN(x) x##_f
#include "nlse_oracle_body.inc"
#undef R
#undef FN

/* ---- scalar precompute, per file, in the precision each file uses ---- */

/* NLSEmagic1D_CPU.c:384-390 */
static void coefs1d_d(coefs_d *c, double s, double a, double h2, double k)
{
    c->s = s; c->a = a; c->k = k;
    c->ah2 = a / h2;
    c->k2 = k / 2.0;
    c->k6 = k / 6.0;
    c->l_a = 1.0 / a;
    c->lh2 = 1.0 / h2;
    c->a76 = a * (7.0 / 6.0);
    c->a112 = a * (1.0 / 12.0);
}

/* NLSEmagic1D_CPU.c:460-466: all in float */
static void coefs1d_f(coefs_f *c, float s, float a, float h2, float k)
{
    c->s = s; c->a = a; c->k = k;
    c->ah2 = a / h2;
    c->k2 = k / 2.0f;
    c->k6 = k / 6.0f;
    c->l_a = 1.0f / a;
    c->lh2 = 1.0f / h2;
    c->a76 = a * (7.0f / 6.0f);
    c->a112 = a * (1.0f / 12.0f);
}

/* NLSEmagic2D_CPU.c:757-763 */
static void coefs2d_d(coefs_d *c, double s, double a, double h2, double k)
{
    c->s = s; c->a = a; c->k = k;
    c->l_a = 1.0 / a;
    c->lh2 = 1.0 / h2;
    c->k2 = k / 2.0;
    c->k6 = k / 6.0;
    c->a1_6h2 = a * (1.0 / (6.0 * h2));
    c->a1_12 = a * (1.0 / 12.0);
    c->ah2 = a / h2;
}

/* NLSEmagic2D_CPU.c:865-871: double expressions rounded to float, except
 * ah2 = a/h2 which is a float division */
static void coefs2d_f(coefs_f *c, float s, float a, float h2, float k)
{
    double ad = a, h2d = h2, kd = k;
    c->s = s; c->a = a; c->k = k;
    c->l_a = (float)(1.0 / ad);
    c->lh2 = (float)(1.0 / h2d);
    c->k2 = (float)(kd / 2.0);
    c->k6 = (float)(kd / 6.0);
    c->a1_6h2 = (float)(ad * (1.0 / (6.0 * h2d)));
    c->a1_12 = (float)(ad * (1.0 / 12.0));
    c->ah2 = a / h2;
}

/* NLSE3D_TAKE_STEPS_CD.c:304-306 and NLSE3D_TAKE_STEPS_2SHOC.c:499-504 */
static void coefs3d_d(coefs_d *c, double s, double a, double h2, double k)
{
    c->s = s; c->a = a; c->k = k;
    c->ah2 = a / h2;
    c->k2 = k / 2.0;
    c->k6 = k / 6.0;
    c->lh2 = 1.0 / h2;
    c->a_6h2 = a / (6.0 * h2);
    c->a_12 = a / 12.0;
    c->l_a = 1.0 / a;
}

/* NLSE3D_TAKE_STEPS_CD_F.c:315-317 and NLSE3D_TAKE_STEPS_2SHOC_F.c:509-514 */
static void coefs3d_f(coefs_f *c, float s, float a, float h2, float k)
{
    c->s = s; c->a = a; c->k = k;
    c->ah2 = a / h2;
    c->k2 = k / 2.0f;
    c->k6 = k / 6.0f;
    c->lh2 = 1.0f / h2;
    c->a_6h2 = a / (6.0f * h2);
    c->a_12 = a / 12.0f;
    c->l_a = 1.0f / a;
}

/* ---- drivers: copy the input, integrate, copy out ---- */

static void run_d(int dim, const coefs_d *c, const double *Ur, const double *Ui,
                  const double *V, double *Unr, double *Uni, int BC, int chunk,
                  long N, long M, long L, int method)
{
    long n = N * M * L;
    double *wr = (double *)malloc(sizeof(double) * (size_t)n * 2), *wi = wr + n;
    memcpy(wr, Ur, sizeof(double) * (size_t)n);
    memcpy(wi, Ui, sizeof(double) * (size_t)n);
    chunk_d(dim, wr, wi, V, c, BC, chunk, N, M, L, method);
    memcpy(Unr, wr, sizeof(double) * (size_t)n);
    memcpy(Uni, wi, sizeof(double) * (size_t)n);
    free(wr);
}

static float *run_f(int dim, const coefs_f *c, const float *Ur, const float *Ui,
                    const float *V, int BC, int chunk, long N, long M, long L, int method)
{
    long n = N * M * L;
    float *w = (float *)malloc(sizeof(float) * (size_t)n * 2);
    memcpy(w, Ur, sizeof(float) * (size_t)n);
    memcpy(w + n, Ui, sizeof(float) * (size_t)n);
    chunk_f(dim, w, w + n, V, c, BC, chunk, N, M, L, method);
    return w;
}

void oracle_nlse1d(const double *Ur, const double *Ui, const double *V, double *Unr,
                   double *Uni, double s, double a, double h2, int BC, int chunk_size,
                   double k, int N, int method)
{
    coefs_d c;
    coefs1d_d(&c, s, a, h2, k);
    run_d(1, &c, Ur, Ui, V, Unr, Uni, BC, chunk_size, N, 1, 1, method);
}

void oracle_nlse1d_f(const float *Ur, const float *Ui, const float *V, float *Unr,
                     float *Uni, float s, float a, float h2, int BC, int chunk_size,
                     float k, int N, int method)
{
    coefs_f c;
    float *w;
    coefs1d_f(&c, s, a, h2, k);
    w = run_f(1, &c, Ur, Ui, V, BC, chunk_size, N, 1, 1, method);
    memcpy(Unr, w, sizeof(float) * (size_t)N);
    memcpy(Uni, w + N, sizeof(float) * (size_t)N);
    free(w);
}

void oracle_nlse2d(const double *Ur, const double *Ui, const double *V, double *Unr,
                   double *Uni, double s, double a, double h2, int BC, int chunk_size,
                   double k, int N, int M, int method)
{
    coefs_d c;
    coefs2d_d(&c, s, a, h2, k);
    run_d(2, &c, Ur, Ui, V, Unr, Uni, BC, chunk_size, N, M, 1, method);
}

void oracle_nlse2d_f(const float *Ur, const float *Ui, const float *V, double *Unr,
                     double *Uni, float s, float a, float h2, int BC, int chunk_size,
                     float k, int N, int M, int method)
{
    coefs_f c;
    long n = (long)N * M, p;
    float *w;
    coefs2d_f(&c, s, a, h2, k);
    w = run_f(2, &c, Ur, Ui, V, BC, chunk_size, N, M, 1, method);
    for (p = 0; p < n; p++) { Unr[p] = (double)w[p]; Uni[p] = (double)w[n + p]; }
    free(w);
}

void oracle_nlse3d(const double *Ur, const double *Ui, const double *V, double *Unr,
                   double *Uni, double s, double a, double h2, int BC, int chunk_size,
                   double k, int N, int M, int L, int method)
{
    coefs_d c;
    coefs3d_d(&c, s, a, h2, k);
    run_d(3, &c, Ur, Ui, V, Unr, Uni, BC, chunk_size, N, M, L, method);
}

void oracle_nlse3d_f(const float *Ur, const float *Ui, const float *V, double *Unr,
                     double *Uni, float s, float a, float h2, int BC, int chunk_size,
                     float k, int N, int M, int L, int method)
{
    coefs_f c;
    long n = (long)N * M * L, p;
    float *w;
    coefs3d_f(&c, s, a, h2, k);
    w = run_f(3, &c, Ur, Ui, V, BC, chunk_size, N, M, L, method);
    for (p = 0; p < n; p++) { Unr[p] = (double)w[p]; Uni[p] = (double)w[n + p]; }
    free(w);
}

static void rhs_once(int dim, const coefs_d *c, const double *Ur, const double *Ui,
                     const double *V, double *Fr, double *Fi, int BC, long N, long M,
                     long L, int method)
{
    long n = N * M * L;
    double *d = (double *)malloc(sizeof(double) * (size_t)n * 2);
    rhs_d(dim, Fr, Fi, Ur, Ui, V, d, d + n, c, BC, N, M, L, method);
    free(d);
}

void oracle_rhs1d(const double *Ur, const double *Ui, const double *V, double *Fr,
                  double *Fi, double s, double a, double h2, int BC, int N, int method)
{
    coefs_d c;
    coefs1d_d(&c, s, a, h2, 0.0);
    rhs_once(1, &c, Ur, Ui, V, Fr, Fi, BC, N, 1, 1, method);
}

void oracle_rhs2d(const double *Ur, const double *Ui, const double *V, double *Fr,
                  double *Fi, double s, double a, double h2, int BC, int N, int M,
                  int method)
{
    coefs_d c;
    coefs2d_d(&c, s, a, h2, 0.0);
    rhs_once(2, &c, Ur, Ui, V, Fr, Fi, BC, N, M, 1, method);
}

void oracle_rhs3d(const double *Ur, const double *Ui, const double *V, double *Fr,
                  double *Fi, double s, double a, double h2, int BC, int N, int M, int L,
                  int method)
{
    coefs_d c;
    coefs3d_d(&c, s, a, h2, 0.0);
    rhs_once(3, &c, Ur, Ui, V, Fr, Fi, BC, N, M, L, method);
}

double oracle_sum_abs2(const double *Ur, const double *Ui, long n)
{
    double acc = 0.0;
    long p;
    for (p = 0; p < n; p++) acc += Ur[p] * Ur[p] + Ui[p] * Ui[p];
    return acc;
}
