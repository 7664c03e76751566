/*
 * nlse_mex_gateway.c -- MATLAB / GNU Octave MEX gateway over the MI355X
 * integrators (libnlsemagic_hip.so, include/nlsemagic_hip.h).
 *
 * One source, compiled once per MATLAB-visible name:
 *
 *   v1, 8 arguments   U = NAME(U, V, s, a, h2, BC, chunk_size, k)
 *       -DNLSE_DIM=d -DNLSE_METHOD=m -DNLSE_PREC=p
 *       names NLSE{d}D_TAKE_STEPS_{CD|2SHOC}_CUDA_{D|F} replace the reference
 *       MEX files of those names (e.g. src_old/NLSE2D_v1/
 *       NLSE2D_TAKE_STEPS_2SHOC_CUDA_D.cu:322-441), so the reference drivers
 *       (src/NLSEmagic2D/NLSEmagic2D.m:256-279, NLSEmagic3D.m:303-325) run
 *       unchanged;
 *   v2, 10 arguments  U = NAME(U, V, s, a, h2, BC, chunk_size, k, precision, method)
 *       -DNLSE_DIM=d -DNLSE_V2
 *       name NLSEmagic{d}D_Take_Steps_GPU replaces
 *       src/NLSEmagic1D/NLSEmagic1D_Take_Steps_GPU.c:5-90 and
 *       src/NLSEmagic2D/NLSEmagic2D_Take_Steps_GPU.c:5-78.
 *
 * Argument meaning is the reference's: U real or complex double (a real U
 * gets a zero imaginary part), V real double of U's size, h2 = h^2, BC 1-4,
 * precision 1 single / 2 double, method 1 CD / 2 2SHOC.  The result is a new
 * complex double array of U's shape (a column for 1D, as the reference).
 * Single precision rounds U, V and the scalars to float and widens the result
 * back (NLSEmagic1D_Take_Steps_GPU.c:50-79).  Unlike the reference the
 * gateway checks its arguments and reports errors with mexErrMsgTxt.
 *
 * Uses the separate-complex MEX API (mxGetPi), which Octave's mkoctfile
 * --mex and MATLAB's default (-R2017b) API provide.  One build per MATLAB
 * name, e.g. NLSE2D_TAKE_STEPS_2SHOC_CUDA_D:
 *   mkoctfile --mex -DNLSE_DIM=2 -DNLSE_METHOD=2 -DNLSE_PREC=2 -Iinclude \
 *     nlse_mex_gateway.c -Lnlsemagic_amd -lnlsemagic_hip -o NLSE2D_TAKE_STEPS_2SHOC_CUDA_D.mex
 * and NLSEmagic2D_Take_Steps_GPU with -DNLSE_DIM=2 -DNLSE_V2.
 */
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "nlsemagic_hip.h"

#ifndef NLSE_DIM
#error "define NLSE_DIM (1, 2 or 3)"
#endif
#ifdef NLSE_V2
#define NLSE_NARGS 10
#else
#define NLSE_NARGS 8
#ifndef NLSE_METHOD
#error "define NLSE_METHOD (1 CD, 2 2SHOC) for an 8-argument gateway"
#endif
#ifndef NLSE_PREC
#error "define NLSE_PREC (1 single, 2 double) for an 8-argument gateway"
#endif
#endif

static void fail_rc(int rc)
{
    char msg[512];
    msg[0] = 0;
    strncat(msg, "nlsemagic_hip: ", sizeof(msg) - 1);
    strncat(msg, nlse_last_error(), sizeof(msg) - strlen(msg) - 1);
    (void)rc;
    mexErrMsgTxt(msg);
}

static int run_double(const double *Ur, const double *Ui, const double *V, double *Or, double *Oi, double s,
                      double a, double h2, int BC, int chunk, double k, int N, int M, int L, int method)
{
#if NLSE_DIM == 1
    (void)M; (void)L;
    return NLSE1D_TAKE_STEPS_HIP(Ur, Ui, V, Or, Oi, s, a, h2, BC, chunk, k, N, method);
#elif NLSE_DIM == 2
    (void)L;
    return NLSE2D_TAKE_STEPS_HIP(Ur, Ui, V, Or, Oi, s, a, h2, BC, chunk, k, N, M, method);
#else
    return NLSE3D_TAKE_STEPS_HIP(Ur, Ui, V, Or, Oi, s, a, h2, BC, chunk, k, N, M, L, method);
#endif
}

static int run_single(const float *Ur, const float *Ui, const float *V, float *Or, float *Oi, float s, float a,
                      float h2, int BC, int chunk, float k, int N, int M, int L, int method)
{
#if NLSE_DIM == 1
    (void)M; (void)L;
    return NLSE1D_TAKE_STEPS_HIP_F(Ur, Ui, V, Or, Oi, s, a, h2, BC, chunk, k, N, method);
#elif NLSE_DIM == 2
    (void)L;
    return NLSE2D_TAKE_STEPS_HIP_F(Ur, Ui, V, Or, Oi, s, a, h2, BC, chunk, k, N, M, method);
#else
    return NLSE3D_TAKE_STEPS_HIP_F(Ur, Ui, V, Or, Oi, s, a, h2, BC, chunk, k, N, M, L, method);
#endif
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    int N, M = 1, L = 1, BC, chunk, precision, method, rc;
    size_t n, p;
    double s, a, h2, k;
    const double *Ur, *Ui, *V;
    double *Or, *Oi;
    mwSize ndims;
    const mwSize *dv;

    (void)nlhs;
    if (nrhs != NLSE_NARGS) mexErrMsgTxt(NLSE_NARGS == 8 ? "expected 8 arguments: U,V,s,a,h2,BC,chunk_size,k"
                                                         : "expected 10 arguments: U,V,s,a,h2,BC,chunk_size,k,"
                                                           "precision,method");
    if (!mxIsDouble(prhs[0]) || !mxIsDouble(prhs[1]) || mxIsComplex(prhs[1]))
        mexErrMsgTxt("U must be a double array and V a real double array");
    ndims = mxGetNumberOfDimensions(prhs[0]);
    dv = mxGetDimensions(prhs[0]);
#if NLSE_DIM == 1
    (void)ndims;
    (void)dv;
    N = (int)(mxGetN(prhs[0]) == 1 ? mxGetM(prhs[0]) : mxGetN(prhs[0]));
    n = (size_t)N;
    if (mxGetM(prhs[0]) != 1 && mxGetN(prhs[0]) != 1) mexErrMsgTxt("U must be a vector");
#elif NLSE_DIM == 2
    if (ndims != 2) mexErrMsgTxt("U must be a 2D array");
    M = (int)dv[0];
    N = (int)dv[1];
    n = (size_t)M * N;
#else
    if (ndims != 3) mexErrMsgTxt("U must be a 3D array");
    M = (int)dv[0];
    N = (int)dv[1];
    L = (int)dv[2];
    n = (size_t)M * N * L;
#endif
    if ((size_t)mxGetNumberOfElements(prhs[1]) != n) mexErrMsgTxt("V must have as many elements as U");
    s = mxGetScalar(prhs[2]);
    a = mxGetScalar(prhs[3]);
    h2 = mxGetScalar(prhs[4]);
    BC = (int)mxGetScalar(prhs[5]);
    chunk = (int)mxGetScalar(prhs[6]);
    k = mxGetScalar(prhs[7]);
#ifdef NLSE_V2
    precision = (int)mxGetScalar(prhs[8]);
    method = (int)mxGetScalar(prhs[9]);
#else
    precision = NLSE_PREC;
    method = NLSE_METHOD;
#endif
    Ur = mxGetPr(prhs[0]);
    Ui = mxIsComplex(prhs[0]) ? mxGetPi(prhs[0]) : NULL;
    V = mxGetPr(prhs[1]);

#if NLSE_DIM == 1
    plhs[0] = mxCreateDoubleMatrix((mwSize)N, 1, mxCOMPLEX);
#elif NLSE_DIM == 2
    plhs[0] = mxCreateDoubleMatrix((mwSize)M, (mwSize)N, mxCOMPLEX);
#else
    plhs[0] = mxCreateNumericArray(3, dv, mxDOUBLE_CLASS, mxCOMPLEX);
#endif
    Or = mxGetPr(plhs[0]);
    Oi = mxGetPi(plhs[0]);

    if (precision == 2) {
        rc = run_double(Ur, Ui, V, Or, Oi, s, a, h2, BC, chunk, k, N, M, L, method);
        if (rc) fail_rc(rc);
    } else if (precision == 1) {
        float *f = (float *)malloc(sizeof(float) * n * 5);
        if (!f) mexErrMsgTxt("out of host memory");
        for (p = 0; p < n; ++p) {
            f[p] = (float)Ur[p];
            f[n + p] = Ui ? (float)Ui[p] : 0.0f;
            f[2 * n + p] = (float)V[p];
        }
        rc = run_single(f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, (float)s, (float)a, (float)h2, BC, chunk,
                        (float)k, N, M, L, method);
        if (!rc)
            for (p = 0; p < n; ++p) {
                Or[p] = (double)f[3 * n + p];
                Oi[p] = (double)f[4 * n + p];
            }
        free(f);
        if (rc) fail_rc(rc);
    } else {
        mexErrMsgTxt("precision must be 1 (single) or 2 (double)");
    }
}
