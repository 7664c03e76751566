/*
 * mex_shim.c -- TEST-ONLY implementation of tests/mex_shim/mex.h plus
 * shim_call(), a C entry point that wraps raw buffers in mxArrays, calls the
 * gateway's mexFunction and copies the result out (errors come back through
 * longjmp as the message text).
 */
#include <setjmp.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

static jmp_buf g_jmp;
static char g_msg[512];

int mxIsDouble(const mxArray *a) { (void)a; return 1; }
int mxIsComplex(const mxArray *a) { return a->is_complex; }
mwSize mxGetM(const mxArray *a) { return a->dims[0]; }
mwSize mxGetN(const mxArray *a)
{
    mwSize n = 1;
    for (int d = 1; d < a->ndims; ++d) n *= a->dims[d];
    return n;
}
mwSize mxGetNumberOfDimensions(const mxArray *a) { return (mwSize)a->ndims; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->dims[0] * mxGetN(a); }
double *mxGetPr(const mxArray *a) { return a->pr; }
double *mxGetPi(const mxArray *a) { return a->pi; }
double mxGetScalar(const mxArray *a) { return a->pr[0]; }

static mxArray *make(int nd, const mwSize *dims, int cplx)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->ndims = nd;
    for (int d = 0; d < 3; ++d) a->dims[d] = d < nd ? dims[d] : 1;
    size_t n = mxGetNumberOfElements(a);
    a->is_complex = cplx;
    a->pr = (double *)calloc(n, sizeof(double));
    a->pi = cplx ? (double *)calloc(n, sizeof(double)) : NULL;
    return a;
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c)
{
    mwSize d[2] = {m, n};
    return make(2, d, c == mxCOMPLEX);
}
mxArray *mxCreateNumericArray(mwSize nd, const mwSize *dims, mxClassID cls, mxComplexity c)
{
    (void)cls;
    return make((int)nd, dims, c == mxCOMPLEX);
}
void mexErrMsgTxt(const char *msg)
{
    strncpy(g_msg, msg, sizeof(g_msg) - 1);
    g_msg[sizeof(g_msg) - 1] = 0;
    longjmp(g_jmp, 1);
}

static void release(mxArray *a)
{
    if (!a) return;
    free(a->pr);
    free(a->pi);
    free(a);
}
void mxDestroyArray(mxArray *a) { release(a); }

/* returns 0 and fills outr/outi (numel of U), or 1 and the error text */
int shim_call(int nd, const size_t *dims, const double *ur, const double *ui, const double *v,
              const double *scalars, int nscalars, double *outr, double *outi, char *err, int errlen)
{
    mxArray *args[16] = {0};
    mxArray *out[1] = {0};
    int rc = 0;
    mwSize d[3];
    for (int i = 0; i < nd; ++i) d[i] = dims[i];
    args[0] = make(nd, d, ui != NULL);
    size_t n = mxGetNumberOfElements(args[0]);
    memcpy(args[0]->pr, ur, n * sizeof(double));
    if (ui) memcpy(args[0]->pi, ui, n * sizeof(double));
    args[1] = make(nd, d, 0);
    memcpy(args[1]->pr, v, n * sizeof(double));
    for (int i = 0; i < nscalars; ++i) {
        mwSize one[2] = {1, 1};
        args[2 + i] = make(2, one, 0);
        args[2 + i]->pr[0] = scalars[i];
    }
    if (setjmp(g_jmp) == 0) {
        mexFunction(1, out, 2 + nscalars, (const mxArray **)args);
        memcpy(outr, out[0]->pr, n * sizeof(double));
        memcpy(outi, out[0]->pi, n * sizeof(double));
    } else {
        strncpy(err, g_msg, (size_t)errlen - 1);
        err[errlen - 1] = 0;
        rc = 1;
    }
    for (int i = 0; i < 2 + nscalars; ++i) release(args[i]);
    release(out[0]);
    return rc;
}

/* a gateway of real outputs: U (ur, ui or NULL), then nscalars scalars;
 * max(nlhs, 1) outputs of U's numel (MATLAB's ans when nlhs = 0) copied to
 * outs[]; 0 or 1 and the error */
int shim_call_real(int nd, const size_t *dims, const double *ur, const double *ui, const double *scalars,
                   int nscalars, int nlhs, double **outs, char *err, int errlen)
{
    mxArray *args[16] = {0};
    mxArray *out[8] = {0};
    int rc = 0;
    mwSize d[3];
    for (int i = 0; i < nd; ++i) d[i] = dims[i];
    args[0] = make(nd, d, ui != NULL);
    size_t n = mxGetNumberOfElements(args[0]);
    memcpy(args[0]->pr, ur, n * sizeof(double));
    if (ui) memcpy(args[0]->pi, ui, n * sizeof(double));
    for (int i = 0; i < nscalars; ++i) {
        mwSize one[2] = {1, 1};
        args[1 + i] = make(2, one, 0);
        args[1 + i]->pr[0] = scalars[i];
    }
    if (setjmp(g_jmp) == 0) {
        mexFunction(nlhs, out, 1 + nscalars, (const mxArray **)args);
        for (int o = 0; o < (nlhs < 1 ? 1 : nlhs); ++o) memcpy(outs[o], out[o]->pr, n * sizeof(double));
    } else {
        strncpy(err, g_msg, (size_t)errlen - 1);
        err[errlen - 1] = 0;
        rc = 1;
    }
    for (int i = 0; i < 1 + nscalars; ++i) release(args[i]);
    for (int o = 0; o < 8; ++o) release(out[o]);
    return rc;
}
