/*
 * mex.h -- TEST-ONLY stand-in for the MATLAB/Octave MEX API, just large
 * enough to compile and drive nlsemagic_amd/mex/nlse_mex_gateway.c on a
 * machine without MATLAB or Octave (tests/test_mex_gateway.py).  It is not
 * used to build anything of the reference.
 */
#ifndef NLSE_TEST_MEX_H
#define NLSE_TEST_MEX_H
#include <stddef.h>

typedef size_t mwSize;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef enum { mxDOUBLE_CLASS = 6 } mxClassID;
typedef struct mxArray_tag {
    int ndims;
    mwSize dims[3];
    int is_complex;
    double *pr, *pi;
} mxArray;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

int mxIsDouble(const mxArray *a);
int mxIsComplex(const mxArray *a);
mwSize mxGetM(const mxArray *a);
mwSize mxGetN(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double *mxGetPr(const mxArray *a);
double *mxGetPi(const mxArray *a);
double mxGetScalar(const mxArray *a);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateNumericArray(mwSize nd, const mwSize *dims, mxClassID cls, mxComplexity c);
void mxDestroyArray(mxArray *a);
void mexErrMsgTxt(const char *msg);
#endif
