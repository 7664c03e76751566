/*
 * nlse_mex_vorticity.c -- MATLAB / GNU Octave MEX gateway rmc_vorticity3D
 * over the MI355X vorticity kernels (NLSE3D_VORTICITY_HIP in
 * include/nlsemagic_hip.h).
 *
 *   [wx, wy, wz, W] = rmc_vorticity3D(U, 1/(2*h), veps)
 *
 * replaces the reference MEX file of that name
 * (src/NLSEmagic3D/rmc_vorticity3D.c:24-354), called by the 3D driver at
 * src/NLSEmagic3D/NLSE3D.m:594 and :1109, with the same arguments and
 * outputs: U a real or complex double M x N x L array, lh2 = 1/(2h), veps
 * the regulariser of the density in the velocity's denominator; four real
 * double arrays of U's size.  Unlike the reference the gateway checks its
 * arguments, and the edge and corner points (which the reference computes
 * from unset memory on the faces next to them) are defined: 0.
 */
#include <string.h>

#include "mex.h"
#include "nlsemagic_hip.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mwSize *dv;
    mxArray *arr[4];
    double *out[4];
    char msg[512];
    int o, rc, nout;

    if (nrhs != 3) mexErrMsgTxt("expected 3 arguments: U, 1/(2h), veps");
    if (nlhs > 4) mexErrMsgTxt("at most 4 outputs: wx, wy, wz, W");
    if (!mxIsDouble(prhs[0])) mexErrMsgTxt("U must be a double array");
    if (mxGetNumberOfDimensions(prhs[0]) != 3) mexErrMsgTxt("U must be a 3D array");
    dv = mxGetDimensions(prhs[0]);
    /* plhs holds max(nlhs, 1) outputs; the others are computed and dropped */
    nout = nlhs < 1 ? 1 : nlhs;
    for (o = 0; o < 4; ++o) {
        arr[o] = mxCreateNumericArray(3, dv, mxDOUBLE_CLASS, mxREAL);
        out[o] = mxGetPr(arr[o]);
        if (o < nout) plhs[o] = arr[o];
    }
    rc = NLSE3D_VORTICITY_HIP(mxGetPr(prhs[0]), mxIsComplex(prhs[0]) ? mxGetPi(prhs[0]) : NULL,
                              mxGetScalar(prhs[1]), mxGetScalar(prhs[2]), (int)dv[0], (int)dv[1], (int)dv[2],
                              out[0], out[1], out[2], out[3]);
    for (o = nout; o < 4; ++o) mxDestroyArray(arr[o]);
    if (rc) {
        msg[0] = 0;
        strncat(msg, "nlsemagic_hip: ", sizeof(msg) - 1);
        strncat(msg, nlse_last_error(), sizeof(msg) - strlen(msg) - 1);
        mexErrMsgTxt(msg);
    }
}
