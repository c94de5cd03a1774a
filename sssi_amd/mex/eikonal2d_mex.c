/*
 * T = eikonal2d_mex(V, dx, sz, sx, iMax)
 *
 * MEX gateway replacing src/fd_cmex/eikonal2d_mex.cpp:27 (same name,
 * arguments and output): V (nz,nx) velocity, dx grid spacing, (sz, sx) the
 * 1-based source cell, iMax iterations of four sweeps; T (nz,nx)
 * traveltimes.  As in the reference, dx, sz, sx and iMax are truncated to
 * integers (eikonal2d_mex.cpp:35-38); a source outside the grid, which the
 * reference writes out of bounds, is an error here.
 */
#include "sssi_mex_common.h"

static int sssi_mex_trunc(const mxArray *a, const char *name, double lo, double hi)
{
    double d = sssi_mex_scalar(a, name);
    if (!(d >= lo && d < hi + 1))
        mexErrMsgIdAndTxt("sssi:range", "%s must lie in [%g, %g]", name, lo, hi);
    return (int)d;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    int64_t nz, nx;
    int sz, sx, iMax, rc;
    double dx;

    if (nrhs != 5) mexErrMsgIdAndTxt("sssi:nargin", "5 input arguments expected");
    if (nlhs > 1) mexErrMsgIdAndTxt("sssi:nargout", "at most 1 output");
    sssi_mex_check_double(prhs[0], "V");
    if (mxGetNumberOfDimensions(prhs[0]) != 2)
        mexErrMsgIdAndTxt("sssi:size", "V must be 2-D");
    nz = (int64_t)mxGetM(prhs[0]);
    nx = (int64_t)mxGetN(prhs[0]);
    if (nz < 2 || nx < 2) mexErrMsgIdAndTxt("sssi:size", "V must be at least 2 x 2");
    dx = (double)sssi_mex_trunc(prhs[1], "dx", 1, 2147483646.0);
    sz = sssi_mex_trunc(prhs[2], "sz", 1, (double)nz);
    sx = sssi_mex_trunc(prhs[3], "sx", 1, (double)nx);
    iMax = sssi_mex_trunc(prhs[4], "iMax", 0, 1 << 24);

    plhs[0] = mxCreateDoubleMatrix((mwSize)nz, (mwSize)nx, mxREAL);
    rc = sssi_eikonal2d(mxGetPr(prhs[0]), nz, nx, dx, sz, sx, iMax, mxGetPr(plhs[0]));
    sssi_mex_fail(rc);
}
