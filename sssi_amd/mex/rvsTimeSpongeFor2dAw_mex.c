/*
 * [model, snapshot] = rvsTimeSpongeFor2dAw_mex(v, data, nBoundary, dz, dx, dt [, opts])
 *
 * MEX gateway for src/rvsTimeSpongeFor2dAw.m:1 (same arguments and
 * outputs; the reference has only the MATLAB function): v (nz,nx); data
 * (nx,nt) injected at the top row from the last step back; model
 * ((nz+2),(nx+2)); snapshot (nz,nx,nt) -- only when requested.
 */
#include "sssi_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mxArray *opts = nrhs > 6 ? prhs[6] : NULL;
    int64_t stride, nz, nx, nt;
    int nB, rc;
    double dz, dx, dt;
    mwSize grid[2];
    mxArray *snap;

    if (nrhs < 6 || nrhs > 7) mexErrMsgIdAndTxt("sssi:nargin", "6 or 7 input arguments expected");
    if (nlhs > 2) mexErrMsgIdAndTxt("sssi:nargout", "at most 2 outputs");
    sssi_mex_check_double(prhs[0], "v");
    sssi_mex_check_double(prhs[1], "data");
    if (mxGetNumberOfDimensions(prhs[0]) != 2 || mxGetNumberOfDimensions(prhs[1]) != 2)
        mexErrMsgIdAndTxt("sssi:size", "v and data must be 2-D");
    nz = (int64_t)mxGetM(prhs[0]);
    nx = (int64_t)mxGetN(prhs[0]);
    if ((int64_t)mxGetM(prhs[1]) != nx)
        mexErrMsgIdAndTxt("sssi:size", "data must be (nx, nt) with nx = size(v, 2)");
    nt = (int64_t)mxGetN(prhs[1]);
    nB = sssi_mex_int_arg(prhs[2], "nBoundary", 0, 1 << 20);
    dz = sssi_mex_scalar(prhs[3], "dz");
    dx = sssi_mex_scalar(prhs[4], "dx");
    dt = sssi_mex_scalar(prhs[5], "dt");
    stride = sssi_mex_opt_int(opts, "snapStride", 1);
    if (stride < 1) mexErrMsgIdAndTxt("sssi:range", "opts.snapStride must be >= 1");

    grid[0] = (mwSize)nz;
    grid[1] = (mwSize)nx;
    plhs[0] = mxCreateDoubleMatrix((mwSize)(nz + 2), (mwSize)(nx + 2), mxREAL);
    snap = sssi_mex_snapshot(nlhs, 2, grid, nt, stride);
    rc = sssi_rvs_sponge_2d_ex(mxGetPr(prhs[0]), nz, nx, mxGetPr(prhs[1]), nt, nB, dz, dx, dt,
                               mxGetPr(plhs[0]), snap ? mxGetPr(snap) : NULL, stride);
    sssi_mex_fail(rc);
    if (nlhs > 1) plhs[1] = snap;
}
