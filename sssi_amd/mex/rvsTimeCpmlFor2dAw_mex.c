/*
 * [model, snapshot] = rvsTimeCpmlFor2dAw_mex(v, data, nDiffOrder, nBoundary, dz, dx, dt [, opts])
 *
 * MEX gateway in place of src/fd_cmex/rvsTimeCpmlFor2dAw_mex.c:38 (same
 * name, arguments and outputs): v (nz,nx); data (nx,nt) injected at the top
 * row from the last step back; model ((nz+2l),(nx+2l)), l = 2*nDiffOrder-1,
 * the padded field one step before the last (rvsTimeCpmlFor2dAw_mex.c:285);
 * snapshot (nz,nx,nt) -- only when requested.
 */
#include "sssi_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mxArray *opts = nrhs > 7 ? prhs[7] : NULL;
    int64_t stride, nz, nx, nt;
    int N, nB, l, rc;
    double dz, dx, dt;
    mwSize grid[2];
    mxArray *snap;

    if (nrhs < 7 || nrhs > 8) mexErrMsgIdAndTxt("sssi:nargin", "7 or 8 input arguments expected");
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
    N = sssi_mex_int_arg(prhs[2], "nDiffOrder", 1, 8);
    nB = sssi_mex_int_arg(prhs[3], "nBoundary", 0, 1 << 20);
    dz = sssi_mex_scalar(prhs[4], "dz");
    dx = sssi_mex_scalar(prhs[5], "dx");
    dt = sssi_mex_scalar(prhs[6], "dt");
    stride = sssi_mex_opt_int(opts, "snapStride", 1);
    if (stride < 1) mexErrMsgIdAndTxt("sssi:range", "opts.snapStride must be >= 1");

    l = 2 * N - 1;
    grid[0] = (mwSize)nz;
    grid[1] = (mwSize)nx;
    plhs[0] = mxCreateDoubleMatrix((mwSize)(nz + 2 * l), (mwSize)(nx + 2 * l), mxREAL);
    snap = sssi_mex_snapshot(nlhs, 2, grid, nt, stride);
    rc = sssi_rvs_cpml_2d_ex(mxGetPr(prhs[0]), nz, nx, mxGetPr(prhs[1]), nt, N, nB, dz, dx, dt,
                             mxGetPr(plhs[0]), snap ? mxGetPr(snap) : NULL, stride);
    sssi_mex_fail(rc);
    if (nlhs > 1) plhs[1] = snap;
}
