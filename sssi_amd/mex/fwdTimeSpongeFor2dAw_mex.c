/*
 * [data, snapshot] = fwdTimeSpongeFor2dAw_mex(v, source, nBoundary, dz, dx, dt [, opts])
 *
 * MEX gateway for src/fwdTimeSpongeFor2dAw.m:1 (same arguments and
 * outputs; the reference has only the MATLAB function): second-order
 * sponge-boundary propagation.  v (nz,nx); source (nz,nx,nt) dense, or a
 * struct {iz, ix, wavelet(nt x nsrc)} of 0-based points; data (nx,nt);
 * snapshot (nz,nx,nt) -- only when requested.
 */
#include "sssi_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mxArray *opts = nrhs > 6 ? prhs[6] : NULL;
    int64_t stride, nz, nx, nt;
    int nB, rc;
    double dz, dx, dt;
    mxArray *snap;

    if (nrhs < 6 || nrhs > 7) mexErrMsgIdAndTxt("sssi:nargin", "6 or 7 input arguments expected");
    if (nlhs > 2) mexErrMsgIdAndTxt("sssi:nargout", "at most 2 outputs");
    sssi_mex_check_double(prhs[0], "v");
    if (mxGetNumberOfDimensions(prhs[0]) != 2) mexErrMsgIdAndTxt("sssi:size", "v must be 2-D");
    nz = (int64_t)mxGetM(prhs[0]);
    nx = (int64_t)mxGetN(prhs[0]);
    nB = sssi_mex_int_arg(prhs[2], "nBoundary", 0, 1 << 20);
    dz = sssi_mex_scalar(prhs[3], "dz");
    dx = sssi_mex_scalar(prhs[4], "dx");
    dt = sssi_mex_scalar(prhs[5], "dt");
    stride = sssi_mex_opt_int(opts, "snapStride", 1);
    if (stride < 1) mexErrMsgIdAndTxt("sssi:range", "opts.snapStride must be >= 1");

    if (mxIsStruct(prhs[1])) {
        sssi_mex_points pts;
        mwSize grid[2];
        sssi_mex_read_points(prhs[1], 0, &pts);
        nt = pts.nt;
        grid[0] = (mwSize)nz;
        grid[1] = (mwSize)nx;
        plhs[0] = mxCreateDoubleMatrix((mwSize)nx, (mwSize)nt, mxREAL);
        snap = sssi_mex_snapshot(nlhs, 2, grid, nt, stride);
        rc = sssi_fwd_sponge_2d_points(mxGetPr(prhs[0]), nz, nx, nt, pts.n, pts.iz, pts.ix,
                                       pts.wavelet, nB, dz, dx, dt, mxGetPr(plhs[0]),
                                       snap ? mxGetPr(snap) : NULL, stride);
    } else {
        const mwSize *d = mxGetDimensions(prhs[1]);
        mwSize nd = mxGetNumberOfDimensions(prhs[1]);
        sssi_mex_check_double(prhs[1], "source");
        if (nd > 3 || (int64_t)d[0] != nz || (int64_t)d[1] != nx)
            mexErrMsgIdAndTxt("sssi:size", "source must be (nz, nx, nt) on the grid of v");
        if (stride != 1)
            mexErrMsgIdAndTxt("sssi:opts", "opts.snapStride needs a struct (point) source");
        nt = nd == 3 ? (int64_t)d[2] : 1;
        plhs[0] = mxCreateDoubleMatrix((mwSize)nx, (mwSize)nt, mxREAL);
        snap = sssi_mex_snapshot(nlhs, 2, d, nt, 1);
        rc = sssi_fwdTimeSpongeFor2dAw(mxGetPr(prhs[0]), nz, nx, mxGetPr(prhs[1]), nt, nB, dz,
                                       dx, dt, mxGetPr(plhs[0]), snap ? mxGetPr(snap) : NULL);
    }
    sssi_mex_fail(rc);
    if (nlhs > 1) plhs[1] = snap;
}
