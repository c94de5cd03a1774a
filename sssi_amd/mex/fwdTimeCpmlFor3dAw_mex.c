/*
 * [data, snapshot] = fwdTimeCpmlFor3dAw_mex(v, source, nDiffOrder, nBoundary, dz, dx, dy, dt [, opts])
 *
 * MEX gateway for src/fwdTimeCpmlFor3dAw.m:1 (same arguments and outputs;
 * the reference has only the MATLAB function): v (nz,nx,ny); source
 * (nz,nx,ny,nt) dense, or a struct {iz, ix, iy, wavelet(nt x nsrc)} of
 * 0-based points; data (nx,ny,nt); snapshot (nz,nx,ny,nt) -- only when
 * requested.
 */
#include "sssi_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mxArray *opts = nrhs > 8 ? prhs[8] : NULL;
    int64_t stride, nz, nx, ny, nt;
    int N, nB, rc;
    double dz, dx, dy, dt;
    const mwSize *vd;
    mwSize ddims[3], grid[3];
    mxArray *snap;

    if (nrhs < 8 || nrhs > 9) mexErrMsgIdAndTxt("sssi:nargin", "8 or 9 input arguments expected");
    if (nlhs > 2) mexErrMsgIdAndTxt("sssi:nargout", "at most 2 outputs");
    sssi_mex_check_double(prhs[0], "v");
    if (mxGetNumberOfDimensions(prhs[0]) > 3) mexErrMsgIdAndTxt("sssi:size", "v must be 3-D");
    vd = mxGetDimensions(prhs[0]);
    nz = (int64_t)vd[0];
    nx = (int64_t)vd[1];
    ny = mxGetNumberOfDimensions(prhs[0]) == 3 ? (int64_t)vd[2] : 1;
    N = sssi_mex_int_arg(prhs[2], "nDiffOrder", 1, 8);
    nB = sssi_mex_int_arg(prhs[3], "nBoundary", 0, 1 << 20);
    dz = sssi_mex_scalar(prhs[4], "dz");
    dx = sssi_mex_scalar(prhs[5], "dx");
    dy = sssi_mex_scalar(prhs[6], "dy");
    dt = sssi_mex_scalar(prhs[7], "dt");
    stride = sssi_mex_opt_int(opts, "snapStride", 1);
    if (stride < 1) mexErrMsgIdAndTxt("sssi:range", "opts.snapStride must be >= 1");
    grid[0] = (mwSize)nz;
    grid[1] = (mwSize)nx;
    grid[2] = (mwSize)ny;
    ddims[0] = (mwSize)nx;
    ddims[1] = (mwSize)ny;

    if (mxIsStruct(prhs[1])) {
        sssi_mex_points pts;
        sssi_mex_read_points(prhs[1], 1, &pts);
        nt = pts.nt;
        ddims[2] = (mwSize)nt;
        plhs[0] = mxCreateNumericArray(3, ddims, mxDOUBLE_CLASS, mxREAL);
        snap = sssi_mex_snapshot(nlhs, 3, grid, nt, stride);
        rc = sssi_fwd_cpml_3d_points(mxGetPr(prhs[0]), nz, nx, ny, nt, pts.n, pts.iz, pts.ix,
                                     pts.iy, pts.wavelet, N, nB, dz, dx, dy, dt,
                                     mxGetPr(plhs[0]), snap ? mxGetPr(snap) : NULL, stride);
    } else {
        const mwSize *d = mxGetDimensions(prhs[1]);
        mwSize nd = mxGetNumberOfDimensions(prhs[1]);
        sssi_mex_check_double(prhs[1], "source");
        if (nd > 4 || (int64_t)d[0] != nz || (int64_t)d[1] != nx ||
            (nd >= 3 ? (int64_t)d[2] : 1) != ny)
            mexErrMsgIdAndTxt("sssi:size", "source must be (nz, nx, ny, nt) on the grid of v");
        if (stride != 1)
            mexErrMsgIdAndTxt("sssi:opts", "opts.snapStride needs a struct (point) source");
        nt = nd == 4 ? (int64_t)d[3] : 1;
        ddims[2] = (mwSize)nt;
        plhs[0] = mxCreateNumericArray(3, ddims, mxDOUBLE_CLASS, mxREAL);
        snap = sssi_mex_snapshot(nlhs, 3, grid, nt, 1);
        rc = sssi_fwdTimeCpmlFor3dAw(mxGetPr(prhs[0]), nz, nx, ny, mxGetPr(prhs[1]), nt, N, nB,
                                     dz, dx, dy, dt, mxGetPr(plhs[0]),
                                     snap ? mxGetPr(snap) : NULL);
    }
    sssi_mex_fail(rc);
    if (nlhs > 1) plhs[1] = snap;
}
