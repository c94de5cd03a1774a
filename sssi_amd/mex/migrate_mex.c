/*
 * m = migrate_mex(travelTime, xRecGrid, xShotAndRecGrid, shot, dt, nz, ixs)
 *
 * MEX gateway for src/migrate.m:1 (with src/shot2RecTime.m:1; the reference
 * has only the MATLAB functions): travelTime (nz,nx[,ntab]) traveltime
 * tables of the positions xShotAndRecGrid, xRecGrid the receiver positions
 * (each must be in xShotAndRecGrid, as ismember finds it), shot (nt,nx) the
 * record, ixs the table of the shot; m (nz,nx) the image.  A sample index
 * past the record, an indexing error in MATLAB, is an error here too.
 */
#include "sssi_mex_common.h"

/* 1-based integer positions of a double vector */
static int *sssi_mex_positions(const mxArray *a, const char *name, int *n)
{
    size_t k, m;
    const double *p;
    int *out;
    sssi_mex_check_double(a, name);
    m = mxGetNumberOfElements(a);
    if (m < 1 || m > (1u << 30)) mexErrMsgIdAndTxt("sssi:size", "%s must not be empty", name);
    if (mxGetM(a) != 1 && mxGetN(a) != 1)
        mexErrMsgIdAndTxt("sssi:size", "%s must be a vector", name);
    p = mxGetPr(a);
    out = (int *)mxCalloc(m, sizeof(int));
    for (k = 0; k < m; ++k) {
        if (p[k] != (double)(int)p[k] || p[k] < 1)
            mexErrMsgIdAndTxt("sssi:range", "%s must hold positive integer positions", name);
        out[k] = (int)p[k];
    }
    *n = (int)m;
    return out;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    const mwSize *d;
    int64_t nz, nx, ntab, nt;
    int nrec, nsr, ixs, rc;
    int *xrec, *xsr;
    double dt;

    if (nrhs != 7) mexErrMsgIdAndTxt("sssi:nargin", "7 input arguments expected");
    if (nlhs > 1) mexErrMsgIdAndTxt("sssi:nargout", "at most 1 output");
    sssi_mex_check_double(prhs[0], "travelTime");
    sssi_mex_check_double(prhs[3], "shot");
    if (mxGetNumberOfDimensions(prhs[0]) > 3)
        mexErrMsgIdAndTxt("sssi:size", "travelTime must be (nz, nx, ntab)");
    d = mxGetDimensions(prhs[0]);
    nz = (int64_t)d[0];
    nx = (int64_t)d[1];
    ntab = mxGetNumberOfDimensions(prhs[0]) == 3 ? (int64_t)d[2] : 1;
    if (mxGetNumberOfDimensions(prhs[3]) != 2 || (int64_t)mxGetN(prhs[3]) != nx)
        mexErrMsgIdAndTxt("sssi:size", "shot must be (nt, nx) with nx = size(travelTime, 2)");
    nt = (int64_t)mxGetM(prhs[3]);
    dt = sssi_mex_scalar(prhs[4], "dt");
    if ((int64_t)sssi_mex_scalar(prhs[5], "nz") != nz)
        mexErrMsgIdAndTxt("sssi:size", "nz must equal size(travelTime, 1)");
    ixs = sssi_mex_int_arg(prhs[6], "ixs", 1, ntab > 2147483647 ? 2147483647 : (int)ntab);
    xrec = sssi_mex_positions(prhs[1], "xRecGrid", &nrec);
    xsr = sssi_mex_positions(prhs[2], "xShotAndRecGrid", &nsr);

    plhs[0] = mxCreateDoubleMatrix((mwSize)nz, (mwSize)nx, mxREAL);
    rc = sssi_kirchhoff_migrate(mxGetPr(prhs[0]), nz, nx, ntab, xrec, nrec, xsr, nsr,
                                mxGetPr(prhs[3]), nt, dt, ixs, mxGetPr(plhs[0]));
    mxFree(xrec);
    mxFree(xsr);
    sssi_mex_fail(rc);
}
