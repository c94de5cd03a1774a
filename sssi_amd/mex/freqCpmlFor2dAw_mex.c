/*
 * [A, snapshot] = freqCpmlFor2dAw_mex(model, source, w, nDiffOrder, nBoundary, dz, dx)
 *
 * MEX gateway for src/freqCpmlFor2dAw.m:1 (same arguments and outputs; the
 * reference has only the MATLAB function).  Built with the interleaved
 * complex API (mex -R2018a).  model (nz,nx) squared slowness; source real or
 * complex with nz*nx*nsrc elements (any shape, MATLAB's reshape(source,
 * nLength, [])), or empty; A the (nz*nx)^2 sparse complex impedance matrix;
 * snapshot = A \ (-source), complex, shaped like source ([] for an empty
 * source), solved on the device.
 */
#include "sssi_mex_common.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    int64_t nz, nx, n, nnz = 0, e, c;
    int N, nB, rc, nsrc;
    double w, dz, dx;
    int64_t *rowptr, *col, *cnt;
    double *val;
    mwIndex *jc, *ir;
    mxComplexDouble *av;

    if (nrhs != 7) mexErrMsgIdAndTxt("sssi:nargin", "7 input arguments expected");
    if (nlhs > 2) mexErrMsgIdAndTxt("sssi:nargout", "at most 2 outputs");
    sssi_mex_check_double(prhs[0], "model");
    if (mxGetNumberOfDimensions(prhs[0]) != 2) mexErrMsgIdAndTxt("sssi:size", "model must be 2-D");
    if (!mxIsDouble(prhs[1])) mexErrMsgIdAndTxt("sssi:class", "source must be a double array");
    nz = (int64_t)mxGetM(prhs[0]);
    nx = (int64_t)mxGetN(prhs[0]);
    n = nz * nx;
    w = sssi_mex_scalar(prhs[2], "w");
    N = sssi_mex_int_arg(prhs[3], "nDiffOrder", 1, 8);
    nB = sssi_mex_int_arg(prhs[4], "nBoundary", 0, 1 << 20);
    dz = sssi_mex_scalar(prhs[5], "dz");
    dx = sssi_mex_scalar(prhs[6], "dx");
    if (n < 1) mexErrMsgIdAndTxt("sssi:size", "model must not be empty");
    if (mxGetNumberOfElements(prhs[1]) % (size_t)n)
        mexErrMsgIdAndTxt("sssi:size", "source must hold nz*nx*nsrc elements");
    nsrc = (int)(mxGetNumberOfElements(prhs[1]) / (size_t)n);

    /* A: CSR from the library, turned into MATLAB's compressed columns */
    rc = sssi_freq_cpml_2d_matrix(mxGetPr(prhs[0]), nz, nx, w, N, nB, dz, dx, &nnz, NULL, NULL,
                                  NULL);
    sssi_mex_fail(rc);
    rowptr = (int64_t *)mxCalloc((size_t)n + 1, sizeof(int64_t));
    col = (int64_t *)mxCalloc((size_t)nnz, sizeof(int64_t));
    val = (double *)mxCalloc(2 * (size_t)nnz, sizeof(double));
    rc = sssi_freq_cpml_2d_matrix(mxGetPr(prhs[0]), nz, nx, w, N, nB, dz, dx, &nnz, rowptr, col,
                                  val);
    sssi_mex_fail(rc);
    plhs[0] = mxCreateSparse((mwSize)n, (mwSize)n, (mwSize)nnz, mxCOMPLEX);
    jc = mxGetJc(plhs[0]);
    ir = mxGetIr(plhs[0]);
    av = mxGetComplexDoubles(plhs[0]);
    cnt = (int64_t *)mxCalloc((size_t)n + 1, sizeof(int64_t));
    for (e = 0; e < nnz; ++e) ++cnt[col[e] + 1];
    for (c = 0; c < n; ++c) cnt[c + 1] += cnt[c];
    for (c = 0; c <= n; ++c) jc[c] = (mwIndex)cnt[c];
    for (int64_t r = 0; r < n; ++r)  /* rows in order: ascending within each column */
        for (e = rowptr[r]; e < rowptr[r + 1]; ++e) {
            const int64_t q = cnt[col[e]]++;
            ir[q] = (mwIndex)r;
            av[q].real = val[2 * e];
            av[q].imag = val[2 * e + 1];
        }
    mxFree(rowptr);
    mxFree(col);
    mxFree(val);
    mxFree(cnt);
    if (nlhs < 2) return;

    /* snapshot = A \ (-source), shaped like source */
    if (nsrc == 0) {
        plhs[1] = mxCreateDoubleMatrix(0, 0, mxREAL);
        return;
    }
    {
        const double *s;
        double *tmp = NULL;
        size_t k, m = (size_t)n * (size_t)nsrc;
        if (mxIsComplex(prhs[1])) {
            s = (const double *)mxGetComplexDoubles(prhs[1]);
        } else {
            const double *re = mxGetPr(prhs[1]);
            tmp = (double *)mxCalloc(2 * m, sizeof(double));
            for (k = 0; k < m; ++k) tmp[2 * k] = re[k];
            s = tmp;
        }
        plhs[1] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[1]),
                                       mxGetDimensions(prhs[1]), mxDOUBLE_CLASS, mxCOMPLEX);
        rc = sssi_freqCpmlFor2dAw(mxGetPr(prhs[0]), nz, nx, s, nsrc, w, N, nB, dz, dx,
                                  (double *)mxGetComplexDoubles(plhs[1]));
        if (tmp) mxFree(tmp);
        sssi_mex_fail(rc);
    }
}
