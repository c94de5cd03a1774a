/*
 * sssi_mex_common.h -- shared argument handling of the MEX gateways that put
 * libsssi (include/sssi.h) behind the reference's MATLAB function names.
 *
 * The gateways keep the reference signatures (name, argument order, output
 * order and shapes) and add what the reference leaves unchecked: class
 * (real double), dimensions, nt >= 1, ranges of nDiffOrder and nBoundary.
 * Errors raise MATLAB errors with identifiers "sssi:<what>".
 *
 * Signature-preserving extensions: `source` may be a struct with fields iz,
 * ix (, iy), wavelet (0-based indices, one wavelet column per point, nt rows)
 * instead of the dense (nz,nx[,ny],nt) array; an optional trailing struct
 * argument `opts` may set snapStride (snapshot slice s holds step
 * s*snapStride).  With nlhs < 2 no snapshot is computed or allocated.
 */
#ifndef SSSI_MEX_COMMON_H
#define SSSI_MEX_COMMON_H

#include "mex.h"
#include "sssi.h"

#include <stdint.h>
#include <stdlib.h>

static void sssi_mex_check_double(const mxArray *a, const char *name)
{
    if (!mxIsDouble(a) || mxIsComplex(a))
        mexErrMsgIdAndTxt("sssi:class", "%s must be a real double array", name);
}

static double sssi_mex_scalar(const mxArray *a, const char *name)
{
    sssi_mex_check_double(a, name);
    if (mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("sssi:size", "%s must be a scalar", name);
    return mxGetScalar(a);
}

/* Library status -> MATLAB error: a non-finite wavefield (the watchdog's
 * SSSI_EUNSTABLE, a time step beyond the stability limit) gets its own
 * identifier; every other failure is sssi:run. */
static void sssi_mex_fail(int rc)
{
    if (rc == SSSI_EUNSTABLE) mexErrMsgIdAndTxt("sssi:unstable", "%s", sssi_last_error());
    if (rc != SSSI_OK) mexErrMsgIdAndTxt("sssi:run", "%s", sssi_last_error());
}

static int64_t sssi_mex_opt_int(const mxArray *opts, const char *field, int64_t dflt)
{
    if (!opts || !mxIsStruct(opts)) return dflt;
    const mxArray *f = mxGetField(opts, 0, field);
    if (!f) return dflt;
    return (int64_t)sssi_mex_scalar(f, field);
}

static int sssi_mex_int_arg(const mxArray *a, const char *name, int lo, int hi)
{
    double d = sssi_mex_scalar(a, name);
    if (d != (double)(int)d || (int)d < lo || (int)d > hi)
        mexErrMsgIdAndTxt("sssi:range", "%s must be an integer in [%d, %d]", name, lo, hi);
    return (int)d;
}

/* Point list from a struct source {iz, ix[, iy], wavelet(nt x nsrc)}; the
 * wavelet is re-ordered to the [t*nsrc + s] layout of the C ABI. */
typedef struct {
    int n;
    int *iz, *ix, *iy;
    double *wavelet;
    int64_t nt;
} sssi_mex_points;

static void sssi_mex_read_points(const mxArray *s, int want_iy, sssi_mex_points *p)
{
    const char *names[3] = {"iz", "ix", "iy"};
    int *dst[3];
    const mxArray *w = mxGetField(s, 0, "wavelet");
    if (!w) mexErrMsgIdAndTxt("sssi:source", "source struct needs a wavelet field");
    sssi_mex_check_double(w, "source.wavelet");
    p->n = (int)mxGetN(w);
    p->nt = (int64_t)mxGetM(w);
    if (p->n < 1 || p->nt < 1) mexErrMsgIdAndTxt("sssi:source", "source.wavelet is empty");
    {
        const double *col = mxGetPr(w);
        p->wavelet = (double *)mxCalloc((size_t)(p->n * p->nt), sizeof(double));
        for (int q = 0; q < p->n; ++q)
            for (int64_t t = 0; t < p->nt; ++t) p->wavelet[t * p->n + q] = col[q * p->nt + t];
    }
    for (int k = 0; k < 2 + want_iy; ++k) {
        const mxArray *f = mxGetField(s, 0, names[k]);
        if (!f) mexErrMsgIdAndTxt("sssi:source", "source struct needs a %s field", names[k]);
        sssi_mex_check_double(f, names[k]);
        if ((int)mxGetNumberOfElements(f) != p->n)
            mexErrMsgIdAndTxt("sssi:source", "source.%s must have one entry per wavelet column",
                              names[k]);
        dst[k] = (int *)mxCalloc(p->n, sizeof(int));
        const double *v = mxGetPr(f);
        for (int q = 0; q < p->n; ++q) {
            /* whole, non-negative and finite (NaN fails the comparison), so the
             * cast is exact; the grid bounds are checked by the library */
            if (!(v[q] >= 0.0 && v[q] <= 2147483647.0) || v[q] != (double)(int)v[q])
                mexErrMsgIdAndTxt("sssi:source", "source.%s entries must be non-negative integers",
                                  names[k]);
            dst[k][q] = (int)v[q];
        }
    }
    p->iz = dst[0];
    p->ix = dst[1];
    p->iy = want_iy ? dst[2] : NULL;
}

/* 3-D snapshot dims (nz, nx[, ny], slices) for nt steps at a stride. */
static mxArray *sssi_mex_snapshot(int nlhs, int ndim, const mwSize *grid, int64_t nt,
                                  int64_t stride)
{
    mwSize dims[4];
    if (nlhs < 2) return NULL;
    for (int k = 0; k < ndim; ++k) dims[k] = grid[k];
    dims[ndim] = (mwSize)((nt + stride - 1) / stride);
    return mxCreateNumericArray(ndim + 1, dims, mxDOUBLE_CLASS, mxREAL);
}

#endif
