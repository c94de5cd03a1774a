/*
 * mexmock.c -- implementation of the test double in mex.h, plus the ctypes
 * entry points the gateway tests call (mock_double, mock_struct,
 * mock_invoke, mock_free).  One shared object per gateway: each links this
 * file with one *_mex.c and libsssi.
 */
#include "mex.h"

#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MOCK_MAX_FIELDS 8
#define MOCK_MAX_ALLOCS 256

struct mxArray_tag {
    mxClassID cls;
    int complex_flag;
    mwSize ndim;
    mwSize dims[8];
    double *pr;     /* complex: interleaved (re, im) */
    int sparse;     /* compressed columns: jc[n+1], ir[nzmax], nzmax values in pr */
    mwIndex *jc, *ir;
    mwSize nzmax;
    int nfields;
    char names[MOCK_MAX_FIELDS][32];
    mxArray *fields[MOCK_MAX_FIELDS];
};

static jmp_buf mock_jmp;
static char mock_errid[128], mock_errmsg[512];
static void *mock_allocs[MOCK_MAX_ALLOCS];
static int mock_nallocs;

static void *track(void *p)
{
    if (mock_nallocs < MOCK_MAX_ALLOCS) mock_allocs[mock_nallocs++] = p;
    return p;
}

int mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsComplex(const mxArray *a) { return a->complex_flag; }
int mxIsStruct(const mxArray *a) { return a->cls == mxSTRUCT_CLASS; }
int mxIsSparse(const mxArray *a) { return a->sparse; }
mxComplexDouble *mxGetComplexDoubles(const mxArray *a)
{
    return a->complex_flag ? (mxComplexDouble *)a->pr : NULL;
}
mwIndex *mxGetJc(const mxArray *a) { return a->jc; }
mwIndex *mxGetIr(const mxArray *a) { return a->ir; }
size_t mxGetM(const mxArray *a) { return a->dims[0]; }
size_t mxGetN(const mxArray *a)
{
    size_t n = 1;
    for (mwSize k = 1; k < a->ndim; ++k) n *= a->dims[k];
    return n;
}
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->ndim; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->dims[0] * mxGetN(a); }
double *mxGetPr(const mxArray *a) { return a->pr; }
double mxGetScalar(const mxArray *a) { return a->pr ? a->pr[0] : 0.0; }

mxArray *mxGetField(const mxArray *a, size_t index, const char *name)
{
    if (a->cls != mxSTRUCT_CLASS || index != 0) return NULL;
    for (int k = 0; k < a->nfields; ++k)
        if (!strcmp(a->names[k], name)) return a->fields[k];
    return NULL;
}

static mxArray *make_array(mwSize ndim, const mwSize *dims, int cplx)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    size_t n = 1;
    /* MATLAB drops trailing singleton dimensions beyond the second */
    while (ndim > 2 && dims[ndim - 1] == 1) --ndim;
    a->cls = mxDOUBLE_CLASS;
    a->ndim = ndim < 2 ? 2 : ndim;
    for (mwSize k = 0; k < a->ndim; ++k) {
        a->dims[k] = k < ndim ? dims[k] : 1;
        n *= a->dims[k];
    }
    a->complex_flag = cplx;
    a->pr = (double *)calloc(n ? (cplx ? 2 * n : n) : 1, sizeof(double));
    return a;
}

static mxArray *make_double(mwSize ndim, const mwSize *dims) { return make_array(ndim, dims, 0); }

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c)
{
    mwSize d[2] = {m, n};
    return make_array(2, d, c == mxCOMPLEX);
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c)
{
    (void)cls;
    return make_array(ndim, dims, c == mxCOMPLEX);
}

mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity c)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->cls = mxDOUBLE_CLASS;
    a->ndim = 2;
    a->dims[0] = m;
    a->dims[1] = n;
    a->sparse = 1;
    a->complex_flag = c == mxCOMPLEX;
    a->nzmax = nzmax ? nzmax : 1;
    a->jc = (mwIndex *)calloc(n + 1, sizeof(mwIndex));
    a->ir = (mwIndex *)calloc(a->nzmax, sizeof(mwIndex));
    a->pr = (double *)calloc(a->complex_flag ? 2 * a->nzmax : a->nzmax, sizeof(double));
    return a;
}

void *mxCalloc(size_t n, size_t size) { return track(calloc(n ? n : 1, size ? size : 1)); }

void mxFree(void *p)
{
    for (int k = 0; k < mock_nallocs; ++k)
        if (mock_allocs[k] == p) mock_allocs[k] = NULL;
    free(p);
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    va_list ap;
    snprintf(mock_errid, sizeof mock_errid, "%s", id);
    va_start(ap, fmt);
    vsnprintf(mock_errmsg, sizeof mock_errmsg, fmt, ap);
    va_end(ap);
    longjmp(mock_jmp, 1);
}

/* ---- ctypes entry points ---- */

mxArray *mock_double(int ndim, const long long *dims, const double *data)
{
    mwSize d[8];
    size_t n = 1;
    for (int k = 0; k < ndim && k < 8; ++k) {
        d[k] = (mwSize)dims[k];
        n *= d[k];
    }
    mxArray *a = make_double((mwSize)ndim, d);
    if (data) memcpy(a->pr, data, n * sizeof(double));
    return a;
}

/* complex double array from interleaved (re, im) data */
mxArray *mock_complex(int ndim, const long long *dims, const double *data)
{
    mwSize d[8];
    size_t n = 1;
    for (int k = 0; k < ndim && k < 8; ++k) {
        d[k] = (mwSize)dims[k];
        n *= d[k];
    }
    mxArray *a = make_array((mwSize)ndim, d, 1);
    if (data) memcpy(a->pr, data, 2 * n * sizeof(double));
    return a;
}

/* flags of an output: bit 0 complex, bit 1 sparse */
int mock_kind(const mxArray *a) { return (a->complex_flag ? 1 : 0) | (a->sparse ? 2 : 0); }

long long mock_nnz(const mxArray *a) { return a->sparse ? (long long)a->jc[a->dims[1]] : 0; }

/* compressed columns of a sparse output: jc[n+1], ir[nnz], values (interleaved if complex) */
void mock_sparse(const mxArray *a, long long *jc, long long *ir, double *val)
{
    const size_t n = a->dims[1], nnz = a->jc[n];
    for (size_t k = 0; k <= n; ++k) jc[k] = (long long)a->jc[k];
    for (size_t k = 0; k < nnz; ++k) ir[k] = (long long)a->ir[k];
    memcpy(val, a->pr, (a->complex_flag ? 2 : 1) * nnz * sizeof(double));
}

mxArray *mock_int32(void)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->cls = mxUNKNOWN_CLASS;
    a->ndim = 2;
    a->dims[0] = a->dims[1] = 1;
    return a;
}

mxArray *mock_struct(int nfields, const char **names, mxArray **values)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->cls = mxSTRUCT_CLASS;
    a->ndim = 2;
    a->dims[0] = a->dims[1] = 1;
    a->nfields = nfields < MOCK_MAX_FIELDS ? nfields : MOCK_MAX_FIELDS;
    for (int k = 0; k < a->nfields; ++k) {
        snprintf(a->names[k], sizeof a->names[k], "%s", names[k]);
        a->fields[k] = values[k];
    }
    return a;
}

void mock_free(mxArray *a)
{
    if (!a) return;
    if (a->cls == mxSTRUCT_CLASS)
        for (int k = 0; k < a->nfields; ++k) mock_free(a->fields[k]);
    free(a->pr);
    free(a->jc);
    free(a->ir);
    free(a);
}

/* Returns 0 on success, 1 on a gateway error (id/message copied out). */
int mock_invoke(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *errid,
                char *errmsg)
{
    int rc = 0;
    mock_nallocs = 0;
    for (int k = 0; k < nlhs; ++k) plhs[k] = NULL;
    if (setjmp(mock_jmp) == 0) {
        mexFunction(nlhs, plhs, nrhs, prhs);
    } else {
        snprintf(errid, 128, "%s", mock_errid);
        snprintf(errmsg, 512, "%s", mock_errmsg);
        rc = 1;
    }
    for (int k = 0; k < mock_nallocs; ++k) free(mock_allocs[k]);
    mock_nallocs = 0;
    return rc;
}

double *mock_data(const mxArray *a, int *ndim, long long *dims)
{
    *ndim = (int)a->ndim;
    for (mwSize k = 0; k < a->ndim; ++k) dims[k] = (long long)a->dims[k];
    return a->pr;
}
