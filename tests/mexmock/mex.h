/*
 * mex.h (test double) -- the subset of MATLAB's MEX/matrix API that the
 * gateways in sssi_amd/mex use, implemented in mexmock.c so that pytest can
 * compile the gateways with gcc and drive them through ctypes.  Arrays are
 * double (column-major; complex ones interleaved, as the -R2018a API keeps
 * them), sparse double (compressed columns) or 1x1 structs; errors longjmp
 * back to mock_invoke, which reports the identifier and message.
 */
#ifndef SSSI_TEST_MEX_H
#define SSSI_TEST_MEX_H

#include <stddef.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct { double real, imag; } mxComplexDouble;
typedef struct mxArray_tag mxArray;
typedef enum { mxUNKNOWN_CLASS = 0, mxSTRUCT_CLASS = 2, mxDOUBLE_CLASS = 6 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

#ifdef __cplusplus
extern "C" {
#endif

int mxIsDouble(const mxArray *a);
int mxIsComplex(const mxArray *a);
int mxIsStruct(const mxArray *a);
int mxIsSparse(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double *mxGetPr(const mxArray *a);
mxComplexDouble *mxGetComplexDoubles(const mxArray *a);
mwIndex *mxGetJc(const mxArray *a);
mwIndex *mxGetIr(const mxArray *a);
mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity c);
double mxGetScalar(const mxArray *a);
mxArray *mxGetField(const mxArray *a, size_t index, const char *name);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c);
void *mxCalloc(size_t n, size_t size);
void mxFree(void *p);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#ifdef __cplusplus
}
#endif

#endif
