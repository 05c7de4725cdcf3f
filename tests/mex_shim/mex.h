/*
 * mex.h -- a minimal stand-in for MATLAB's MEX/MX API, for compiling and
 * exercising aspire_amd/matlab/clm_mex.c without MATLAB.  TEST
 * INFRASTRUCTURE ONLY.  It declares the subset of functions the gateway
 * calls, with the signatures MathWorks documents for the R2018a
 * interleaved-complex API (mwSize/mwIndex as size_t, as in -largeArrayDims
 * builds); tests/mex_shim/mex_shim.c implements them on plain C memory.
 */
#ifndef CLM_TEST_MEX_H
#define CLM_TEST_MEX_H

#include <stdbool.h>
#include <stddef.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;

typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxDOUBLE_CLASS = 6,
    mxSINGLE_CLASS = 7,
    mxINT32_CLASS = 12
} mxClassID;

typedef double mxDouble;
typedef struct { double real, imag; } mxComplexDouble;
typedef struct { float real, imag; } mxComplexSingle;

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID classid, mxComplexity flag);
mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity flag);
void mxDestroyArray(mxArray *pa);

mxDouble *mxGetDoubles(const mxArray *pa);
mxComplexDouble *mxGetComplexDoubles(const mxArray *pa);
mxComplexSingle *mxGetComplexSingles(const mxArray *pa);
mwIndex *mxGetJc(const mxArray *pa);
mwIndex *mxGetIr(const mxArray *pa);

mwSize mxGetNumberOfDimensions(const mxArray *pa);
const mwSize *mxGetDimensions(const mxArray *pa);
size_t mxGetM(const mxArray *pa);
size_t mxGetN(const mxArray *pa);
size_t mxGetNumberOfElements(const mxArray *pa);
bool mxIsComplex(const mxArray *pa);
bool mxIsDouble(const mxArray *pa);
bool mxIsSingle(const mxArray *pa);
bool mxIsNumeric(const mxArray *pa);
bool mxIsEmpty(const mxArray *pa);
double mxGetScalar(const mxArray *pa);

void *mxMalloc(size_t n);
void mxFree(void *ptr);

void mexErrMsgIdAndTxt(const char *errorid, const char *errormsg, ...) __attribute__((noreturn));
void mexWarnMsgIdAndTxt(const char *warningid, const char *warningmsg, ...);
int mexPrintf(const char *message, ...);
int mexAtExit(void (*exit_fcn)(void));

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#endif
