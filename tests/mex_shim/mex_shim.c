/*
 * mex_shim.c -- implementation of the stand-in MEX/MX API declared in
 * tests/mex_shim/mex.h, plus shim_* entry points through which a test
 * (Python ctypes) builds mxArrays, calls mexFunction and reads the results.
 * TEST INFRASTRUCTURE ONLY.  mexErrMsgIdAndTxt long-jumps back to
 * shim_call, as MATLAB's does to the interpreter; mxMalloc'd blocks still
 * live at that point are counted, so a test sees a gateway that leaks.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    int complex_;
    int sparse;
    mwSize ndim;
    mwSize dims[8];
    void *data;    /* interleaved complex or real */
    mwIndex *jc, *ir;
    mwSize nzmax;
};

static jmp_buf g_jmp;
static int g_in_call = 0;
static char g_err_id[128], g_err_msg[1024];
static char g_out[1 << 16];
static size_t g_out_len = 0;
static int g_warnings = 0;
static void (*g_exit)(void) = NULL;
static long g_live_blocks = 0;

static size_t esize(const mxArray *a) {
    size_t s = a->cls == mxSINGLE_CLASS ? 4 : a->cls == mxINT32_CLASS ? 4 : 8;
    return a->complex_ ? 2 * s : s;
}

size_t mxGetNumberOfElements(const mxArray *a) {
    size_t n = 1;
    for (mwSize i = 0; i < a->ndim; i++) n *= a->dims[i];
    return n;
}

static mxArray *make(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity flag) {
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->cls = cls;
    a->complex_ = flag == mxCOMPLEX;
    a->ndim = ndim < 2 ? 2 : ndim;
    a->dims[0] = ndim > 0 ? dims[0] : 1;
    a->dims[1] = ndim > 1 ? dims[1] : 1;
    for (mwSize i = 2; i < ndim && i < 8; i++) a->dims[i] = dims[i];
    while (a->ndim > 2 && a->dims[a->ndim - 1] == 1) a->ndim--;  /* MATLAB drops trailing singletons */
    size_t n = mxGetNumberOfElements(a);
    a->data = calloc(n ? n : 1, esize(a));
    return a;
}

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag) {
    mwSize d[2] = {m, n};
    return make(2, d, mxDOUBLE_CLASS, flag);
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity flag) {
    return make(ndim, dims, cls, flag);
}

mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity flag) {
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->cls = mxDOUBLE_CLASS;
    a->complex_ = flag == mxCOMPLEX;
    a->sparse = 1;
    a->ndim = 2;
    a->dims[0] = m;
    a->dims[1] = n;
    a->nzmax = nzmax ? nzmax : 1;
    a->data = calloc(a->nzmax, esize(a));
    a->ir = (mwIndex *)calloc(a->nzmax, sizeof(mwIndex));
    a->jc = (mwIndex *)calloc(n + 1, sizeof(mwIndex));
    return a;
}

void mxDestroyArray(mxArray *a) {
    if (!a) return;
    free(a->data);
    free(a->jc);
    free(a->ir);
    free(a);
}

mxDouble *mxGetDoubles(const mxArray *a) {
    return a->cls == mxDOUBLE_CLASS && !a->complex_ ? (mxDouble *)a->data : NULL;
}
mxComplexDouble *mxGetComplexDoubles(const mxArray *a) {
    return a->cls == mxDOUBLE_CLASS && a->complex_ ? (mxComplexDouble *)a->data : NULL;
}
mxComplexSingle *mxGetComplexSingles(const mxArray *a) {
    return a->cls == mxSINGLE_CLASS && a->complex_ ? (mxComplexSingle *)a->data : NULL;
}
mwIndex *mxGetJc(const mxArray *a) { return a->jc; }
mwIndex *mxGetIr(const mxArray *a) { return a->ir; }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->ndim; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
size_t mxGetM(const mxArray *a) { return a->dims[0]; }
size_t mxGetN(const mxArray *a) {
    size_t n = 1;
    for (mwSize i = 1; i < a->ndim; i++) n *= a->dims[i];
    return n;
}
bool mxIsComplex(const mxArray *a) { return a->complex_ != 0; }
bool mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsSingle(const mxArray *a) { return a->cls == mxSINGLE_CLASS; }
bool mxIsNumeric(const mxArray *a) { return a->cls == mxDOUBLE_CLASS || a->cls == mxSINGLE_CLASS || a->cls == mxINT32_CLASS; }
bool mxIsEmpty(const mxArray *a) { return mxGetNumberOfElements(a) == 0; }
double mxGetScalar(const mxArray *a) {
    if (mxIsEmpty(a)) return 0.0;
    if (a->cls == mxSINGLE_CLASS) return ((float *)a->data)[0];
    if (a->cls == mxINT32_CLASS) return ((int32_t *)a->data)[0];
    return ((double *)a->data)[0];
}

void *mxMalloc(size_t n) {
    g_live_blocks++;
    return malloc(n ? n : 1);
}
void mxFree(void *p) {
    if (p) g_live_blocks--;
    free(p);
}

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    snprintf(g_err_id, sizeof g_err_id, "%s", id);
    vsnprintf(g_err_msg, sizeof g_err_msg, fmt, ap);
    va_end(ap);
    if (g_in_call) longjmp(g_jmp, 1);
    abort();
}

static void append(const char *fmt, va_list ap) {
    if (g_out_len >= sizeof g_out - 1) return;
    int k = vsnprintf(g_out + g_out_len, sizeof g_out - g_out_len, fmt, ap);
    if (k > 0) g_out_len += (size_t)k;
    if (g_out_len > sizeof g_out - 1) g_out_len = sizeof g_out - 1;
}

void mexWarnMsgIdAndTxt(const char *id, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    g_warnings++;
    char head[160];
    snprintf(head, sizeof head, "Warning [%s]: ", id);
    size_t hl = strlen(head);
    if (g_out_len + hl < sizeof g_out - 1) {
        memcpy(g_out + g_out_len, head, hl);
        g_out_len += hl;
    }
    append(fmt, ap);
    va_end(ap);
}

int mexPrintf(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    size_t before = g_out_len;
    append(fmt, ap);
    va_end(ap);
    return (int)(g_out_len - before);
}

int mexAtExit(void (*fn)(void)) {
    g_exit = fn;
    return 0;
}

/* ---- test entry points ---- */

mxArray *shim_create(int cls, int complex_, int ndim, const int64_t *dims, const void *data) {
    mwSize d[8];
    for (int i = 0; i < ndim && i < 8; i++) d[i] = (mwSize)dims[i];
    mxArray *a = make((mwSize)ndim, d, (mxClassID)cls, complex_ ? mxCOMPLEX : mxREAL);
    if (data) memcpy(a->data, data, mxGetNumberOfElements(a) * esize(a));
    return a;
}

/* 0 on success, 1 when the gateway raised an error (id/message kept) */
int shim_call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
    g_err_id[0] = g_err_msg[0] = 0;
    g_in_call = 1;
    if (setjmp(g_jmp)) {
        g_in_call = 0;
        return 1;
    }
    mexFunction(nlhs, plhs, nrhs, prhs);
    g_in_call = 0;
    return 0;
}

const char *shim_error_id(void) { return g_err_id; }
const char *shim_error_msg(void) { return g_err_msg; }
const char *shim_output(void) {
    g_out[g_out_len] = 0;
    return g_out;
}
int shim_warnings(void) { return g_warnings; }
long shim_live_blocks(void) { return g_live_blocks; }
void shim_clear_mex(void) {
    if (g_exit) g_exit();
    g_exit = NULL;
}

int shim_ndim(const mxArray *a) { return (int)a->ndim; }
int64_t shim_dim(const mxArray *a, int i) { return (int64_t)a->dims[i]; }
int shim_is_sparse(const mxArray *a) { return a->sparse; }
void *shim_data(const mxArray *a) { return a->data; }
mwIndex *shim_jc(const mxArray *a) { return a->jc; }
mwIndex *shim_ir(const mxArray *a) { return a->ir; }
void shim_destroy(mxArray *a) { mxDestroyArray(a); }
