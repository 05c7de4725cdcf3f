/*
 * clm_mex.c -- MEX gateway of libclm for MATLAB/Octave.
 *
 *   [clstack,corrstack,shift_equations,shift_equations_map,clstack_mask] =
 *       clm_mex(pf, verbose, max_shift, shift_step, precision,
 *               map_filter_radius, ref_clmatrix)
 *
 * The five outputs of cryo_clmatrix (sinograms/cryo_clmatrix.m:1-3, in the
 * form of cryo_clmatrix_cpu.m:764-772) from libclm's C ABI (include/clm.h):
 * clm_clmatrix (or clm_clmatrix_filtered for map_filter_radius > 0) on the
 * GPU, then clm_assemble and clm_shift_equations_csc on the host in FP64.
 * precision: 0 FP32, 1 BF16 + re-score, 2 FP16 + re-score (default), 3 FP64.
 * clstack_mask is filled by clm_clstack_mask when verbose has bit 2 and a
 * K x K ref_clmatrix is given (cryo_clmatrix_cpu.m:567-617,697-698), else 0.
 * verbose bit 1 prints libclm's reports with mexPrintf; lines of near-zero
 * norm raise the warning clm:raynormalize (cryo_raynormalize.m:20-22).
 *
 * Build (R2018a interleaved-complex API):
 *   mex -O -R2018a clm_mex.c -I<repo>/include -L<repo>/aspire_amd -lclm \
 *       LDFLAGS='$LDFLAGS -Wl,-rpath,<repo>/aspire_amd'
 *
 * One context (HIP stream and cached device buffers) is created on first use
 * and destroyed at `clear mex`.  MATLAB calls mexFunction on one thread and
 * no mx* call is made off it; errors are raised only after libclm has
 * returned (mexErrMsgIdAndTxt does not return), with every mxMalloc'd
 * buffer freed first.
 */
#include <stdint.h>
#include <string.h>

#include "mex.h"
#include "clm.h"

static clm_ctx *g_ctx = NULL;

static void cleanup(void) {
    if (g_ctx) {
        clm_destroy(g_ctx);
        g_ctx = NULL;
    }
}

static void report(void *user, int level, const char *msg) {
    (void)user;
    if (level) mexWarnMsgIdAndTxt("clm:raynormalize", "%s", msg);
    else mexPrintf("%s\n", msg);
}

static double scalar_arg(int nrhs, const mxArray *prhs[], int i, double dflt) {
    if (i >= nrhs || mxIsEmpty(prhs[i])) return dflt;
    if (!mxIsNumeric(prhs[i]) || mxGetNumberOfElements(prhs[i]) != 1)
        mexErrMsgIdAndTxt("clm:arg", "argument %d must be a numeric scalar", i + 1);
    return mxGetScalar(prhs[i]);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 1) mexErrMsgIdAndTxt("clm:arg", "pf is required");
    if (nlhs > 5) mexErrMsgIdAndTxt("clm:arg", "at most five outputs");
    const mxArray *pf = prhs[0];
    const mwSize nd = mxGetNumberOfDimensions(pf);
    const mwSize *d = mxGetDimensions(pf);
    if (nd > 3) mexErrMsgIdAndTxt("clm:arg", "pf must be n_r x n_theta x K");
    clm_params p;
    clm_default_params(&p);
    p.n_r = (int32_t)d[0];
    p.n_theta = (int32_t)d[1];
    p.K = nd > 2 ? (int32_t)d[2] : 1;
    if (p.n_theta % 2) mexErrMsgIdAndTxt("clm:arg", "n_theta must be even");  /* cryo_clmatrix_cpu.m:84-86 */
    p.verbose = (int32_t)scalar_arg(nrhs, prhs, 1, 1.0);
    p.max_shift = scalar_arg(nrhs, prhs, 2, 15.0);
    p.shift_step = scalar_arg(nrhs, prhs, 3, 1.0);
    p.precision = (int32_t)scalar_arg(nrhs, prhs, 4, (double)CLM_FP16_RESCORE);
    const int32_t radius = (int32_t)scalar_arg(nrhs, prhs, 5, 0.0);
    const mxArray *ref = nrhs > 6 && !mxIsEmpty(prhs[6]) ? prhs[6] : NULL;
    if (ref && (!mxIsDouble(ref) || mxIsComplex(ref) || mxGetM(ref) != (size_t)p.K || mxGetN(ref) != (size_t)p.K))
        mexErrMsgIdAndTxt("clm:arg", "ref_clmatrix must be a real K x K double matrix");

    /* interleaved complex: exactly the layout clm.h expects; a real pf is
       not a polar Fourier transform */
    const void *data = NULL;
    if (!mxIsComplex(pf)) mexErrMsgIdAndTxt("clm:arg", "pf must be complex");
    if (mxIsDouble(pf)) {
        p.pf_dtype = CLM_C128;
        data = mxGetComplexDoubles(pf);
    } else if (mxIsSingle(pf)) {
        p.pf_dtype = CLM_C64;
        data = mxGetComplexSingles(pf);
    } else {
        mexErrMsgIdAndTxt("clm:arg", "pf must be single or double");
    }

    if (!g_ctx) {
        int err = 0;
        g_ctx = clm_create(0, &err);
        if (!g_ctx) mexErrMsgIdAndTxt("clm:nodev", "clm_create failed (%d): no usable GPU", err);
        clm_set_log(g_ctx, report, NULL);
        mexAtExit(cleanup);
    }
    const int64_t P = clm_num_pairs(p.K);
    const size_t np = (size_t)(P > 0 ? P : 1);
    int32_t *cl1 = (int32_t *)mxMalloc(np * sizeof(int32_t));
    int32_t *cl2 = (int32_t *)mxMalloc(np * sizeof(int32_t));
    double *corr = (double *)mxMalloc(np * sizeof(double));
    double *shift = (double *)mxMalloc(np * sizeof(double));
    const int rc = radius > 0 ? clm_clmatrix_filtered(g_ctx, &p, radius, data, cl1, cl2, NULL, corr, shift, NULL)
                              : clm_clmatrix(g_ctx, &p, data, cl1, cl2, NULL, corr, shift, NULL);
    if (rc) {
        mxFree(cl1);
        mxFree(cl2);
        mxFree(corr);
        mxFree(shift);
        mexErrMsgIdAndTxt("clm:run", "libclm error %d: %s", rc, clm_last_error(g_ctx));
    }

    const mwSize K = (mwSize)p.K;
    mxArray *clstack = mxCreateDoubleMatrix(K, K, mxREAL);
    mxArray *corrstack = mxCreateDoubleMatrix(K, K, mxREAL);
    mxArray *emap = mxCreateDoubleMatrix(K, K, mxREAL);
    clm_assemble(p.K, cl1, cl2, corr, mxGetDoubles(clstack), mxGetDoubles(corrstack), mxGetDoubles(emap));

    int64_t nnz = 0;
    clm_shift_equations_csc(p.K, p.n_theta, cl1, cl2, shift, NULL, NULL, NULL, &nnz);
    mxArray *eqs = mxCreateSparse((mwSize)P, 2 * K + 1, (mwSize)(nnz > 0 ? nnz : 1), mxREAL);
    int64_t *jc = (int64_t *)mxMalloc((2 * (size_t)K + 2) * sizeof(int64_t));
    int64_t *ir = (int64_t *)mxMalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
    clm_shift_equations_csc(p.K, p.n_theta, cl1, cl2, shift, jc, ir, mxGetDoubles(eqs), &nnz);
    mwIndex *Jc = mxGetJc(eqs), *Ir = mxGetIr(eqs);
    for (mwSize c = 0; c < 2 * K + 2; c++) Jc[c] = (mwIndex)jc[c];
    for (int64_t i = 0; i < nnz; i++) Ir[i] = (mwIndex)ir[i];
    mxFree(jc);
    mxFree(ir);
    mxFree(cl1);
    mxFree(cl2);
    mxFree(corr);
    mxFree(shift);

    mxArray *mask = mxCreateDoubleMatrix(K, K, mxREAL);
    if ((p.verbose & 2) && ref) {
        int64_t matched = 0;
        clm_clstack_mask(p.K, p.n_theta, mxGetDoubles(clstack), mxGetDoubles(ref), mxGetDoubles(mask), &matched);
        mexPrintf("Matched common-lines=%lld/%lld  (%3.1f%%)\n", (long long)matched, (long long)P,
                  P > 0 ? 100.0 * (double)matched / (double)P : 0.0);
    }

    mxArray *outs[5] = {clstack, corrstack, eqs, emap, mask};
    for (int i = 0; i < 5; i++) {
        if (i < (nlhs > 0 ? nlhs : 1)) plhs[i] = outs[i];
        else mxDestroyArray(outs[i]);
    }
}
