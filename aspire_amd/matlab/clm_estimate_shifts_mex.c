/*
 * clm_estimate_shifts_mex.c -- MEX gateway of libclm for
 *
 *   [est_shifts, shift_equations] = clm_estimate_shifts_mex(pf, rotations,
 *       max_shift, shift_step, memoryfactor, seed)
 *
 * the outputs of cryo_estimate_shifts (sinograms/cryo_estimate_shifts.m:1-2,
 * 225-230) from libclm's C ABI (include/clm.h): clm_n_equations and
 * clm_equation_pairs pick the pairs (:58-94,130-137; a seeded uniform draw
 * instead of MATLAB's randperm stream, seed default 0),
 * clm_cryo_estimate_shifts finds their common lines and 1-D shifts from the
 * rotations and solves the equations on the GPU (:142-228: lsqr with tol
 * 1e-8 and maxit 100), and clm_shift_equations_csc_ex with
 * CLM_FLIP_BY_INDEX assembles shift_equations on the host in FP64 (the
 * selected pairs' rows in pair order, the right-hand side as the last
 * column).  pf is n_r x n_theta x K complex (double or single), rotations
 * 3 x 3 x K double; max_shift and shift_step default to the reference's
 * (shift_step 1, memoryfactor 10000).  The reference's shifts_2d_ref and
 * verbose arguments only drive its diagnostics and are not taken.
 *
 * Build (R2018a interleaved-complex API):
 *   mex -O -R2018a clm_estimate_shifts_mex.c -I<repo>/include \
 *       -L<repo>/aspire_amd -lclm LDFLAGS='$LDFLAGS -Wl,-rpath,<repo>/aspire_amd'
 *
 * One context is created on first use and destroyed at `clear mex`; errors
 * are raised only after libclm has returned, with every mxMalloc'd buffer
 * freed first.
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

static double scalar_arg(int nrhs, const mxArray *prhs[], int i, double dflt) {
    if (i >= nrhs || mxIsEmpty(prhs[i])) return dflt;
    if (!mxIsNumeric(prhs[i]) || mxGetNumberOfElements(prhs[i]) != 1)
        mexErrMsgIdAndTxt("clm:arg", "argument %d must be a numeric scalar", i + 1);
    return mxGetScalar(prhs[i]);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 3) mexErrMsgIdAndTxt("clm:arg", "pf, rotations and max_shift are required");
    if (nlhs > 2) mexErrMsgIdAndTxt("clm:arg", "at most two outputs");
    const mxArray *pf = prhs[0], *rot = prhs[1];
    const mwSize nd = mxGetNumberOfDimensions(pf);
    const mwSize *d = mxGetDimensions(pf);
    if (nd > 3) mexErrMsgIdAndTxt("clm:arg", "pf must be n_r x n_theta x K");
    clm_params p;
    clm_default_params(&p);
    p.n_r = (int32_t)d[0];
    p.n_theta = (int32_t)d[1];
    p.K = nd > 2 ? (int32_t)d[2] : 1;
    if (p.n_theta % 2) mexErrMsgIdAndTxt("clm:arg", "n_theta must be even");
    if (p.K < 2) mexErrMsgIdAndTxt("clm:arg", "at least two images are needed");
    const mwSize rn = mxGetNumberOfDimensions(rot);
    const mwSize *rd = mxGetDimensions(rot);
    if (!mxIsDouble(rot) || mxIsComplex(rot) || rd[0] != 3 || rd[1] != 3 || (rn > 2 ? rd[2] : 1) != (mwSize)p.K)
        mexErrMsgIdAndTxt("clm:arg", "rotations must be a real 3 x 3 x K double array");
    p.max_shift = scalar_arg(nrhs, prhs, 2, 0.0);
    p.shift_step = scalar_arg(nrhs, prhs, 3, 1.0);
    const double memoryfactor = scalar_arg(nrhs, prhs, 4, 10000.0);
    const double seed = scalar_arg(nrhs, prhs, 5, 0.0);
    if (!(seed >= 0.0 && seed < 18446744073709551616.0)) mexErrMsgIdAndTxt("clm:arg", "seed must be a non-negative integer");
    int64_t neq = 0;
    if (clm_n_equations(p.K, memoryfactor, &neq))  /* cryo_estimate_shifts.m:58-60 */
        mexErrMsgIdAndTxt("clm:arg", "subsamplingfactor must be between 0 and 1 or larger than 100");

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
        mexAtExit(cleanup);
    }
    const int64_t P = clm_num_pairs(p.K);
    uint8_t *sel = NULL;
    int64_t nrow = P;
    if (neq < P) {
        sel = (uint8_t *)mxMalloc((size_t)P);
        clm_equation_pairs(p.K, neq, (uint64_t)seed, sel, &nrow);
    }
    int32_t *cl1 = (int32_t *)mxMalloc((size_t)P * sizeof(int32_t));
    int32_t *cl2 = (int32_t *)mxMalloc((size_t)P * sizeof(int32_t));
    double *shift = (double *)mxMalloc((size_t)P * sizeof(double));
    double *x = (double *)mxMalloc((size_t)p.K * 2 * sizeof(double));
    const int rc = clm_cryo_estimate_shifts(g_ctx, &p, data, mxGetDoubles(rot), sel, 1.0e-8, 100, cl1, cl2, shift,
                                            x, NULL);
    if (rc) {
        if (sel) mxFree(sel);
        mxFree(cl1);
        mxFree(cl2);
        mxFree(shift);
        mxFree(x);
        mexErrMsgIdAndTxt("clm:run", "libclm error %d: %s", rc, clm_last_error(g_ctx));
    }

    /* est_shifts = full(transpose(reshape(x, 2, K))): row k = (x(2k-1), x(2k)) */
    const mwSize K = (mwSize)p.K;
    mxArray *est = mxCreateDoubleMatrix(K, 2, mxREAL);
    double *e = mxGetDoubles(est);
    for (mwSize k = 0; k < K; k++) {
        e[k] = x[2 * k];
        e[K + k] = x[2 * k + 1];
    }
    plhs[0] = est;
    if (nlhs > 1) {
        int64_t nnz = 0;
        clm_shift_equations_csc_ex(p.K, p.n_theta, sel, CLM_FLIP_BY_INDEX, cl1, cl2, shift, NULL, NULL, NULL, &nnz);
        mxArray *eqs = mxCreateSparse((mwSize)nrow, 2 * K + 1, (mwSize)(nnz > 0 ? nnz : 1), mxREAL);
        int64_t *jc = (int64_t *)mxMalloc((2 * K + 2) * sizeof(int64_t));
        int64_t *ir = (int64_t *)mxMalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
        clm_shift_equations_csc_ex(p.K, p.n_theta, sel, CLM_FLIP_BY_INDEX, cl1, cl2, shift, jc, ir,
                                   mxGetDoubles(eqs), &nnz);
        mwIndex *mjc = mxGetJc(eqs), *mir = mxGetIr(eqs);
        for (mwSize q = 0; q < 2 * K + 2; q++) mjc[q] = (mwIndex)jc[q];
        for (int64_t q = 0; q < nnz; q++) mir[q] = (mwIndex)ir[q];
        mxFree(jc);
        mxFree(ir);
        plhs[1] = eqs;
    }
    if (sel) mxFree(sel);
    mxFree(cl1);
    mxFree(cl2);
    mxFree(shift);
    mxFree(x);
}
