/*
 * clm.h -- C ABI of the MI355X common-lines engine (libclm.so).
 *
 * The library is a drop-in for ASPIRE's
 *   [clstack,corrstack,shift_equations,shift_equations_map,clstack_mask] =
 *       cryo_clmatrix(pf, NK, verbose, max_shift, shift_step)
 * (sinograms/cryo_clmatrix.m:1-3, which dispatches to
 * sinograms/cryo_clmatrix_cpu.m:1-3 or sinograms/cryo_clmatrix_gpu.m:1-3).
 * ASPIRE has no FFI for this path; its "operator API" is that MATLAB
 * function.  The entry points below are what a MEX gateway (or ctypes)
 * binds to replace it.
 *
 * Plain C types only: pointers, sizes, enums and POD structs.  Every call
 * returns 0 on success or a negative CLM_E* code; the message of the last
 * failure on a context is returned by clm_last_error().  Nothing throws
 * across this boundary.
 *
 * Data layout (MATLAB's): pf is complex, interleaved (re,im), column-major
 * n_r x n_theta x K; element (r,t,k) starts at scalar offset
 * 2*((k*n_theta + t)*n_r + r).  Per-pair outputs are indexed by equation
 * number e-1, pairs ordered (k1 ascending, k2 ascending), k1 < k2 -- the row
 * order of shift_equations (cryo_clmatrix_cpu.m:722-741).
 */
#ifndef CLM_H
#define CLM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLM_VERSION 10000 /* 1.0.0 */

/* error codes */
#define CLM_OK 0
#define CLM_E_ARG (-1)      /* invalid argument (e.g. odd n_theta)       */
#define CLM_E_HIP (-2)      /* HIP runtime failure                        */
#define CLM_E_NOMEM (-3)    /* device or host allocation failed           */
#define CLM_E_UNSUP (-4)    /* shape outside what the kernels support     */
#define CLM_E_NODEV (-5)    /* no usable GPU                              */

/* arithmetic of the pair contraction */
enum clm_precision {
    CLM_FP32 = 0,         /* FP32 MFMA contraction, FP32 argmax (plain GPU path) */
    CLM_BF16_RESCORE = 1, /* BF16 MFMA + FP32 re-score + FP64 near-tie re-score  */
    CLM_FP16_RESCORE = 2, /* FP16 MFMA + FP32 re-score + FP64 near-tie re-score  */
    CLM_FP64_REF = 3      /* FP64 on every candidate, the reference's arithmetic */
};

/* element type of pf */
enum clm_dtype {
    CLM_C128 = 0, /* complex double (MATLAB double)  */
    CLM_C64 = 1   /* complex float  (MATLAB single)  */
};

typedef struct clm_params {
    int32_t n_r;         /* rows of pf                                        */
    int32_t n_theta;     /* columns of pf; must be even (cryo_clmatrix_cpu.m:84) */
    int32_t K;           /* number of images                                  */
    int32_t pf_dtype;    /* enum clm_dtype                                    */
    double max_shift;    /* cryo_clmatrix arg 4 (default 15)                  */
    double shift_step;   /* cryo_clmatrix arg 5 (default 1)                   */
    int32_t precision;   /* enum clm_precision                                */
    int32_t verbose;     /* cryo_clmatrix arg 3 bitmask; bit 1 = progress log  */
    double tie_margin;   /* FP64 re-score of candidates within this of the
                            FP32 best (correlation units), default 1e-3       */
    int32_t variant;     /* 0 = tuned/default kernel variant, else a variant id */
    int32_t reserved;    /* tuning bits: bit 0 disables the lagged wave schedule;
                            bit 1 disables the FP32 refinement of uncertain
                            BF16/FP16 pairs (they take a second BF16/FP16
                            pass instead); bit 2 disables the raised wave
                            priority during MFMA issue                      */
} clm_params;

typedef struct clm_stats {
    int64_t pairs;             /* pairs computed                                  */
    int64_t cells_rescored;    /* FP32-rescored cells (16 rows x 2 maps each)     */
    int64_t uncertain_pairs;   /* pairs sent to the exact-threshold second pass   */
    int64_t emitted;           /* candidates emitted by the second pass           */
    int64_t fp64_candidates;   /* candidates re-scored in FP64                    */
    int64_t fallback_pairs;    /* pairs finished by the FP64 exhaustive kernel    */
    int64_t small_norm_lines;  /* lines with norm < 1e-13 (cryo_raynormalize.m:20) */
    double ms_prep;            /* preprocessing kernel                            */
    double ms_pairs;           /* main contraction (all pairs)                    */
    double ms_rescore;         /* re-score + second pass + fallback               */
    double ms_total;           /* whole device job                                */
    int32_t variant;           /* kernel variant that ran                         */
    int32_t n_eval_shifts;     /* shifts the contraction evaluated                */
    int64_t refined_pairs;     /* uncertain BF16/FP16 pairs re-run with FP32
                                  operands                                      */
} clm_stats;

/* Library version (CLM_VERSION of the build). */
int clm_version(void);

/* Fill p with cryo_clmatrix's defaults (max_shift 15, shift_step 1,
 * verbose 1, tie margin 1e-3) and FP16 re-score: FP16 runs at the BF16
 * MFMA rate with an 8x tighter rounding bound, so on high-SNR data (broad
 * correlation peaks) almost no pair needs the FP32 refinement that BF16
 * sends ~45% of pairs to at SNR 1. */
void clm_default_params(clm_params *p);

/* Number of pairs K(K-1)/2. */
int64_t clm_num_pairs(int32_t K);

/* Create a context on HIP device `device` (one HIP stream per context).
 * Replaces is_gpu()/cryo_selectGPU (common/is_gpu.m:9-11,
 * common/cryo_selectGPU.m:56).  Returns NULL and sets *err on failure. */
typedef struct clm_ctx clm_ctx;
clm_ctx *clm_create(int device, int *err);
void clm_destroy(clm_ctx *ctx);
const char *clm_last_error(const clm_ctx *ctx);

/*
 * Reports of the verbose argument (cryo_clmatrix_cpu.m:155-204).  With
 * verbose bit 1 set in clm_params, each finished call sends the call's shape
 * and then its stages (preprocessing, contraction, re-score: times, pair and
 * candidate counts) to fn with level 0; the reference prints per-pair
 * progress instead (:744-749), which a single device job has no host point
 * for.  Lines of near-zero norm (cryo_raynormalize.m:20-22) are sent with
 * level 1 whatever the verbose bits.  fn runs on the calling thread before
 * the entry point returns (the MEX gateway passes mexPrintf through it).
 * fn = NULL turns reports off (the default).
 */
typedef void (*clm_log_fn)(void *user, int level, const char *msg);
void clm_set_log(clm_ctx *ctx, clm_log_fn fn, void *user);

/*
 * The hot path: all K(K-1)/2 pairs of cryo_clmatrix_cpu.m:296-751
 * (preprocessing :96,:277-293 included).
 *
 * Host-buffer form: pf is host memory; cl1, cl2, jidx, corr, shift1d are host
 * arrays of clm_num_pairs(K) entries (any may be NULL to skip it):
 *   cl1[e]   = clstack(k1,k2) in [0, n_theta/2]   (1-based, 0 = not set)
 *   cl2[e]   = clstack(k2,k1) in [0, n_theta]
 *   jidx[e]  = 0-based shift index of the best match, -1 if not set
 *   corr[e]  = best correlation (before corrstack's 1-x of :764), 0 if not set
 *   shift1d[e] = shifts_1d(k1,k2), the RHS of equation e
 * stats may be NULL.
 */
int clm_clmatrix(clm_ctx *ctx, const clm_params *p, const void *pf,
                 int32_t *cl1, int32_t *cl2, int32_t *jidx, double *corr,
                 double *shift1d, clm_stats *stats);

/* Device-buffer form of clm_clmatrix: pf and every output are device
 * pointers (hipMalloc or torch allocations on the context's device).  Runs
 * on the context's stream and returns once the results are in place. */
int clm_clmatrix_device(clm_ctx *ctx, const clm_params *p, const void *d_pf,
                        int32_t *d_cl1, int32_t *d_cl2, int32_t *d_jidx,
                        double *d_corr, double *d_shift1d, clm_stats *stats);

/*
 * Allocation-free form of clm_clmatrix_device.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as clm_clmatrix_device, but every
 * device buffer the job needs (rays, operand planes, A-stacks, pair records,
 * lists, counters, tiles) is carved from one static device array of
 * CLM_NOALLOC_SCRATCH_BYTES that each GPU holds from library load until the
 * process ends; the call makes no hipMalloc/hipFree.  A shape whose buffers
 * do not fit returns CLM_E_UNSUP with the sizes in clm_last_error, before
 * any kernel runs.  Every call on one GPU shares that GPU's array (all
 * contexts of the process): calls that overlap on one GPU race.  Counters
 * and lists are cleared by the job itself; other contents persist from the
 * last call and are overwritten before use.
 */
#define CLM_NOALLOC_SCRATCH_BYTES ((size_t)512 << 20)
int clm_clmatrix_device_noalloc(clm_ctx *ctx, const clm_params *p,
                                const void *d_pf, int32_t *d_cl1,
                                int32_t *d_cl2, int32_t *d_jidx,
                                double *d_corr, double *d_shift1d,
                                clm_stats *stats);

/*
 * cryo_clmatrix with map_filter_radius > 0 (its sixth argument,
 * cryo_clmatrix_cpu.m:24-29,383-385): every shift's L x 2L correlation map
 * is filtered by cryo_average_clmap (sinograms/cryo_average_clmap.m:20-48,
 * default weights W0 = 1.25.^(21:-1:1)) before the first-maximum search, and
 * corr[e] is the filtered value.  Radius d in [1, 10].  Every map entry and
 * the stencil are evaluated in FP64 with the oracle's arithmetic (the
 * precision field is not used).  Host-buffer and device-buffer forms, with
 * the outputs of clm_clmatrix / clm_clmatrix_device.
 */
int clm_clmatrix_filtered(clm_ctx *ctx, const clm_params *p,
                          int32_t map_filter_radius, const void *pf,
                          int32_t *cl1, int32_t *cl2, int32_t *jidx,
                          double *corr, double *shift1d, clm_stats *stats);
int clm_clmatrix_filtered_device(clm_ctx *ctx, const clm_params *p,
                                 int32_t map_filter_radius, const void *d_pf,
                                 int32_t *d_cl1, int32_t *d_cl2,
                                 int32_t *d_jidx, double *d_corr,
                                 double *d_shift1d, clm_stats *stats);

/*
 * Several independent stacks in one device job (batched cryo_clmatrix calls,
 * e.g. the per-class or per-dataset calls a pipeline makes in a loop around
 * cryo_clmatrix, abinitio/C1/cryo_abinitio_C1_worker.m:99).  Stacks may
 * differ in K, n_r, n_theta, shift grid and precision; stacks whose kernel
 * variant agrees share one persistent launch.  io[i] holds stack i's device
 * pointers (pf and outputs, as for clm_clmatrix_device); p[i] its
 * parameters.  stats (optional) is the sum over stacks.
 * flags: CLM_GROUP_GRAPH captures the job in a HIP graph on the first call
 * and replays it on later calls with identical parameters and pointers (the
 * graph holds no memset nodes: counters are cleared by a kernel).
 */
typedef struct clm_stack_io {
    const void *pf;
    int32_t *cl1, *cl2, *jidx;
    double *corr, *shift1d;
} clm_stack_io;

#define CLM_GROUP_GRAPH 1

int clm_clmatrix_grouped(clm_ctx *ctx, int32_t n, const clm_params *p,
                         const clm_stack_io *io, int32_t flags, clm_stats *stats);

/*
 * Autotuner (no counterpart in the reference, whose GPU path has no tunable
 * kernel).  Times every pair-kernel variant that fits the shape of p -- output
 * tile width (32 x CT columns per wave), pairs per workgroup and the lagged
 * wave schedule -- on the first sample_pairs pairs (<= 0: 200000) of pf
 * (device memory, as for clm_clmatrix_device), and keeps the fastest per pair
 * for later calls on this context whose variant is 0 and whose precision,
 * n_r, n_theta and shift grid match.  entries (capacity max_entries, may be
 * NULL when 0) receive each variant's timing; *n_entries their count.
 * Leaves the outputs of earlier calls untouched.
 */
typedef struct clm_tune_entry {
    int32_t variant;  /* CT*100 + waves per workgroup                */
    int32_t lag;      /* 1 = lagged wave schedule                     */
    double ms;        /* pass-1 kernel time on the sample             */
    int64_t pairs;    /* pairs in the sample                          */
} clm_tune_entry;

int clm_autotune(clm_ctx *ctx, const clm_params *p, const void *d_pf,
                 int32_t sample_pairs, clm_tune_entry *entries,
                 int32_t max_entries, int32_t *n_entries);

/*
 * Learned form: windowed shift search (no counterpart in the reference, whose
 * search always covers the full grid of cryo_clmatrix_cpu.m:140,360-399).
 * d_prop holds one proposed 2-D shift per image (K x 2 float, image axes
 * x, y in pixels, the convention of the shift equations :722-738), e.g. from
 * the shift network below.  Each line l of image k is moved by its 1-D share
 * a_k(l) = sin(l*pi/L)*x_k + cos(l*pi/L)*y_k, rounded to a multiple of
 * shift_step, and the pair search covers only the shifts -window..window
 * around a1 - a2 (C1 columns) or -a1 - a2 (C2 columns): every window shift is
 * a shift of the reference grid.  Pairs whose window peak is below `gate`,
 * or whose best shift lies off the reference grid, are searched again over
 * the full grid (the exhaustive search is the fallback).  Outputs as for
 * clm_clmatrix_device; d_flag (optional, P bytes) receives 1 for every pair
 * that was searched over the full grid, 0 for pairs the window settled.
 * Requires max_shift and window to be multiples of shift_step; precision
 * BF16/FP16/FP32.
 */
typedef struct clm_window_stats {
    int64_t pairs;          /* K(K-1)/2                                        */
    int64_t accepted;       /* pairs settled by the windowed search            */
    int64_t below_gate;     /* window peak below the gate -> exhaustive        */
    int64_t off_grid;       /* best shift off the reference grid -> exhaustive */
    int32_t window_shifts;  /* shifts per window (2*window/step + 1)           */
    int32_t full_shifts;    /* shifts of the reference grid                    */
    double ms_window;       /* per-line shifts, windowed search, merge         */
    double ms_fallback;     /* exhaustive search of the rejected pairs         */
    double ms_total;
} clm_window_stats;

int clm_clmatrix_windowed(clm_ctx *ctx, const clm_params *p, const void *d_pf,
                          const float *d_prop, double window, double gate,
                          int32_t *d_cl1, int32_t *d_cl2, int32_t *d_jidx,
                          double *d_corr, double *d_shift1d, uint8_t *d_flag,
                          clm_window_stats *stats);

/*
 * Partner subset: the search of chosen pairs only, as
 * cryo_clmatrix_gpu_1(pf, NK, ...) (sinograms/cryo_clmatrix_gpu_1.m:22-24,
 * 346-350) searches each image against at most NK of its later images.
 * sel holds one byte per pair in equation order ((k1,k2), k1 < k2, k1
 * outer); a nonzero byte selects the pair (clm_partner_subset draws the
 * reference's subset).  Each selected pair is searched exactly as by
 * clm_clmatrix_device (cryo_clmatrix_cpu.m:360-399 semantics and parity);
 * every other pair gets cl1 = cl2 = 0, jidx = -1, corr = shift1d = 0 (the
 * reference leaves clstack and corrstack at zero there).  stats->pairs is
 * the number of selected pairs.  The _device form takes device pointers
 * (pf, sel and outputs); the other, host buffers.
 */
int clm_clmatrix_subset_device(clm_ctx *ctx, const clm_params *p,
                               const void *d_pf, const uint8_t *d_sel,
                               int32_t *d_cl1, int32_t *d_cl2, int32_t *d_jidx,
                               double *d_corr, double *d_shift1d,
                               clm_stats *stats);
int clm_clmatrix_subset(clm_ctx *ctx, const clm_params *p, const void *pf,
                        const uint8_t *sel, int32_t *cl1, int32_t *cl2,
                        int32_t *jidx, double *corr, double *shift1d,
                        clm_stats *stats);

/*
 * The partner subset of cryo_clmatrix_gpu_1.m:150-151,346-350 (host):
 * image k1 (0-based) is paired with n2 = min(K-1-k1, NK) of its K-1-k1
 * later images, drawn uniformly without replacement (every n2-subset equally
 * likely) and kept in increasing order; NK = -1 selects every pair, NK = 0
 * none.  The draw depends on (seed, k1) only: MATLAB's randperm stream is
 * not reproduced, its distribution is.  sel receives P bytes (1 = selected);
 * count (optional) the number selected.
 */
int clm_partner_subset(int32_t K, int32_t NK, uint64_t seed, uint8_t *sel,
                       int64_t *count);

/*
 * Shift network of the learned form (no counterpart in the reference).  It
 * reads the CLM_SN_NR lowest radial samples of all CLM_SN_T rays of each
 * image and proposes the image's 2-D shift for clm_clmatrix_windowed.
 * Two circular 1-D convolutions (32 -> 32 -> 8 channels, kernel 5, ReLU)
 * and a dense head (2880 -> 16 -> 2); CLM_SN_NPARAM parameters packed as
 *   W1[32][32][5] b1[32] W2[8][32][5] b2[8] F1[16][2880] c1[16] F2[2][16] c2[2].
 * Weights and stored activations are BF16 (uint16_t bit patterns), sums are
 * FP32; gradients and optimizer state FP32.  Every call runs on the
 * context's stream and returns when its results are in place.
 *   features: pf (device, as for clm_clmatrix_device; n_theta must be
 *     CLM_SN_T, n_r >= CLM_SN_NR) -> x, K x 32 x CLM_SN_T BF16: channel 2r+c
 *     is the real (c=0) or imaginary part of pf(r, t) divided by the root
 *     mean square of the image's 16 x 360 values.
 *   forward: x (B images) -> out_px, B x 2 float, proposed shift in pixels.
 *   grad: loss = 1/(2B) sum |y - target_px/8|^2 with y = out_px/8; writes
 *     the gradient (CLM_SN_NPARAM floats) and the loss (1 float).
 *   adam: one Adam step (bias-corrected, step >= 1) on the FP32 master
 *     weights, refreshing the BF16 copy.
 */
#define CLM_SN_NR 16
#define CLM_SN_T 360
#define CLM_SN_NPARAM 52570

int clm_shiftnet_nparam(void);
int clm_shiftnet_features(clm_ctx *ctx, const void *d_pf, int32_t n_r,
                          int32_t n_theta, int32_t K, int32_t pf_dtype,
                          uint16_t *d_x);
int clm_shiftnet_forward(clm_ctx *ctx, const uint16_t *d_w,
                         const uint16_t *d_x, int32_t B, float *d_out_px);
int clm_shiftnet_grad(clm_ctx *ctx, const uint16_t *d_w, const uint16_t *d_x,
                      const float *d_target_px, int32_t B, float *d_grad,
                      float *d_loss);
int clm_shiftnet_adam(clm_ctx *ctx, float *d_master, uint16_t *d_w,
                      float *d_m, float *d_v, const float *d_grad,
                      int32_t step, float lr, float beta1, float beta2,
                      float eps);

/*
 * clstack_mask, the fifth output, as the reference fills it under verbose
 * bit 2 when ref_clmatrix is given (cryo_clmatrix_cpu.m:567-617,697-698):
 * mask(k1,k2) = 1 for k1 < k2 when both estimated lines lie within 5 degrees
 * of the reference's lines (or both within 5 degrees of their antipodes),
 * else 0.  clstack, ref_clmatrix and mask are K x K column-major doubles
 * (1-based line indices in [1, n_theta]); matched (optional) receives the
 * number of matched pairs.  Host FP64.
 */
int clm_clstack_mask(int32_t K, int32_t n_theta, const double *clstack,
                     const double *ref_clmatrix, double *mask,
                     int64_t *matched);

/*
 * Host-side FP64 assembly of the reference outputs from per-pair records
 * (cryo_clmatrix_cpu.m:392-399,722-741,764-772).  Dense outputs are K x K
 * column-major doubles (MATLAB layout); any may be NULL.
 *   clstack(k1,k2)=cl1, clstack(k2,k1)=cl2; corrstack(k1,k2)=1-corr where
 *   corr!=0; shift_equations_map(k1,k2)=e.
 */
int clm_assemble(int32_t K, const int32_t *cl1, const int32_t *cl2,
                 const double *corr, double *clstack, double *corrstack,
                 double *eqmap);

/*
 * shift_equations as a compressed-sparse-column matrix of size
 * P x (2K+1) (P = K(K-1)/2), the form MATLAB's mxCreateSparse takes.
 * Explicit zeros are not stored (MATLAB sparse semantics).
 * Call with ir = pr = NULL to get *nnz; then with arrays of that size.
 * jc has 2K+2 entries.  Row indices are 0-based (mwIndex convention).
 */
int clm_shift_equations_csc(int32_t K, int32_t n_theta, const int32_t *cl1,
                            const int32_t *cl2, const double *shift1d,
                            int64_t *jc, int64_t *ir, double *pr, int64_t *nnz);

/*
 * Outputs of a partner-subset search (cryo_clmatrix_gpu_1.m:809-827,
 * 853-858): as clm_assemble and clm_shift_equations_csc, with equations for
 * the selected pairs only, numbered in pair order.  shift_equations is
 * n_sel x (2K+1) (n_sel = number of nonzero bytes of sel);
 * shift_equations_map(k1,k2) is the 1-based equation number of a selected
 * pair and 0 elsewhere.
 */
int clm_assemble_subset(int32_t K, const uint8_t *sel, const int32_t *cl1,
                        const int32_t *cl2, const double *corr,
                        double *clstack, double *corrstack, double *eqmap);
int clm_shift_equations_csc_subset(int32_t K, int32_t n_theta,
                                   const uint8_t *sel, const int32_t *cl1,
                                   const int32_t *cl2, const double *shift1d,
                                   int64_t *jc, int64_t *ir, double *pr,
                                   int64_t *nnz);

/*
 * Orientation rule of an equation's second line.  0: the angle test of
 * cryo_clmatrix_cpu.m:733 (beta = (cl2-1)*pi/L < pi keeps the row,
 * otherwise beta -= pi and the first two coefficients change sign).
 * CLM_FLIP_BY_INDEX: the index test of cryo_estimate_shifts.m:171-176,
 * 215-222 (cl2 > n_theta/2 flips).  The two differ only for cl2 =
 * n_theta/2 + 1 at the n_theta where (n_theta/2) * RN(pi/(n_theta/2)) <
 * RN(pi) (e.g. 150, 300, 350; not 360).
 */
#define CLM_FLIP_BY_INDEX 1

/*
 * clm_shift_equations_csc_ex flag: with a selection, cl1, cl2 and shift1d
 * hold only the selected pairs' records, in pair order (one per row of the
 * result), instead of K(K-1)/2 records.
 */
#define CLM_COMPACT_RECORDS 2

/*
 * clm_shift_equations_csc with an optional selection (sel null: every
 * pair; else the rows of the selected pairs, as the _subset form) and
 * flags: CLM_FLIP_BY_INDEX (the orientation rule of the second line) and
 * CLM_COMPACT_RECORDS (records of the selected pairs only; needs sel).
 */
int clm_shift_equations_csc_ex(int32_t K, int32_t n_theta, const uint8_t *sel,
                               int32_t flags, const int32_t *cl1,
                               const int32_t *cl2, const double *shift1d,
                               int64_t *jc, int64_t *ir, double *pr,
                               int64_t *nnz);

/*
 * Polar Fourier transform of a projection stack on the device: the step
 * before cryo_clmatrix.  Replaces [pf, freqs] = cryo_pft(p, n_r, n_theta)
 * (fourier/nufft/cryo_pft.m:1-75) with the exact non-uniform DFT that
 * cryo_pft's nufft2 falls back to (fourier/nufft/nudft2.m:40-47):
 *   freqs(m,:) = [k*w0*sin(j*dtheta), k*w0*cos(j*dtheta)], m = j*n_r + k,
 *       w0 = 2*pi/(2*n_r-1), dtheta = 2*pi/n_theta, j < n_theta/2 (:49-58)
 *   pf(k,j,i) = sum_{u,v} p(u,v,i) exp(+1i*(freqs(m,1)*g(u) + freqs(m,2)*g(v))),
 *       g(t) = ceil(t - n/2)  (nudft2 evaluated at -freqs, :70)
 *   pf(:, n_theta/2 + j, :) = conj(pf(:, j, :))  (:75)
 * Images are real, square n x n, column-major n x n x K (MATLAB layout,
 * element (u,v,i) at u + n*v + n*n*i), double or single; pf is written in
 * the layout clm_clmatrix reads (see the top of this file).  Arithmetic is
 * FP64 throughout (the pf of single images is rounded to complex single at
 * the end, as MATLAB returns single for single input).
 */
enum clm_real_dtype {
    CLM_R64 = 0, /* double images (MATLAB double) */
    CLM_R32 = 1  /* single images (MATLAB single) */
};

/*
 * Method: CLM_PFT_EXACT evaluates the sums above exactly (FP64 MFMA
 * contraction, O(n^2) per polar point); CLM_PFT_NUFFT evaluates them by a
 * type-2 non-uniform FFT, as nufft2 does with a NUFFT library installed
 * (fourier/nufft/nufft_transform.m:108-150; epsilon 1e-15 by default,
 * nufft_initialize.m:29): exponential-of-semicircle kernel of width 16 on
 * a 2x-oversampled grid, hipFFT for the grid transforms, O(n^2 log n) per
 * image plus 256 grid reads per polar point.  The two agree to about
 * 1e-14 of max |pf|.  n <= 2048 for the NUFFT method.
 */
enum clm_pft_method {
    CLM_PFT_EXACT = 0,
    CLM_PFT_NUFFT = 1
};

typedef struct clm_pft_params {
    int32_t n;         /* image side (square images, nudft2.m:27-31)       */
    int32_t K;         /* number of images                                 */
    int32_t n_r;       /* radial samples per ray                           */
    int32_t n_theta;   /* rays; must be even (cryo_pft.m:44-46)            */
    int32_t img_dtype; /* enum clm_real_dtype                              */
    int32_t pf_dtype;  /* enum clm_dtype of the output                     */
    int32_t method;    /* enum clm_pft_method                              */
} clm_pft_params;

typedef struct clm_pft_stats {
    float ms_tables;   /* exact: per-axis exponential tables; NUFFT:
                          kernel weights of the polar points           */
    float ms_stage;    /* exact: images to the pixel-major FP64 staging
                          area; NUFFT: deconvolution onto the grids     */
    float ms_pft;      /* exact: the contraction (FP64 MFMA); NUFFT: the
                          interpolation at the polar points            */
    int32_t chunks;    /* image chunks (exact: at most 8192 images;
                          NUFFT: at most 2048 images and 4 GiB of grid) */
    float ms_fft;      /* NUFFT: the batched 2-D FFTs                   */
    int32_t grid_side; /* NUFFT: oversampled grid side N (2,3,5-smooth,
                          >= 2n)                                        */
} clm_pft_stats;

/* Device-buffer form: d_img and d_pf are device pointers on the context's
 * device; returns once pf is written.  stats may be NULL. */
int clm_pft_device(clm_ctx *ctx, const clm_pft_params *p, const void *d_img,
                   void *d_pf, clm_pft_stats *stats);
/* Host-buffer form (img and pf in host memory). */
int clm_pft(clm_ctx *ctx, const clm_pft_params *p, const void *img, void *pf,
            clm_pft_stats *stats);
/* freqs of cryo_pft.m:52-58, (n_r*n_theta/2) x 2 column-major doubles. */
int clm_pft_freqs(int32_t n_r, int32_t n_theta, double *freqs);

/*
 * 2-D shifts from the shift equations, on the device: the step after
 * cryo_clmatrix.  Replaces
 *   est_shifts = lsqr(shift_equations(:,1:end-1), shift_equations(:,end), tol, maxit)
 * of sinograms/cryo_estimate_shifts.m:229-230 (tol 1e-8, maxit 100 there).
 * The matrix is not formed: its rows (cryo_clmatrix_cpu.m:722-741) are
 * evaluated from the per-pair outputs of clm_clmatrix_device (d_cl1, d_cl2,
 * d_shift1d, pair order k1 < k2).  LSQR (Paige and Saunders 1982) from
 * x0 = 0 with damping `damp` (1 gives the eye(2K) block of
 * sinograms/test_shifts.m:24-27, without its three gauge rows), stopping
 * when ||b - Ax|| <= tol ||b|| or ||A'(b - Ax)|| <= tol ||A||_F ||b - Ax||
 * (running estimates), or after maxit iterations.  A pair whose line
 * numbers lie outside [0, n_theta] contributes a zero row.
 * d_x (device, 2K doubles): est_shifts(k,:) = [d_x[2k], d_x[2k+1]].
 */
typedef struct clm_lsqr_stats {
    int32_t iterations; /* LSQR iterations run                              */
    int32_t flag;       /* 0: a stopping test was met; 1: maxit reached     */
    double relres;      /* estimate of ||b - Ax|| / ||b||                   */
    double normar;      /* estimate of ||A'(b - Ax)|| / (||A||_F ||b - Ax||) */
    double norm_a;      /* running Frobenius-norm estimate of A             */
    float ms;           /* device time of the solve                         */
} clm_lsqr_stats;

int clm_estimate_shifts(clm_ctx *ctx, int32_t K, int32_t n_theta,
                        const int32_t *d_cl1, const int32_t *d_cl2,
                        const double *d_shift1d, double tol, int32_t maxit,
                        double damp, double *d_x, clm_lsqr_stats *stats);

/* clm_estimate_shifts with an orientation rule (flags above). */
int clm_estimate_shifts_ex(clm_ctx *ctx, int32_t K, int32_t n_theta,
                           const int32_t *d_cl1, const int32_t *d_cl2,
                           const double *d_shift1d, double tol, int32_t maxit,
                           double damp, int32_t flags, double *d_x,
                           clm_lsqr_stats *stats);

/*
 * Common lines and 1-D shifts from known rotations: the per-pair loop of
 * cryo_estimate_shifts(pf, rotations, max_shift, shift_step, ...)
 * (sinograms/cryo_estimate_shifts.m:142-223), the equations' lines and
 * right-hand sides for clm_estimate_shifts_ex with CLM_FLIP_BY_INDEX.
 * p gives K, n_r, n_theta, pf_dtype, max_shift and shift_step (the other
 * fields are not used).  d_rot: K rotations, 3 x 3 x K doubles in MATLAB
 * order (R_k(r,c) at r + 3c + 9k).  For every pair k1 < k2 with
 * d_sel[e] != 0 (d_sel null: every pair), in equation order:
 *   [cij, cji] = commonline_R(R_k1.', R_k2.', n_theta) (:151,
 *     commonline_R.m:22-35, round half away from zero), moved to
 *     cij < n_theta/2 (:155-163),
 *     1-based: d_cl1[e] = cij in [1, n_theta/2], d_cl2[e] = cji in
 *     [1, n_theta];
 *   the rays of those lines, filtered and normalized as cryo_clmatrix does
 *   (:166-188), and C1, C2 = 2 real((ray_i .* phases)' ray_j), 2 real((conj(
 *   ray_i) .* phases)' ray_j) over the shift grid (:190-195; FP64, sums over
 *   the ray in order); d_shift1d[e] = the shift of the first maximum of C1
 *   if max C1 > max C2, else of C2 (:197-204).
 * Pairs not selected get d_cl1 = d_cl2 = -1 (a zero row for the solver)
 * and d_shift1d = 0.
 */
int clm_common_lines_from_rotations(clm_ctx *ctx, const clm_params *p,
                                    const void *d_pf, const double *d_rot,
                                    const uint8_t *d_sel, int32_t *d_cl1,
                                    int32_t *d_cl2, double *d_shift1d);

/*
 * The whole cryo_estimate_shifts on host buffers
 * (sinograms/cryo_estimate_shifts.m:1-230, less the sparse assembly, which
 * clm_shift_equations_csc_ex with CLM_FLIP_BY_INDEX does on the host): pf
 * (as for clm_clmatrix) and rot (3 x 3 x K doubles, MATLAB order) are
 * uploaded, clm_common_lines_from_rotations finds the pairs' lines and 1-D
 * shifts (sel: K(K-1)/2 host bytes, nonzero = use the pair, or null for
 * every pair), and clm_estimate_shifts_ex solves the equations
 * (lsqr(A, b, tol, maxit) of :228 with tol 1e-8 and maxit 100 there).
 * Outputs on the host: cl1, cl2, shift1d (K(K-1)/2 each; -1, -1, 0 for the
 * pairs left out) and x (2K: x[2k], x[2k+1] are the shift of image k, the
 * rows of est_shifts = transpose(reshape(x, 2, K)), :229).
 */
int clm_cryo_estimate_shifts(clm_ctx *ctx, const clm_params *p, const void *pf,
                             const double *rot, const uint8_t *sel, double tol,
                             int32_t maxit, int32_t *cl1, int32_t *cl2,
                             double *shift1d, double *x, clm_lsqr_stats *st);

/*
 * Number of equations cryo_estimate_shifts uses for K images
 * (cryo_estimate_shifts.m:58-94, in the same double arithmetic): every pair,
 * the fraction memoryfactor of them (memoryfactor <= 1), or as many as a
 * dense K(K-1)/2 x 2K double system holds in memoryfactor megabytes
 * (memoryfactor >= 100); at least K.  CLM_E_ARG for a memoryfactor below 0,
 * between 1 and 100, or NaN (the reference's error, :58-60).  Host only.
 */
int clm_n_equations(int32_t K, double memoryfactor, int64_t *n);

/*
 * A uniform draw of min(n, K(K-1)/2) pairs without replacement, as a
 * selection (sel: K(K-1)/2 bytes, 1 = drawn), the pairs that
 * cryo_estimate_shifts.m:130-137 picks with randperm (MATLAB's stream is
 * not reproduced; the draw is a function of seed).  count: pairs drawn.
 * Host only.
 */
int clm_equation_pairs(int32_t K, int64_t n, uint64_t seed, uint8_t *sel,
                       int64_t *count);

/*
 * The voting synchronization matrix on the device: the consumer of the
 * common lines in the C1 ab-initio pipeline.  Replaces
 *   S = cryo_syncmatrix_vote(clstack, n_theta)
 * (abinitio/C1/sync2n/cryo_syncmatrix_vote.m:38-57 with
 * cryo_syncmatrixIJ_vote.m, cryo_voteIJ.m and rotratio_eulerangle_vec.m,
 * no reference rotations), reading the common lines from the per-pair
 * outputs of clm_clmatrix_device (d_cl1, d_cl2).  d_S (device, 2K x 2K
 * doubles, column-major) receives S: identity diagonal blocks, the voted
 * 2x2 block of every pair at (k1, k2) and its transpose at (k2, k1).
 * cos_phi2 and c_alpha are clamped to [-1, 1] before acos.
 */
typedef struct clm_vote_stats {
    int64_t pairs;  /* K(K-1)/2                           */
    float ms_build; /* clstack tables and diagonal blocks */
    float ms_vote;  /* the vote                           */
} clm_vote_stats;

int clm_syncmatrix_vote(clm_ctx *ctx, int32_t K, int32_t n_theta,
                        const int32_t *d_cl1, const int32_t *d_cl2,
                        double *d_S, clm_vote_stats *stats);

/*
 * sync3n triangle kernels (the other consumer of the common lines,
 * abinitio/C1/cryo_abinitio_C1_worker.m:116-147), on relative rotations
 * d_R: 3 x 3 x N-choose-2 doubles in MATLAB layout (pair (i,j), i<j, 0-based,
 * at 9*e with e = (2N-i-1)i/2 + j-i-1, each matrix column-major).  Every
 * triangle i<j<k is matched in its four J-configurations and scored as in
 * the reference; per-pair sums add a pair's triangles in ascending order of
 * the third image (the reference's single-thread order).  Device pointers;
 * N in [3, 46340].  The kernels use a symmetric copy of R (9 N^2 doubles).
 *
 * clm_sync3n_triangles_scores replaces
 *   [cum_scores, scores_hist] = triangles_scores_mex(N, Rij, n_intervals, ~, 1, ~)
 * (abinitio/C1/sync3n/triangles_scores_mex.c:1-346, called from
 * triangles_scores.m:62): d_cum (N-choose-2, or NULL) gets each pair's sum
 * of its triangles' scores, d_hist (n_intervals) the histogram of all
 * 3 N-choose-3 scores, bins as the reference forms them.
 *
 * clm_sync3n_pairs_probabilities replaces
 *   [ln_f_ind, ln_f_arb] = pairs_probabilities_mex(N, Rij, P2, A, a, B, b, x0, ~, ~)
 * (pairs_probabilities_mex.c:1-339); prm (host) = {P2, A, a, B, b, x0}.
 *
 * clm_sync3n_signs_times_v replaces
 *   v2 = signs_times_v_mex(N, Rij, v, n_eigs, scores_as_entries, pairs_scores, ~)
 * (signs_times_v_mex.c:1-329), the product of the power method
 * (cryo_sync3n_Jsync_power_method.m:62): d_v and d_v2 are N-choose-2 x n_eigs
 * column-major, n_eigs in [1, 8]; scores_as_entries 0 (signs), 1 (triangle
 * scores) or 2 (products of d_pair_scores, N-choose-2).
 */
typedef struct clm_sync3n_stats {
    int64_t triangles; /* N-choose-3                                 */
    float ms_layout;   /* symmetric copies of R (and v), tables      */
    float ms_kernel;   /* the triangle kernel                        */
} clm_sync3n_stats;

int clm_sync3n_triangles_scores(clm_ctx *ctx, int32_t N, const double *d_R,
                                int32_t n_intervals, double *d_cum,
                                double *d_hist, clm_sync3n_stats *stats);
int clm_sync3n_pairs_probabilities(clm_ctx *ctx, int32_t N, const double *d_R,
                                   const double *prm, double *d_ln_f_ind,
                                   double *d_ln_f_arb, clm_sync3n_stats *stats);
int clm_sync3n_signs_times_v(clm_ctx *ctx, int32_t N, const double *d_R,
                             const double *d_v, int32_t n_eigs,
                             int32_t scores_as_entries,
                             const double *d_pair_scores, double *d_v2,
                             clm_sync3n_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* CLM_H */
