/*
 * vote_oracle.c -- FP64 CPU restatement of ASPIRE's common-lines voting
 * synchronization matrix, the consumer of cryo_clmatrix's clstack.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for libclm's
 * clm_syncmatrix_vote.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker; the product path
 * (aspire_amd/, libclm.so) never links or calls it.
 *
 * Follows abinitio/C1/sync2n/cryo_syncmatrix_vote.m:38-57 (S = eye(2K), the
 * 2x2 block of every pair and its transpose),
 * cryo_syncmatrixIJ_vote.m:24-31,125-164 without reference rotations,
 * cryo_voteIJ.m:36-124 (angles of the triangles, Gaussian histogram on 60
 * bins over [0,180] degrees with sigma 3, peak, images within 360/60 degrees
 * of it) and rotratio_eulerangle_vec.m:22-78 (relative rotation of each
 * voted triangle; its top-left 2x2 block averaged).  Sums run over k3 in
 * ascending order, as MATLAB's sum over the selected images does.
 *
 * Two choices where MATLAB's behaviour would leave the reals: cos_phi2 and
 * c_alpha are clamped to [-1, 1] before acos (cryo_voteIJ.m clamps only when
 * every excess is below 1e-12 and otherwise continues with complex angles;
 * rotratio_eulerangle_vec.m does not clamp).  x = linspace(0,180,60) is
 * formed as d*180/59 for d < 59 and exactly 180 for the last point.
 *
 * Layouts: cl is the K x K clstack as doubles (column-major, cl(i,j) at
 * cl[j*K + i]; 0 = no common line); S is 2K x 2K column-major.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VOTO_PI 3.141592653589793238462643383279502884
#define NTICS 60

static inline double CL(const double *cl, int K, int i, int j) { return cl[(int64_t)j * K + i]; }

static void ticks(double *x)
{
    for (int d = 0; d < NTICS - 1; d++) x[d] = d * (180.0 / (NTICS - 1));
    x[NTICS - 1] = 180.0;
}

/* cryo_voteIJ.m:40-63: the angle (degrees) of triangle (k1,k2,k3), or NaN
 * when k3 takes no part (no line, or cond <= 1e-5). */
static double vote_angle(const double *cl, int K, double L, int k1, int k2, int k3)
{
    if (k3 == k1 || CL(cl, K, k1, k3) == 0 || CL(cl, K, k2, k3) == 0) return NAN;
    const double l12 = CL(cl, K, k1, k2), l21 = CL(cl, K, k2, k1);
    const double l13 = CL(cl, K, k1, k3), l31 = CL(cl, K, k3, k1);
    const double l23 = CL(cl, K, k2, k3), l32 = CL(cl, K, k3, k2);
    const double theta1 = (l13 - l12) * 2 * VOTO_PI / L;
    const double theta2 = (l21 - l23) * 2 * VOTO_PI / L;
    const double theta3 = (l32 - l31) * 2 * VOTO_PI / L;
    const double c1 = cos(theta1), c2 = cos(theta2), c3 = cos(theta3);
    const double cond = 1 + 2 * c1 * c2 * c3 - (c1 * c1 + c2 * c2 + c3 * c3);
    if (!(cond > 1.0e-5)) return NAN;
    double cp = (c3 - c1 * c2) / (sin(theta1) * sin(theta2));
    if (cp > 1) cp = 1;
    if (cp < -1) cp = -1;
    return acos(cp) * 180 / VOTO_PI;
}

/* The 2x2 block R22 of pair (k1, k2): R22 = [r00 r01; r10 r11]. */
void clmo_vote_pair(const double *cl, int K, double L, int k1, int k2, double *r22)
{
    r22[0] = r22[1] = r22[2] = r22[3] = 0.0;
    if (k1 == k2 || CL(cl, K, k1, k2) == 0) return;
    double x[NTICS], h[NTICS];
    ticks(x);
    for (int d = 0; d < NTICS; d++) h[d] = 0.0;
    const double sigma = 3.0;
    int any = 0;
    for (int k3 = 0; k3 < K; k3++) {
        const double a = vote_angle(cl, K, L, k1, k2, k3);
        if (isnan(a)) continue;
        any = 1;
        for (int d = 0; d < NTICS; d++) {
            double t = 2 * a * x[d];                   /* h = 2*angles*x        :98  */
            t = t - a * a;                             /* h - angles.^2         :99  */
            t = t - x[d] * x[d];                       /* h - x.^2              :100 */
            h[d] = h[d] + exp(t / (2 * sigma * sigma)); /* exp(h./(2*sigma.^2)) :101 */
        }
    }
    if (!any) return;
    int peak = 0;
    for (int d = 1; d < NTICS; d++)
        if (h[d] > h[peak]) peak = d;
    /* rotratio_eulerangle_vec over the voted images */
    const double n = L;
    const double l12 = CL(cl, K, k1, k2), l21 = CL(cl, K, k2, k1);
    const double aa = (l12 - 1) * 2 * VOTO_PI / n, bb = (l21 - 1) * 2 * VOTO_PI / n;
    const double ang1 = VOTO_PI - bb, ang3 = aa - VOTO_PI;
    const double sa = sin(ang1), ca = cos(ang1), sc = sin(ang3), cc = cos(ang3);
    const double TOL_idx = 1e-12;
    double s00 = 0, s01 = 0, s10 = 0, s11 = 0;
    int cnt = 0;
    for (int k3 = 0; k3 < K; k3++) {
        const double ang = vote_angle(cl, K, L, k1, k2, k3);
        if (isnan(ang) || !(fabs(ang - x[peak]) < 360.0 / NTICS)) continue;
        cnt++;
        const double idx1 = CL(cl, K, k3, k2) - CL(cl, K, k3, k1);
        const double idx2 = CL(cl, K, k2, k3) - CL(cl, K, k2, k1);
        const double idx3 = CL(cl, K, k1, k3) - CL(cl, K, k1, k2);
        const double a = cos(2 * VOTO_PI * idx1 / n);
        const double b = cos(2 * VOTO_PI * idx2 / n);
        const double c = cos(2 * VOTO_PI * idx3 / n);
        const double cond = 1 + 2 * a * b * c - (a * a + b * b + c * c);
        if (!(cond > 1.0e-5)) continue; /* R = 0, still counted */
        double c_alpha = (a - b * c) / sqrt(1 - b * b) / sqrt(1 - c * c);
        const int ind1 = (idx3 > n / 2 + TOL_idx) || (idx3 < -TOL_idx && idx3 > -n / 2);
        const int ind2 = (idx2 > n / 2 + TOL_idx) || (idx2 < -TOL_idx && idx2 > -n / 2);
        if (ind1 != ind2) c_alpha = -c_alpha;
        if (c_alpha > 1) c_alpha = 1;
        if (c_alpha < -1) c_alpha = -1;
        const double cb = cos(acos(c_alpha));
        s00 = s00 + (cc * ca - sc * cb * sa);
        s01 = s01 + (-cc * sa - sc * cb * ca);
        s10 = s10 + (sc * ca + cc * cb * sa);
        s11 = s11 + (-sa * sc + cc * cb * ca);
    }
    if (cnt == 0) return;
    r22[0] = s00 / cnt;
    r22[1] = s01 / cnt;
    r22[2] = s10 / cnt;
    r22[3] = s11 / cnt;
}

void clmo_syncmatrix_vote(const double *cl, int K, double L, double *S, int nthreads)
{
    const int64_t n2 = 2 * (int64_t)K;
    memset(S, 0, sizeof(double) * n2 * n2);
    for (int64_t i = 0; i < n2; i++) S[i * n2 + i] = 1.0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int k1 = 0; k1 < K - 1; k1++)
        for (int k2 = k1 + 1; k2 < K; k2++) {
            double r[4];
            clmo_vote_pair(cl, K, L, k1, k2, r);
            /* S(2k1+a, 2k2+b) = R22(a,b); S(2k2+b, 2k1+a) = R22(a,b) */
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    S[(2 * (int64_t)k2 + b) * n2 + 2 * k1 + a] = r[2 * a + b];
                    S[(2 * (int64_t)k1 + a) * n2 + 2 * k2 + b] = r[2 * a + b];
                }
        }
}
