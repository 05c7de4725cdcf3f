/*
 * sync3n_oracle.c -- FP64 CPU restatement of ASPIRE's sync3n triangle
 * kernels, the consumer of cryo_clmatrix's common lines in the C1 ab-initio
 * pipeline (abinitio/C1/cryo_abinitio_C1_worker.m:116-147).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for libclm's
 * clm_sync3n_* entry points.  Only tests/ and bench.py's cpu_baseline leg may
 * load it, and only as the checker; the product path never links or calls it.
 *
 * The reference's kernels are MEX files (abinitio/C1/sync3n/triangles_scores_mex.c and its two siblings): they
 * need MATLAB's mex.h/matrix.h, which this image lacks, so they cannot be
 * built here and are restated instead:
 *   triangles_scores_mex.c:83-204     -> s3o_triangles_scores
 *   pairs_probabilities_mex.c:84-210  -> s3o_pairs_probabilities
 *   signs_times_v_mex.c:84-213        -> s3o_signs_times_v
 * For every triangle i<j<k of the relative rotations R_pairs (3 x 3 x
 * N-choose-2, column-major, pair (i,j) at PAIR_IDX), four J-configurations
 * are matched (||Rij Rjk - Rik||^2, with J Rij J, J Rjk J, or J Rik J); the
 * best one, and per pair of sides its best alternative, give three scores
 * 1 - sqrt(best/alternative).  The arithmetic follows the reference
 * operation for operation (3x3 products as (a + b) + c, squared norms summed
 * in index order, no fused multiply-add).
 *
 * Accumulation order: with nthreads == 1 the triangles are visited in the
 * reference's loop order (i, then j, then k ascending), which for every pair
 * adds its triangles in ascending order of the third image -- the order
 * libclm's kernels use, so results are bit-identical.  With nthreads > 1 the
 * i range is split as the reference splits it (:262-275 of
 * triangles_scores_mex.c) into per-thread arrays summed in thread order,
 * which reproduces the reference at that thread count (timing baseline).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int64_t pidx(int64_t N, int64_t i, int64_t j) { return (2 * N - i - 1) * i / 2 + j - i - 1; }

static inline void mult3(double *out, const double *A, const double *B) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[3 * j + i] = A[3 * 0 + i] * B[3 * j + 0] + A[3 * 1 + i] * B[3 * j + 1] + A[3 * 2 + i] * B[3 * j + 2];
}

static inline void jrj(const double *R, double *A) {
    A[0] = R[0]; A[1] = R[1]; A[2] = -R[2];
    A[3] = R[3]; A[4] = R[4]; A[5] = -R[5];
    A[6] = -R[6]; A[7] = -R[7]; A[8] = R[8];
}

static inline double dnorm(const double *A, const double *B) {
    double n = 0;
    for (int i = 0; i < 9; i++) n += (A[i] - B[i]) * (A[i] - B[i]);
    return n;
}

/* ALTS[a][best][side pair] and the sign configurations (triangles_scores_mex.c:238-240,
   signs_times_v_mex.c:245-249) */
static const int ALTS[2][4][3] = {{{1, 2, 1}, {0, 2, 0}, {0, 0, 1}, {1, 0, 0}},
                                  {{2, 3, 3}, {3, 3, 2}, {3, 1, 3}, {2, 1, 2}}};
static const int SIGNS[4][3] = {{1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, -1}};

/* The triangle (Rij, Rjk, Rik): its best configuration and the scores of its
   side pairs s[0] = (ij,jk), s[1] = (ik,jk), s[2] = (ij,ik). */
static inline int triangle(const double *Rij, const double *Rjk, const double *Rik, double c[4], double s[3]) {
    double JRijJ[9], JRjkJ[9], JRikJ[9], tmp[9];
    jrj(Rij, JRijJ);
    jrj(Rjk, JRjkJ);
    jrj(Rik, JRikJ);
    mult3(tmp, Rij, Rjk);
    c[0] = dnorm(tmp, Rik);
    mult3(tmp, JRijJ, Rjk);
    c[1] = dnorm(tmp, Rik);
    mult3(tmp, Rij, JRjkJ);
    c[2] = dnorm(tmp, Rik);
    mult3(tmp, Rij, Rjk);
    c[3] = dnorm(tmp, JRikJ);
    int b = 0;
    double bv = c[0];
    if (c[1] < bv) { b = 1; bv = c[1]; }
    if (c[2] < bv) { b = 2; bv = c[2]; }
    if (c[3] < bv) { b = 3; bv = c[3]; }
    for (int q = 0; q < 3; q++) {
        double alt = c[ALTS[0][b][q]];
        if (c[ALTS[1][b][q]] < alt) alt = c[ALTS[1][b][q]];
        s[q] = 1 - sqrt(bv / alt);
    }
    return b;
}

static inline int bin_of(double s, const double *thr, int n) {
    int l;
    for (l = 0; l < n - 1; l++)
        if (s < thr[l]) break;
    return l;
}

/* i ranges of the reference's threads (triangles_scores_mex.c:262-275) */
static int split(int64_t N, int nth, unsigned *starts, unsigned *stops) {
    const unsigned long long ntri = (unsigned long long)N * (N - 1) * (N - 2) / 6;
    const unsigned long long per = (ntri + nth - 1) / nth;
    unsigned stop = 0;
    for (int t = 0; t < nth; t++) {
        unsigned start = stop, i;
        unsigned long long n = 0;
        for (i = start; i < N - 2; i++) {
            n += (unsigned long long)(N - i - 1) * (N - i - 2) / 2;
            if (n >= per) break;
        }
        i++;
        stop = i;
        if (stop > N - 2) stop = (unsigned)(N - 2);
        starts[t] = start;
        stops[t] = stop;
    }
    return nth;
}

#define MODE_SCORES 0
#define MODE_PROBS 1
#define MODE_SIGNS 2

typedef struct {
    int mode;
    int n_intervals;
    const double *thr;
    double P2, A, a, B, b, x0;
    const double *v;
    int n_eigs;
    int entries;
    const double *pair_scores;
} Job;

static void loop(int64_t N, const double *R, unsigned start, unsigned stop, const Job *jb, double *o1, double *o2) {
    const int64_t P = N * (N - 1) / 2;
    for (int64_t i = start; i < stop; i++)
        for (int64_t j = i + 1; j < N - 1; j++) {
            const int64_t ij = pidx(N, i, j);
            for (int64_t k = j + 1; k < N; k++) {
                const int64_t jk = pidx(N, j, k), ik = pidx(N, i, k);
                double c[4], s[3];
                const int b = triangle(R + 9 * ij, R + 9 * jk, R + 9 * ik, c, s);
                if (jb->mode == MODE_SCORES) {
                    if (o1) {
                        o1[ij] += s[0] + s[2];
                        o1[jk] += s[0] + s[1];
                        o1[ik] += s[1] + s[2];
                    }
                    for (int q = 0; q < 3; q++) o2[bin_of(s[q], jb->thr, jb->n_intervals)] += 1;
                } else if (jb->mode == MODE_PROBS) {
                    double fi[3], fa[3];
                    for (int q = 0; q < 3; q++) {
                        const double x = s[q];
                        fi[q] = log(jb->P2 * (jb->B * pow(1 - x, jb->b) * exp(-jb->b / (1 - jb->x0) * (1 - x))) +
                                    (1 - jb->P2) * jb->A * pow((1 - x), jb->a));
                        fa[q] = log(jb->A * pow((1 - x), jb->a));
                    }
                    o1[ij] += fi[0] + fi[2];
                    o1[jk] += fi[0] + fi[1];
                    o1[ik] += fi[1] + fi[2];
                    o2[ij] += fa[0] + fa[2];
                    o2[jk] += fa[0] + fa[1];
                    o2[ik] += fa[1] + fa[2];
                } else {
                    double e[3];
                    if (jb->entries == 1) {
                        for (int q = 0; q < 3; q++) e[q] = SIGNS[b][q] * s[q];
                    } else if (jb->entries == 2) {
                        const double t = jb->pair_scores[ij] * jb->pair_scores[jk] * jb->pair_scores[ik];
                        for (int q = 0; q < 3; q++) e[q] = SIGNS[b][q] * t;
                    } else {
                        for (int q = 0; q < 3; q++) e[q] = SIGNS[b][q];
                    }
                    for (int g = 0; g < jb->n_eigs; g++) {
                        const double *v = jb->v + P * g;
                        double *v2 = o1 + P * g;
                        v2[ij] += e[0] * v[jk] + e[2] * v[ik];
                        v2[jk] += e[0] * v[ij] + e[1] * v[ik];
                        v2[ik] += e[2] * v[ij] + e[1] * v[jk];
                    }
                }
            }
        }
}

static int run(int64_t N, const double *R, const Job *jb, double *o1, size_t n1, double *o2, size_t n2, int nthreads) {
    if (N < 3) return 0;
    if (o1) memset(o1, 0, n1 * sizeof(double));
    if (o2) memset(o2, 0, n2 * sizeof(double));
    if (nthreads <= 1) {
        loop(N, R, 0, (unsigned)(N - 2), jb, o1, o2);
        return 1;
    }
    unsigned *st = (unsigned *)malloc(sizeof(unsigned) * nthreads), *sp = (unsigned *)malloc(sizeof(unsigned) * nthreads);
    split(N, nthreads, st, sp);
    double **p1 = (double **)calloc(nthreads, sizeof(double *)), **p2 = (double **)calloc(nthreads, sizeof(double *));
    for (int t = 0; t < nthreads; t++) {
        if (o1) p1[t] = (double *)calloc(n1, sizeof(double));
        if (o2) p2[t] = (double *)calloc(n2, sizeof(double));
    }
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int t = 0; t < nthreads; t++) loop(N, R, st[t], sp[t], jb, p1[t], p2[t]);
    for (int t = 0; t < nthreads; t++) {
        for (size_t x = 0; o1 && x < n1; x++) o1[x] += p1[t][x];
        for (size_t x = 0; o2 && x < n2; x++) o2[x] += p2[t][x];
        free(p1[t]);
        free(p2[t]);
    }
    free(p1);
    free(p2);
    free(st);
    free(sp);
    return nthreads;
}

/* Thresholds of the histogram: threshold += h, n-1 times (triangles_scores_mex.c:179-183). */
static void thresholds(int n, double *thr) {
    const double h = 1. / n;
    double t = 0;
    for (int l = 0; l < n - 1; l++) {
        t += h;
        thr[l] = t;
    }
}

int s3o_triangles_scores(int64_t N, const double *R, int n_intervals, double *cum, double *hist, int nthreads) {
    if (n_intervals < 1) return -1;
    double *thr = (double *)malloc(sizeof(double) * (n_intervals > 1 ? n_intervals - 1 : 1));
    thresholds(n_intervals, thr);
    Job jb = {0};
    jb.mode = MODE_SCORES;
    jb.n_intervals = n_intervals;
    jb.thr = thr;
    const int used = run(N, R, &jb, cum, cum ? (size_t)(N * (N - 1) / 2) : 0, hist, (size_t)n_intervals, nthreads);
    free(thr);
    return used;
}

int s3o_pairs_probabilities(int64_t N, const double *R, const double *prm, double *ln_f_ind, double *ln_f_arb,
                            int nthreads) {
    Job jb = {0};
    jb.mode = MODE_PROBS;
    jb.P2 = prm[0]; jb.A = prm[1]; jb.a = prm[2]; jb.B = prm[3]; jb.b = prm[4]; jb.x0 = prm[5];
    const size_t P = (size_t)(N * (N - 1) / 2);
    return run(N, R, &jb, ln_f_ind, P, ln_f_arb, P, nthreads);
}

int s3o_signs_times_v(int64_t N, const double *R, const double *v, int n_eigs, int scores_as_entries,
                      const double *pair_scores, double *v2, int nthreads) {
    Job jb = {0};
    jb.mode = MODE_SIGNS;
    jb.v = v;
    jb.n_eigs = n_eigs;
    jb.entries = scores_as_entries;
    jb.pair_scores = pair_scores;
    const size_t P = (size_t)(N * (N - 1) / 2);
    return run(N, R, &jb, v2, P * (size_t)n_eigs, NULL, 0, nthreads);
}

/* One triangle, for tests: c[4], s[3], returns the best configuration. */
int s3o_triangle(const double *Rij, const double *Rjk, const double *Rik, double *c, double *s) {
    return triangle(Rij, Rjk, Rik, c, s);
}
