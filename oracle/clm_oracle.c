/*
 * clm_oracle.c -- FP64 CPU restatement of ASPIRE's cryo_clmatrix_cpu.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP
 * common-lines engine.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker; the product path
 * (aspire_amd/, libclm.so) never links or calls it.
 *
 * Every function follows sinograms/cryo_clmatrix_cpu.m line by line
 * (citations inline).  Arithmetic is IEEE double with -ffp-contract=off so
 * the per-element summation order is fixed: sum over r = 0..R-1 in order.
 * MATLAB's zgemm (MKL) may sum in another order, so results can differ from
 * MATLAB in the last bits; only exact ties can change an index because of it.
 *
 * Conventions (0-based C indices, 1-based values where the reference stores
 * MATLAB indices):
 *   pf   : complex double, interleaved (re,im), column-major n_r x T x K,
 *          element (r,t,k) at 2*((k*T + t)*n_r + r).
 *   P    : normalized half lines, complex double interleaved, layout
 *          [k][l][r] with l = 0..L-1 (L = T/2), r = 0..R-1 (R = n_r-1);
 *          row r carries radial index rk2 = r - R (i.e. -R..-1).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CLMO_PI 3.141592653589793238462643383279502884

/* n_shifts = ceil(2*max_shift/shift_step+1)   cryo_clmatrix_cpu.m:140 */
int clmo_n_shifts(double max_shift, double shift_step)
{
    return (int)ceil(2.0 * max_shift / shift_step + 1.0);
}

/* shift = -max_shift+(shiftidx-1)*shift_step   cryo_clmatrix_cpu.m:361 */
double clmo_shift_value(double max_shift, double shift_step, int j)
{
    return -max_shift + (double)j * shift_step;
}

/* H = sqrt(abs(rk)).*exp(-rk.^2/(2*(rmax/4).^2))   cryo_clmatrix_cpu.m:279 */
static double filt(int rk, int rmax)
{
    double r = (double)rk;
    double w = (double)rmax / 4.0;
    return sqrt(fabs(r)) * exp(-(r * r) / (2.0 * (w * w)));
}

/*
 * Line assembly (:96), filter (:279-289), band zeroing (:290), per-line
 * normalization (cryo_raynormalize.m:17-24) and half-line extraction
 * (:299,:312).  Writes P ([k][l][r] complex) and returns the number of lines
 * whose norm is below 1e-13 (the reference warns for these and still
 * divides, cryo_raynormalize.m:20-23).
 */
int64_t clmo_preprocess(const double *pf, int n_r, int T, int K, double *P)
{
    const int L = T / 2, R = n_r - 1;
    int64_t n_small = 0;
    double *H = (double *)malloc(sizeof(double) * (2 * n_r - 1));
    for (int i = 0; i < 2 * n_r - 1; i++) {
        int rk = i - R; /* rk = -rmax:rmax */
        /* rows rmax:rmax+2 (1-based) = rk in {-1,0,1} are zeroed  :290 */
        H[i] = (rk >= -1 && rk <= 1) ? 0.0 : filt(rk, R);
    }
    for (int k = 0; k < K; k++) {
        for (int l = 0; l < L; l++) {
            /* full line G(i), i=0..2n_r-2 (0-based), rk = i-R:
             * G(0..n_r-2) = flip(pf(2:n_r, L+l))  -> G(i) = pf(n_r-1-i, L+l)
             * G(n_r-1..)  = pf(1:n_r, l)          -> G(i) = pf(i-(n_r-1), l) */
            double ss = 0.0;
            for (int i = 0; i < 2 * n_r - 1; i++) {
                int rr, tt;
                if (i < n_r - 1) { rr = n_r - 1 - i; tt = L + l; }
                else             { rr = i - (n_r - 1); tt = l; }
                const double *z = pf + 2 * ((int64_t)(k * (int64_t)T + tt) * n_r + rr);
                double re = z[0] * H[i], im = z[1] * H[i];
                ss += re * re + im * im;
            }
            double nr = sqrt(ss);
            if (nr < 1.0e-13) n_small++;
            double *out = P + 2 * ((int64_t)(k * (int64_t)L + l) * R);
            for (int r = 0; r < R; r++) {
                int i = r; /* half line rows 1..rmax of G */
                int rr = n_r - 1 - i;
                const double *z = pf + 2 * ((int64_t)(k * (int64_t)T + L + l) * n_r + rr);
                out[2 * r]     = (z[0] * H[i]) / nr;
                out[2 * r + 1] = (z[1] * H[i]) / nr;
            }
        }
    }
    free(H);
    return n_small;
}

/* shift_phases = exp(-2*pi*sqrt(-1).*rk2.*shift./(2*rmax+1))  :362 */
static void phases(int R, double shift, double *ph /* R complex */)
{
    for (int r = 0; r < R; r++) {
        double rk = (double)(r - R);
        double th = ((-2.0 * CLMO_PI * rk) * shift) / (double)(2 * R + 1);
        ph[2 * r] = cos(th);
        ph[2 * r + 1] = sin(th);
    }
}

/*
 * One pair (k1,k2): the shift loop :360-399.  For each shift j:
 *   C1 = 2*real((P1.*ph)'*P2), C2 = 2*real((conj(P1).*ph)'*P2), C=[C1,C2];
 *   [sval,sidx] = max(C(:))  -- first maximum in column-major order,
 *                               NaN entries ignored (MATLAB max);
 *   if sval > best: record (cl1,cl2) = ind2sub([L 2L], sidx), shift.
 * best starts at 0 (corrstack init, :209) so a pair whose maximum is <= 0
 * keeps cl1=cl2=0 and shift 0.
 * Scratch: work must hold 4*R*L + 2*L + 2*R doubles.
 */
static void pair_eval(const double *P1, const double *P2, int R, int L,
                      double max_shift, double shift_step, int S,
                      double *work, int32_t *cl1, int32_t *cl2,
                      int32_t *jbest, double *best, double *shift1d)
{
    double *ar = work;              /* [r][l1] Re(P1.*ph)        */
    double *ai = ar + R * L;        /* [r][l1] Im(P1.*ph)        */
    double *fr = ai + R * L;        /* [r][l1] Re(conj(P1).*ph)  */
    double *fi = fr + R * L;        /* [r][l1] Im(conj(P1).*ph)  */
    double *c1 = fi + R * L;        /* one column of C1 (L)      */
    double *c2 = c1 + L;            /* one column of C2 (L)      */
    double *ph = c2 + L;            /* R complex                 */
    double bv = 0.0;
    int32_t b1 = 0, b2 = 0, bj = -1;
    double bs = 0.0;
    for (int j = 0; j < S; j++) {
        double shift = clmo_shift_value(max_shift, shift_step, j);
        phases(R, shift, ph);
        for (int l1 = 0; l1 < L; l1++) {
            const double *p = P1 + 2 * (int64_t)l1 * R;
            for (int r = 0; r < R; r++) {
                double pr = p[2 * r], pi = p[2 * r + 1];
                double qr = ph[2 * r], qi = ph[2 * r + 1];
                /* P1.*ph */
                ar[r * L + l1] = pr * qr - pi * qi;
                ai[r * L + l1] = pr * qi + pi * qr;
                /* conj(P1).*ph */
                fr[r * L + l1] = pr * qr + pi * qi;
                fi[r * L + l1] = pr * qi - pi * qr;
            }
        }
        /* first max over C(:) in column-major order: columns 0..L-1 are C1,
         * columns L..2L-1 are C2; rows l1 fastest. */
        double sv = NAN;
        int64_t si = 0;
        for (int half = 0; half < 2; half++) {
            const double *xr = half ? fr : ar, *xi = half ? fi : ai;
            double *cc = half ? c2 : c1;
            for (int l2 = 0; l2 < L; l2++) {
                const double *q = P2 + 2 * (int64_t)l2 * R;
                for (int l1 = 0; l1 < L; l1++) cc[l1] = 0.0;
                for (int r = 0; r < R; r++) {
                    double br = q[2 * r], bi = q[2 * r + 1];
                    const double *xrr = xr + r * L, *xir = xi + r * L;
                    /* real(conj(a)*b) = ar*br + ai*bi, accumulated over r */
                    for (int l1 = 0; l1 < L; l1++) {
                        double t = xrr[l1] * br;
                        double u = xir[l1] * bi;
                        cc[l1] = cc[l1] + (t + u);
                    }
                }
                int64_t col = (int64_t)half * L + l2;
                for (int l1 = 0; l1 < L; l1++) {
                    double v = 2.0 * cc[l1];
                    if (isnan(v)) continue;
                    if (isnan(sv) || v > sv) { sv = v; si = col * L + l1; }
                }
            }
        }
        if (sv > bv) { /* strict >, NaN compares false  :392 */
            bv = sv;
            b1 = (int32_t)(si % L) + 1;   /* ind2sub([L 2L],sidx)  :393 */
            b2 = (int32_t)(si / L) + 1;
            bj = j;
            bs = shift;
        }
    }
    *cl1 = b1; *cl2 = b2; *jbest = bj; *best = bv; *shift1d = bs;
}

/*
 * Evaluate the pair loop for an explicit list of pairs (k1 < k2, 0-based).
 * Outputs per pair: cl1 in [0,L] and cl2 in [0,2L] (1-based MATLAB values,
 * 0 when no correlation exceeded 0), jbest (0-based shift index or -1),
 * best (the raw maximal correlation, before corrstack's 1-x), shift1d.
 */
void clmo_pairs(const double *P, int R, int L, int K, double max_shift,
                double shift_step, const int32_t *k1s, const int32_t *k2s,
                int64_t npairs, int32_t *cl1, int32_t *cl2, int32_t *jbest,
                double *best, double *shift1d, int nthreads)
{
    (void)K;
    int S = clmo_n_shifts(max_shift, shift_step);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *work = (double *)malloc(sizeof(double) * (4 * (size_t)R * L + 2 * (size_t)L + 2 * (size_t)R));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t p = 0; p < npairs; p++) {
            const double *P1 = P + 2 * (int64_t)k1s[p] * L * R;
            const double *P2 = P + 2 * (int64_t)k2s[p] * L * R;
            pair_eval(P1, P2, R, L, max_shift, shift_step, S, work,
                      cl1 + p, cl2 + p, jbest + p, best + p, shift1d + p);
        }
        free(work);
    }
}

/*
 * Value of one entry of the correlation map, exactly as computed in the pair
 * loop: C(l1, l2p) at shift index j for images (k1,k2); l1 in [0,L),
 * l2p in [0,2L) (l2p >= L selects C2).  Used by tests to measure how close a
 * disagreement is to a tie.
 */
double clmo_entry(const double *P, int R, int L, double max_shift,
                  double shift_step, int k1, int k2, int j, int l1, int l2p)
{
    double ph[2 * 1024];
    if (R > 1024) return NAN;
    phases(R, clmo_shift_value(max_shift, shift_step, j), ph);
    const double *p = P + 2 * ((int64_t)k1 * L + l1) * R;
    const double *q = P + 2 * ((int64_t)k2 * L + (l2p % L)) * R;
    int flip = l2p >= L;
    double acc = 0.0;
    for (int r = 0; r < R; r++) {
        double pr = p[2 * r], pi = p[2 * r + 1], qr = ph[2 * r], qi = ph[2 * r + 1];
        double xr, xi;
        if (!flip) { xr = pr * qr - pi * qi; xi = pr * qi + pi * qr; }
        else       { xr = pr * qr + pi * qi; xi = pr * qi - pi * qr; }
        double t = xr * q[2 * r], u = xi * q[2 * r + 1];
        acc = acc + (t + u);
    }
    return 2.0 * acc;
}

/*
 * The whole correlation map C = [C1, C2] (L x 2L, column-major: C(l1, l2p) at
 * map[l2p*L + l1]) of pair (k1, k2) at shift index j, each entry computed as
 * clmo_entry computes it.
 */
void clmo_map(const double *P, int R, int L, double max_shift, double shift_step,
              int k1, int k2, int j, double *map)
{
    for (int l2p = 0; l2p < 2 * L; l2p++)
        for (int l1 = 0; l1 < L; l1++)
            map[(int64_t)l2p * L + l1] = clmo_entry(P, R, L, max_shift, shift_step, k1, k2, j, l1, l2p);
}

/*
 * cryo_average_clmap(map, d) with its default weights
 * (sinograms/cryo_average_clmap.m:20-48), as the direct sum its FFT
 * evaluates:
 *   W0 = 1.25.^(21:-1:1);  W(i,j) = W0(1 + |i-d| + |j-d|), i,j = 0..2d   (:24-26)
 *   map_ext = [map; map(:, [L+1:2L, 1:L])]  (2L x 2L)                   (:30)
 *   out(l1, c) = sum_i sum_j W(i,j) * map_ext((l1-d+i) mod 2L, (c-d+j) mod 2L)
 * for l1 < L, c < 2L -- the circular cross-correlation of :33-35 moved by d
 * (:48).  Sums run i outer, j inner.  out is L x 2L column-major.
 */
void clmo_filter_map(const double *map, int L, int d, double *out)
{
    const int w = 2 * d + 1, L2 = 2 * L;
    double W[21 * 21];
    for (int i = 0; i < w; i++)
        for (int j = 0; j < w; j++)
            W[i * w + j] = pow(1.25, (double)(21 - (abs(i - d) + abs(j - d))));
    for (int c = 0; c < L2; c++)
        for (int l1 = 0; l1 < L; l1++) {
            double v = 0.0;
            for (int i = 0; i < w; i++) {
                const int rr = ((l1 - d + i) % L2 + L2) % L2;
                for (int j = 0; j < w; j++) {
                    int cc = ((c - d + j) % L2 + L2) % L2;
                    int r = rr;
                    if (r >= L) { r -= L; cc = (cc + L) % L2; }  /* rows L..2L-1: halves swapped */
                    v = v + W[i * w + j] * map[(int64_t)cc * L + r];
                }
            }
            out[(int64_t)c * L + l1] = v;
        }
}

/*
 * The pair loop with map_filter_radius = d > 0 (cryo_clmatrix_cpu.m:383-385):
 * every shift's map is filtered before the first-max search, and the stored
 * correlation is the filtered value.  Outputs as clmo_pairs.
 */
void clmo_pairs_filtered(const double *P, int R, int L, int K, double max_shift,
                         double shift_step, int d, const int32_t *k1s, const int32_t *k2s,
                         int64_t npairs, int32_t *cl1, int32_t *cl2, int32_t *jbest,
                         double *best, double *shift1d, int nthreads)
{
    (void)K;
    const int S = clmo_n_shifts(max_shift, shift_step);
    const size_t msz = (size_t)2 * L * L;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *map = (double *)malloc(sizeof(double) * msz);
        double *out = (double *)malloc(sizeof(double) * msz);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t p = 0; p < npairs; p++) {
            double bv = 0.0, bs = 0.0;
            int32_t b1 = 0, b2 = 0, bj = -1;
            for (int j = 0; j < S; j++) {
                clmo_map(P, R, L, max_shift, shift_step, k1s[p], k2s[p], j, map);
                clmo_filter_map(map, L, d, out);
                double sv = NAN;
                int64_t si = 0;
                for (int64_t q = 0; q < (int64_t)msz; q++) {
                    const double v = out[q];
                    if (isnan(v)) continue;
                    if (isnan(sv) || v > sv) { sv = v; si = q; }
                }
                if (sv > bv) {
                    bv = sv;
                    b1 = (int32_t)(si % L) + 1;
                    b2 = (int32_t)(si / L) + 1;
                    bj = j;
                    bs = clmo_shift_value(max_shift, shift_step, j);
                }
            }
            cl1[p] = b1; cl2[p] = b2; jbest[p] = bj; best[p] = bv; shift1d[p] = bs;
        }
        free(map);
        free(out);
    }
}

/*
 * Shift-equation rows :722-741.  For each pair in (k1 asc, k2 asc) order,
 * given cl1=clstack(k1,k2), cl2=clstack(k2,k1) (1-based, possibly 0) writes
 * four coefficients and four 1-based column indices.  dtheta = pi/L (:245).
 */
void clmo_equations(int L, const int32_t *k1s, const int32_t *k2s,
                    const int32_t *cl1, const int32_t *cl2, int64_t npairs,
                    double *coef, int32_t *cols)
{
    double dtheta = CLMO_PI / (double)L;
    for (int64_t p = 0; p < npairs; p++) {
        double a = (double)(cl1[p] - 1) * dtheta;
        double b = (double)(cl2[p] - 1) * dtheta;
        double *c = coef + 4 * p;
        if (b < CLMO_PI) {
            c[0] = sin(a); c[1] = cos(a); c[2] = -sin(b); c[3] = -cos(b);
        } else {
            b = b - CLMO_PI;
            c[0] = -sin(a); c[1] = -cos(a); c[2] = -sin(b); c[3] = -cos(b);
        }
        int32_t *q = cols + 4 * p;
        q[0] = 2 * (k1s[p] + 1) - 1; q[1] = 2 * (k1s[p] + 1);
        q[2] = 2 * (k2s[p] + 1) - 1; q[3] = 2 * (k2s[p] + 1);
    }
}

int clmo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
