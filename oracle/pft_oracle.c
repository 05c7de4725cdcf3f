/*
 * pft_oracle.c -- FP64 CPU restatement of ASPIRE's cryo_pft with the exact
 * non-uniform DFT (the reference's fallback when no NUFFT library exists).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for libclm's
 * clm_pft.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker; the product path (aspire_amd/,
 * libclm.so) never links or calls it.
 *
 * Follows fourier/nufft/cryo_pft.m:49-75 and fourier/nufft/nudft2.m:37-47:
 *   freqs((j-1)*n_r + k, :) = [(k-1)*omega0*sin((j-1)*dtheta),
 *                              (k-1)*omega0*cos((j-1)*dtheta)]   (:52-58)
 *   pf = nufft2(p, -freqs')                                      (:70)
 *   nudft2: grid = ceil([-N/2:N/2-1]); [grid_x, grid_y] = ndgrid(grid, grid);
 *           pts = [grid_x(:) grid_y(:)]';
 *           im_f(k,:) = exp(-i*(fourier_pts(:,k)'*pts)) * im_to_vec(im)
 *   pf = reshape(pf, [n_r n_theta/2 n_proj]); pf = cat(2, pf, conj(pf)) (:73-75)
 * pts is column-major over the image (first index fastest), so pixel
 * (u, v) of image i is p[u + n*v + n*n*i] and sits at (grid[u], grid[v]).
 * The phase is formed as MATLAB forms fourier_pts(:,k)'*pts (two products,
 * one sum) and each sum runs over the pixels in that column-major order.
 *
 * Layouts: p is n x n x K doubles, column-major; pf is complex double,
 * interleaved, column-major n_r x n_theta x K; freqs is (n_r*n_theta/2) x 2
 * column-major.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PFTO_PI 3.141592653589793238462643383279502884

void clmo_pft_freqs(int n_r, int n_theta, double *freqs)
{
    const int L = n_theta / 2;
    const int64_t M = (int64_t)n_r * L;
    const double omega0 = 2 * PFTO_PI / (2 * n_r - 1); /* cryo_pft.m:49 */
    const double dtheta = 2 * PFTO_PI / n_theta;       /* cryo_pft.m:50 */
    for (int j = 0; j < L; j++)                      /* cryo_pft.m:53-58 */
        for (int k = 0; k < n_r; k++) {
            freqs[(int64_t)j * n_r + k] = k * omega0 * sin(j * dtheta);
            freqs[M + (int64_t)j * n_r + k] = k * omega0 * cos(j * dtheta);
        }
}

/* pf of K images: one row of exp(-i*fourier_pts'*pts) at a time (shared by
 * every image), then its product with each image. */
void clmo_pft(const double *p, int n, int K, int n_r, int n_theta, double *pf, int nthreads)
{
    const int L = n_theta / 2;
    const int64_t M = (int64_t)n_r * L;
    const int64_t N2 = (int64_t)n * n;
    double *freqs = malloc(sizeof(double) * 2 * M);
    double *grid = malloc(sizeof(double) * n);
    clmo_pft_freqs(n_r, n_theta, freqs);
    for (int t = 0; t < n; t++) grid[t] = ceil(-n / 2.0 + t); /* nudft2.m:40 */
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel
    {
        double *er = malloc(sizeof(double) * N2), *ei = malloc(sizeof(double) * N2);
#pragma omp for schedule(dynamic, 4)
        for (int64_t m = 0; m < M; m++) {
            /* fourier_pts = -freqs' (cryo_pft.m:70) */
            const double fx = -freqs[m], fy = -freqs[M + m];
            for (int v = 0; v < n; v++)
                for (int u = 0; u < n; u++) {
                    const double arg = fx * grid[u] + fy * grid[v];
                    /* exp(-i*arg) */
                    er[u + (int64_t)n * v] = cos(arg);
                    ei[u + (int64_t)n * v] = -sin(arg);
                }
            const int64_t j = m / n_r, k = m % n_r;
            for (int i = 0; i < K; i++) {
                const double *img = p + (int64_t)i * N2;
                double re = 0.0, im = 0.0;
                for (int64_t q = 0; q < N2; q++) {
                    re += er[q] * img[q];
                    im += ei[q] * img[q];
                }
                const int64_t a = ((int64_t)i * n_theta + j) * n_r + k;
                const int64_t b = ((int64_t)i * n_theta + j + L) * n_r + k;
                pf[2 * a] = re;
                pf[2 * a + 1] = im;
                pf[2 * b] = re; /* cat(2, pf, conj(pf)), cryo_pft.m:75 */
                pf[2 * b + 1] = -im;
            }
        }
        free(er);
        free(ei);
    }
    free(freqs);
    free(grid);
}
