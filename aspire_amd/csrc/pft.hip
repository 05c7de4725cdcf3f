// pft.hip -- polar Fourier transform of a projection stack on the device:
// the step before the common-lines search.
//
// Replaces cryo_pft(p, n_r, n_theta) (fourier/nufft/cryo_pft.m:1-75) with its
// exact evaluation, the non-uniform DFT of fourier/nufft/nudft2.m:40-47 (the
// reference's own fallback when no NUFFT library is installed):
//
//   freqs(m) = ((k)*w0*sin(j*dtheta), (k)*w0*cos(j*dtheta)),  m = j*n_r + k,
//       w0 = 2*pi/(2*n_r - 1), dtheta = 2*pi/n_theta, j < n_theta/2  (:49-58)
//   pf(k, j, i) = sum_{u,v} p(u, v, i) * exp(+1i*(fx(m)*g(u) + fy(m)*g(v)))
//       g(t) = ceil(t - n/2)      (nudft2.m:40-47 evaluated at -freqs, :70)
//   pf(:, n_theta/2 + j, :) = conj(pf(:, j, :))                        (:75)
//
// u is the first (row) index of an image, paired with fx = r*sin(theta); v is
// the second, paired with fy = r*cos(theta).
//
// On the device this is a contraction over the n*n pixels of a complex weight
// matrix E (M x n^2, M = n_r*n_theta/2) with the real image stack (n^2 x K):
// FP64 MFMA (v_mfma_f64_16x16x4f64).  E is never stored: the exponential
// factors per axis, E(m,(u,v)) = ax(u,m) * ay(v,m), come from two small tables
// and a per-point rotation, and each lane forms its operand with complex
// products in registers.  Images are first staged pixel-major (pixel rows of Kp images,
// FP64) so the B fragments load as 128-byte rows.
//
// CLM_PFT_NUFFT: the same sums by a type-2 non-uniform FFT, the method the
// reference's nufft2 uses when a NUFFT library is installed
// (fourier/nufft/nufft_transform.m:108-150, finufft with epsilon 1e-15 by
// default, nufft_initialize.m:29).  Own implementation of the published
// exponential-of-semicircle method (Barnett, Magland and af Klinteberg, SIAM
// J. Sci. Comput. 41(5), 2019): kernel phi(z) = exp(beta (sqrt(1 - z^2) - 1))
// of width NW = 16 grid points, beta = 2.30 NW, on an oversampled grid of
// side N >= 2n (2,3,5-smooth).  Per image:
//   a(x mod N, y mod N) = p(u, v) / (Phi(x) Phi(y)),  x = g(u), y = g(v)
//   b = the unnormalized backward 2-D FFT of a (sum a e^{+2 pi i l x / N})
//   pf(m) = sum_{t, r < NW} phi(2 (sx - lx - t)/NW) phi(2 (sy - ly - r)/NW) b(lx + t, ly + r)
// with sx = fx(m) N / (2 pi), lx = ceil(sx - NW/2) (sy, ly alike, indices mod
// N) and Phi(x) = int phi(2s/NW) cos(2 pi s x / N) ds over |s| <= NW/2, the
// kernel's Fourier transform (Gauss-Legendre quadrature, on the host).  The
// result differs from the exact sums by about 1e-14 of max |pf|.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "clm_ctx.h"

namespace pft {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int MB = 128;  // polar points per workgroup (4 waves x 32)
constexpr int KB = 64;   // images per workgroup (4 column tiles of 16)
constexpr int KCHUNK = 8192;  // images per staged chunk

struct Geo {
    int n, n_r, n_theta, M, Mp;  // M = n_r * n_theta/2 polar points, Mp = roundup(M, MB)
    int nup;                     // roundup(n, 4): pixel rows per image column in the staged stack
    int Kc, Kp;                  // images in this chunk, roundup(Kc, KB)
};

// ax[u*Mp + m] = exp(+1i*fx(m)*g(u)) (u < 4; the first four rows),
// ay[v*Mp + m] = exp(+1i*fy(m)*g(v)) (v < n), and
// ax[4*Mp + m] = exp(+4i*fx(m)), the step from pixel u to u + 4 (zero for
// m >= M).  grid: (Mp/256, 5 + n).
__global__ void tables_kernel(Geo g, double2* ax, double2* ay) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (m >= g.Mp) return;
    const bool is_x = t < 5;
    const int idx = is_x ? t : t - 5;
    double2 out = make_double2(0.0, 0.0);
    if (m < g.M) {
        const int k = m % g.n_r, j = m / g.n_r;
        const double omega0 = 2.0 * M_PI / (2.0 * g.n_r - 1.0);  // cryo_pft.m:49
        const double dtheta = 2.0 * M_PI / g.n_theta;             // cryo_pft.m:50
        const double f = is_x ? (k * omega0) * sin(j * dtheta) : (k * omega0) * cos(j * dtheta);
        const double gt = idx == 4 && is_x ? 4.0 : ceil(-0.5 * g.n + idx);  // nudft2.m:40
        double s, c;
        sincos(f * gt, &s, &c);
        out = make_double2(c, s);
    }
    (is_x ? ax : ay)[(size_t)idx * g.Mp + m] = out;
}

// Staged stack: xt[(v*nup + u)*Kp + i] = p(u, v, k0 + i), zero-padded in u and
// i.  One workgroup moves 64 pixel rows x 64 images through LDS so that both
// the reads (along u) and the writes (along i) are contiguous.
template <typename T>
__global__ __launch_bounds__(256) void stage_kernel(Geo g, const T* __restrict__ img, long long k0, int K,
                                                    double* __restrict__ xt) {
    __shared__ double tile[64][65];
    const int q0 = blockIdx.x * 64;  // staged pixel row
    const int i0 = blockIdx.y * 64;  // image within the chunk
    const int npix = g.n * g.nup;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int i = i0 + r, q = q0 + tx;
        double val = 0.0;
        if (q < npix && i < g.Kc && k0 + i < K) {
            const int v = q / g.nup, u = q - v * g.nup;
            if (u < g.n) val = (double)img[(size_t)(k0 + i) * g.n * g.n + (size_t)v * g.n + u];
        }
        tile[r][tx] = val;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int q = q0 + r, i = i0 + tx;
        if (q < npix && i < g.Kp) xt[(size_t)q * g.Kp + i] = tile[tx][r];
    }
}

template <typename OutT>
__device__ inline void put(OutT* pf, size_t at, double re, double im);
template <>
__device__ inline void put<double2>(double2* pf, size_t at, double re, double im) {
    pf[at] = make_double2(re, im);
}
template <>
__device__ inline void put<float2>(float2* pf, size_t at, double re, double im) {
    pf[at] = make_float2((float)re, (float)im);
}

// One workgroup: MB polar points x KB images (tile t of a 1-D grid; the tiles
// a workgroup's XCD runs are contiguous, polar-point blocks fastest, so they
// stream the same images through that XCD's L2); wave w takes points
// [32w, 32w + 32) of the block as two 16-row tiles, each split into its real
// and imaginary weight rows, against 4 column tiles of 16 images: 16
// accumulator tiles (16 x 16 FP64 each).  Each step contracts 4 pixels (one
// MFMA k-block): lane l holds pixel u0 + l/16 for point (l % 16) of a row
// tile (A) and image (l % 16) of a column tile (B).  Along u the lane's
// exponential factor advances by the rotation exp(+4i*fx(m)) (one complex
// product); it restarts from the table at every v.  Staged pixels u >= n
// are zero, so the factor's value there does not matter.
template <typename OutT>
__global__ __launch_bounds__(256, 2) void pft_kernel(Geo g, const double* __restrict__ xt,
                                                  const double2* __restrict__ ax, const double2* __restrict__ ay,
                                                  OutT* __restrict__ pf, long long k0, int nmb, int ntiles,
                                                  int per_xcd) {
    const int b = blockIdx.x;
    const int t = (b % 8) * per_xcd + b / 8;  // dispatch goes round-robin over the 8 XCDs
    if (t >= ntiles) return;
    const int mb = t % nmb, kb = t / nmb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int mA = mb * MB + wave * 32 + r16, mB = mA + 16;
    const int ib = kb * KB;
    d4 acc[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[a][b][c] = d4{0.0, 0.0, 0.0, 0.0};
    const double2 rA = ax[(size_t)4 * g.Mp + mA], rB = ax[(size_t)4 * g.Mp + mB];
    const double2 x0A = ax[(size_t)kq * g.Mp + mA], x0B = ax[(size_t)kq * g.Mp + mB];
    auto cmul = [](double2 p, double2 q) { return make_double2(p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x); };
    for (int v = 0; v < g.n; v++) {
        const double2 yA = ay[(size_t)v * g.Mp + mA], yB = ay[(size_t)v * g.Mp + mB];
        double2 xA = x0A, xB = x0B;
        const double* xr = xt + ((size_t)v * g.nup + kq) * g.Kp + ib + r16;
        for (int u0 = 0; u0 < g.nup; u0 += 4) {
            double bx[4];
#pragma unroll
            for (int c = 0; c < 4; c++) bx[c] = xr[(size_t)u0 * g.Kp + 16 * c];
            const double2 eA = cmul(xA, yA), eB = cmul(xB, yB);  // E = ax * ay
#pragma unroll
            for (int c = 0; c < 4; c++) {
                acc[0][0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(eA.x, bx[c], acc[0][0][c], 0, 0, 0);
                acc[0][1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(eA.y, bx[c], acc[0][1][c], 0, 0, 0);
                acc[1][0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(eB.x, bx[c], acc[1][0][c], 0, 0, 0);
                acc[1][1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(eB.y, bx[c], acc[1][1][c], 0, 0, 0);
            }
            xA = cmul(xA, rA);
            xB = cmul(xB, rB);
        }
    }
    // D tile layout of the FP64 16x16x4 MFMA: lane l, element t holds row
    // 4*t + l/16, column l % 16 (one row per element, unlike the FP32 tiles).
    const int L = g.n_theta / 2;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int m = mb * MB + wave * 32 + a * 16 + 4 * e + kq;
            if (m >= g.M) continue;
            const int k = m % g.n_r, j = m / g.n_r;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int i = ib + 16 * c + r16;
                if (i >= g.Kc) continue;
                const size_t img = (size_t)(k0 + i) * g.n_theta;
                const double re = acc[a][0][c][e], im = acc[a][1][c][e];
                put<OutT>(pf, (img + j) * g.n_r + k, re, im);
                put<OutT>(pf, (img + j + L) * g.n_r + k, re, -im);
            }
        }
}

// ------------------------------------------------------------ type-2 NUFFT
constexpr int NW = 16;                  // kernel width, grid points per axis
constexpr double NBETA = 2.30 * NW;     // kernel shape for oversampling 2
constexpr int NCHUNK = 2048;            // images per batch of FFTs (at most)
constexpr int NIPW = 4;                 // images per interpolation workgroup

struct NGeo {
    int n, n_r, n_theta, M;  // image side, rays, polar points M = n_r * n_theta/2
    int N;                   // oversampled grid side
    int x0;                  // g(0) = ceil(-n/2): pixel u sits at x = x0 + u
    int Kc;                  // images in this chunk
};

__device__ inline double es_kernel(double z) {
    return fabs(z) <= 1.0 ? exp(NBETA * (sqrt(1.0 - z * z) - 1.0)) : 0.0;
}

// Per polar point m: the first grid index of its window on each axis and
// the NW kernel weights along x, then along y.  Frequencies as in
// tables_kernel (cryo_pft.m:49-58).
__global__ void nweights_kernel(NGeo g, int2* base, double* wt) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.M) return;
    const int k = m % g.n_r, j = m / g.n_r;
    const double omega0 = 2.0 * M_PI / (2.0 * g.n_r - 1.0);
    const double dtheta = 2.0 * M_PI / g.n_theta;
    const double fx = (k * omega0) * sin(j * dtheta), fy = (k * omega0) * cos(j * dtheta);
    const double sx = fx * (g.N / (2.0 * M_PI)), sy = fy * (g.N / (2.0 * M_PI));
    const int lx = (int)ceil(sx - 0.5 * NW), ly = (int)ceil(sy - 0.5 * NW);
    base[m] = make_int2(lx, ly);
    double* w = wt + (size_t)m * 2 * NW;
    for (int t = 0; t < NW; t++) {
        w[t] = es_kernel(2.0 * (sx - (lx + t)) / NW);
        w[NW + t] = es_kernel(2.0 * (sy - (ly + t)) / NW);
    }
}

// Grid of the chunk's images, a[i][y mod N][x mod N] (x contiguous):
// p(u, v, k0 + i) / (dec[u] dec[v]) where a pixel sits, zero elsewhere.
template <typename T>
__global__ void deconv_kernel(NGeo g, const T* __restrict__ img, long long k0, const double* __restrict__ dec,
                              double2* __restrict__ a) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long NN = (long long)g.N * g.N;
    if (q >= NN * g.Kc) return;
    const int i = (int)(q / NN);
    const int r = (int)(q - (long long)i * NN);
    const int ly = r / g.N, lx = r - ly * g.N;
    const int xmax = g.x0 + g.n - 1;  // pixel offsets x0..xmax; N >= 2n keeps x and x - N apart
    const int x = lx <= xmax ? lx : lx - g.N, y = ly <= xmax ? ly : ly - g.N;
    double val = 0.0;
    if (x >= g.x0 && y >= g.x0) {
        const int u = x - g.x0, v = y - g.x0;
        val = (double)img[((size_t)(k0 + i) * g.n + v) * g.n + u] / (dec[u] * dec[v]);
    }
    a[q] = make_double2(val, 0.0);
}

// Sum over the 16 lanes of each row (quad_perm [1,0,3,2], [2,3,0,1],
// row_ror 4, row_ror 8): every lane of the row gets it.  All lanes active.
template <int CTRL>
__device__ inline double dpp64(double x) {
    const unsigned long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffu), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double row16_sum(double x) {
    x += dpp64<0xb1>(x);
    x += dpp64<0x4e>(x);
    x += dpp64<0x124>(x);
    x += dpp64<0x128>(x);
    return x;
}

__device__ inline void put_pf(double2* o, double re, double im) { *o = make_double2(re, im); }
__device__ inline void put_pf(float2* o, double re, double im) { *o = make_float2((float)re, (float)im); }

// pf of the chunk's images at the polar points.  Workgroup: 16 points (4
// per wave, 16 lanes each; lane t = x offset of the window) x NIPW images;
// each lane walks the NW window rows in order, then the point's 16 lanes are
// summed.  pf(k, j) and its conjugate pf(k, j + n_theta/2) (cryo_pft.m:75).
template <typename OutT>
__global__ __launch_bounds__(256) void interp_kernel(NGeo g, const double2* __restrict__ b,
                                                     const int2* __restrict__ base, const double* __restrict__ wt,
                                                     OutT* __restrict__ pf, long long k0) {
    __shared__ double s_wy[16][NW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane & 15, ps = wave * 4 + (lane >> 4);
    const int m = blockIdx.x * 16 + ps;
    const bool act = m < g.M;
    int cx = 0, ry0 = 0;
    double wx = 0.0;
    if (act) {
        const int2 bs = base[m];
        wx = wt[(size_t)m * 2 * NW + t];
        s_wy[ps][t] = wt[(size_t)m * 2 * NW + NW + t];
        cx = (bs.x + t) % g.N;
        if (cx < 0) cx += g.N;
        ry0 = bs.y % g.N;
        if (ry0 < 0) ry0 += g.N;
    }
    __syncthreads();
    const int L = g.n_theta / 2;
    const int k = act ? m % g.n_r : 0, j = act ? m / g.n_r : 0;
    const long long NN = (long long)g.N * g.N;
    for (int ii = 0; ii < NIPW; ii++) {
        const int i = blockIdx.y * NIPW + ii;
        if (i >= g.Kc) break;  // the same for the whole workgroup
        const double2* bi = b + (size_t)i * NN;
        double re = 0.0, im = 0.0;
        if (act) {
            int ry = ry0;
#pragma unroll 4
            for (int r = 0; r < NW; r++) {
                const double w = s_wy[ps][r] * wx;
                const double2 v = bi[(size_t)ry * g.N + cx];
                re += w * v.x;
                im += w * v.y;
                ry = ry + 1 == g.N ? 0 : ry + 1;
            }
        }
        re = row16_sum(re);
        im = row16_sum(im);
        if (act && t == 0) {
            const size_t col = (size_t)(k0 + i) * g.n_theta;
            put_pf(pf + (col + j) * g.n_r + k, re, im);
            put_pf(pf + (col + j + L) * g.n_r + k, re, -im);
        }
    }
}

}  // namespace pft

using namespace clmh;

namespace {

int pft_check(clm_ctx* c, const clm_pft_params* p) {
    if (!p) return fail(c, CLM_E_ARG, "null params");
    if (p->n_theta <= 0 || p->n_theta % 2) return fail(c, CLM_E_ARG, "`n_theta` must be even.");
    if (p->n < 1 || p->n > 4096) return fail(c, CLM_E_UNSUP, "image side n = %d out of range [1, 4096]", p->n);
    if (p->n_r < 1 || p->K < 0) return fail(c, CLM_E_ARG, "invalid n_r = %d or K = %d", p->n_r, p->K);
    if ((long long)p->n_r * (p->n_theta / 2) > (1LL << 26))
        return fail(c, CLM_E_UNSUP, "n_r * n_theta/2 = %lld polar points exceeds 2^26",
                    (long long)p->n_r * (p->n_theta / 2));
    if (p->img_dtype != CLM_R64 && p->img_dtype != CLM_R32)
        return fail(c, CLM_E_ARG, "unknown image dtype %d", p->img_dtype);
    if (p->pf_dtype != CLM_C128 && p->pf_dtype != CLM_C64) return fail(c, CLM_E_ARG, "unknown pf dtype %d", p->pf_dtype);
    if (p->method != CLM_PFT_EXACT && p->method != CLM_PFT_NUFFT)
        return fail(c, CLM_E_ARG, "unknown pft method %d", p->method);
    if (p->method == CLM_PFT_NUFFT && p->n > 2048)
        return fail(c, CLM_E_UNSUP, "image side n = %d: the NUFFT method takes n <= 2048", p->n);
    return CLM_OK;
}

int pft_run(clm_ctx* c, const clm_pft_params* p, const void* d_img, void* d_pf, clm_pft_stats* st) {
    using namespace pft;
    Geo g{};
    g.n = p->n;
    g.n_r = p->n_r;
    g.n_theta = p->n_theta;
    g.M = p->n_r * (p->n_theta / 2);
    g.Mp = (g.M + MB - 1) / MB * MB;
    g.nup = (p->n + 3) / 4 * 4;
    const int K = p->K;
    const int kc = K < KCHUNK ? K : KCHUNK;
    const int kcp = (kc + KB - 1) / KB * KB;
    int rc;
    if ((rc = ensure(c, c->pft_ax, (size_t)5 * g.Mp * sizeof(double2))) ||
        (rc = ensure(c, c->pft_ay, (size_t)g.n * g.Mp * sizeof(double2))) ||
        (rc = ensure(c, c->pft_xt, (size_t)g.n * g.nup * (size_t)kcp * sizeof(double))))
        return rc;
    hipEvent_t* ev = c->ev;
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    tables_kernel<<<dim3(g.Mp / 256 + (g.Mp % 256 != 0), 5 + g.n), 256, 0, c->stream>>>(
        g, (double2*)c->pft_ax.p, (double2*)c->pft_ay.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    float ms_stage = 0.f, ms_pft = 0.f;
    int chunks = 0;
    for (long long k0 = 0; k0 < K; k0 += KCHUNK) {
        g.Kc = (int)((K - k0) < KCHUNK ? (K - k0) : KCHUNK);
        g.Kp = (g.Kc + KB - 1) / KB * KB;
        const int npix = g.n * g.nup;
        dim3 sgrid((npix + 63) / 64, g.Kp / 64);
        HIPCHK(c, hipEventRecord(ev[2], c->stream));
        if (p->img_dtype == CLM_R64)
            stage_kernel<double><<<sgrid, 256, 0, c->stream>>>(g, (const double*)d_img, k0, K, (double*)c->pft_xt.p);
        else
            stage_kernel<float><<<sgrid, 256, 0, c->stream>>>(g, (const float*)d_img, k0, K, (double*)c->pft_xt.p);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(ev[3], c->stream));
        const int nmb = g.Mp / MB, ntiles = nmb * (g.Kp / KB);
        const int per = (ntiles + 7) / 8;
        const dim3 grid(8 * per);
        if (p->pf_dtype == CLM_C128)
            pft_kernel<double2><<<grid, 256, 0, c->stream>>>(g, (const double*)c->pft_xt.p, (const double2*)c->pft_ax.p,
                                                               (const double2*)c->pft_ay.p, (double2*)d_pf, k0, nmb,
                                                               ntiles, per);
        else
            pft_kernel<float2><<<grid, 256, 0, c->stream>>>(g, (const double*)c->pft_xt.p, (const double2*)c->pft_ax.p,
                                                              (const double2*)c->pft_ay.p, (float2*)d_pf, k0, nmb,
                                                              ntiles, per);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(ev[4], c->stream));
        HIPCHK(c, hipEventSynchronize(ev[4]));
        float a = 0.f, b = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, ev[2], ev[3]));
        HIPCHK(c, hipEventElapsedTime(&b, ev[3], ev[4]));
        ms_stage += a;
        ms_pft += b;
        chunks++;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st) {
        float t = 0.f;
        HIPCHK(c, hipEventElapsedTime(&t, ev[0], ev[1]));
        st->ms_tables = t;
        st->ms_stage = ms_stage;
        st->ms_pft = ms_pft;
        st->chunks = chunks;
    }
    return CLM_OK;
}

// ------------------------------------------------------------ type-2 NUFFT
// Gauss-Legendre rule of q nodes on [-1, 1]: roots of P_q by Newton's method
// (P_q from the three-term recurrence), weights 2 / ((1 - z^2) P_q'(z)^2).
void gauss_legendre(int q, std::vector<double>& x, std::vector<double>& w) {
    x.assign(q, 0.0);
    w.assign(q, 0.0);
    for (int i = 0; i < (q + 1) / 2; i++) {
        double z = std::cos(M_PI * (i + 0.75) / (q + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p0 = 0.0;
            for (int jj = 1; jj <= q; jj++) {
                const double pm = p0;
                p0 = p1;
                p1 = ((2.0 * jj - 1.0) * z * p0 - (jj - 1.0) * pm) / jj;
            }
            dp = q * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        x[q - 1 - i] = z;
        w[i] = w[q - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// Smallest 2^a 3^b 5^c >= n.
int smooth_at_least(int n) {
    for (int m = n;; m++) {
        int r = m;
        for (int f : {2, 3, 5})
            while (r % f == 0) r /= f;
        if (r == 1) return m;
    }
}

hipfftHandle fft_plan(clm_ctx* c, int N, int batch, int* rc) {
    for (int s = 0; s < 2; s++)
        if (c->fft_plan[s] && c->fft_side[s] == N && c->fft_batch[s] == batch) return (hipfftHandle)c->fft_plan[s];
    // replace the older of the two (slot 1), keeping slot 0 the most recent
    if (c->fft_plan[1]) (void)hipfftDestroy((hipfftHandle)c->fft_plan[1]);
    c->fft_plan[1] = c->fft_plan[0];
    c->fft_side[1] = c->fft_side[0];
    c->fft_batch[1] = c->fft_batch[0];
    c->fft_plan[0] = nullptr;
    hipfftHandle h = nullptr;
    int dims[2] = {N, N};
    hipfftResult r = hipfftPlanMany(&h, 2, dims, nullptr, 1, N * N, nullptr, 1, N * N, HIPFFT_Z2Z, batch);
    if (r == HIPFFT_SUCCESS) r = hipfftSetStream(h, c->stream);
    if (r != HIPFFT_SUCCESS) {
        if (h) (void)hipfftDestroy(h);
        *rc = fail(c, CLM_E_HIP, "hipFFT plan of %d x %d x %d failed (hipfftResult %d)", N, N, batch, (int)r);
        return nullptr;
    }
    c->fft_plan[0] = (void*)h;
    c->fft_side[0] = N;
    c->fft_batch[0] = batch;
    return h;
}

int nufft_run(clm_ctx* c, const clm_pft_params* p, const void* d_img, void* d_pf, clm_pft_stats* st) {
    using namespace pft;
    NGeo g{};
    g.n = p->n;
    g.n_r = p->n_r;
    g.n_theta = p->n_theta;
    g.M = p->n_r * (p->n_theta / 2);
    g.N = smooth_at_least(std::max(2 * p->n, 2 * NW));
    g.x0 = (int)std::ceil(-0.5 * p->n);
    const int K = p->K;
    const size_t NN = (size_t)g.N * g.N;
    // images per batch: at most NCHUNK, and grids of at most 4 GiB
    const int kc = (int)std::max<long long>(1, std::min<long long>({(long long)K, (long long)NCHUNK,
                                                                     (4LL << 30) / (long long)(NN * sizeof(double2))}));
    int rc;
    if ((rc = ensure(c, c->pft_grid, (size_t)kc * NN * sizeof(double2))) ||
        (rc = ensure(c, c->pft_base, (size_t)g.M * sizeof(int2))) ||
        (rc = ensure(c, c->pft_wt, (size_t)g.M * 2 * NW * sizeof(double))) ||
        (rc = ensure(c, c->pft_dec, (size_t)g.n * sizeof(double))))
        return rc;
    // deconvolution factors Phi(x0 + u), u < n
    std::vector<double> gx, gw, dec(g.n);
    gauss_legendre(200, gx, gw);
    for (int u = 0; u < g.n; u++) {
        const double x = g.x0 + u;
        double acc = 0.0;
        for (size_t q = 0; q < gx.size(); q++)
            acc += gw[q] * std::exp(NBETA * (std::sqrt(1.0 - gx[q] * gx[q]) - 1.0)) *
                   std::cos(M_PI * gx[q] * NW * x / g.N);
        dec[u] = 0.5 * NW * acc;
    }
    hipEvent_t* ev = c->ev;
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    HIPCHK(c, hipMemcpyAsync(c->pft_dec.p, dec.data(), dec.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    nweights_kernel<<<(g.M + 255) / 256, 256, 0, c->stream>>>(g, (int2*)c->pft_base.p, (double*)c->pft_wt.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    float ms_stage = 0.f, ms_fft = 0.f, ms_pft = 0.f;
    int chunks = 0;
    double2* grid = (double2*)c->pft_grid.p;
    for (long long k0 = 0; k0 < K; k0 += kc) {
        g.Kc = (int)((K - k0) < kc ? (K - k0) : kc);
        hipfftHandle plan = fft_plan(c, g.N, g.Kc, &rc);
        if (!plan) return rc;
        HIPCHK(c, hipEventRecord(ev[2], c->stream));
        const long long nq = (long long)NN * g.Kc;
        if (p->img_dtype == CLM_R64)
            deconv_kernel<double><<<(unsigned)((nq + 255) / 256), 256, 0, c->stream>>>(
                g, (const double*)d_img, k0, (const double*)c->pft_dec.p, grid);
        else
            deconv_kernel<float><<<(unsigned)((nq + 255) / 256), 256, 0, c->stream>>>(
                g, (const float*)d_img, k0, (const double*)c->pft_dec.p, grid);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(ev[3], c->stream));
        const hipfftResult fr = hipfftExecZ2Z(plan, (hipfftDoubleComplex*)grid, (hipfftDoubleComplex*)grid,
                                              HIPFFT_BACKWARD);
        if (fr != HIPFFT_SUCCESS) return fail(c, CLM_E_HIP, "hipfftExecZ2Z failed (hipfftResult %d)", (int)fr);
        HIPCHK(c, hipEventRecord(ev[4], c->stream));
        const dim3 igrid((g.M + 15) / 16, (g.Kc + NIPW - 1) / NIPW);
        if (p->pf_dtype == CLM_C128)
            interp_kernel<double2><<<igrid, 256, 0, c->stream>>>(g, grid, (const int2*)c->pft_base.p,
                                                                 (const double*)c->pft_wt.p, (double2*)d_pf, k0);
        else
            interp_kernel<float2><<<igrid, 256, 0, c->stream>>>(g, grid, (const int2*)c->pft_base.p,
                                                                (const double*)c->pft_wt.p, (float2*)d_pf, k0);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(ev[5], c->stream));
        HIPCHK(c, hipEventSynchronize(ev[5]));
        float a = 0.f, b = 0.f, d = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, ev[2], ev[3]));
        HIPCHK(c, hipEventElapsedTime(&b, ev[3], ev[4]));
        HIPCHK(c, hipEventElapsedTime(&d, ev[4], ev[5]));
        ms_stage += a;
        ms_fft += b;
        ms_pft += d;
        chunks++;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));  // dec is released on return
    if (st) {
        float t = 0.f;
        HIPCHK(c, hipEventElapsedTime(&t, ev[0], ev[1]));
        st->ms_tables = t;
        st->ms_stage = ms_stage;
        st->ms_fft = ms_fft;
        st->ms_pft = ms_pft;
        st->chunks = chunks;
        st->grid_side = g.N;
    }
    return CLM_OK;
}

int pft_dispatch(clm_ctx* c, const clm_pft_params* p, const void* d_img, void* d_pf, clm_pft_stats* st) {
    return p->method == CLM_PFT_NUFFT ? nufft_run(c, p, d_img, d_pf, st) : pft_run(c, p, d_img, d_pf, st);
}

}  // namespace

namespace clmh {
void pft_release(clm_ctx* c) {
    for (int s = 0; s < 2; s++) {
        if (c->fft_plan[s]) (void)hipfftDestroy((hipfftHandle)c->fft_plan[s]);
        c->fft_plan[s] = nullptr;
        c->fft_side[s] = c->fft_batch[s] = 0;
    }
}
}  // namespace clmh

extern "C" {

int clm_pft_freqs(int32_t n_r, int32_t n_theta, double* freqs) {
    if (n_r < 1 || n_theta <= 0 || n_theta % 2 || !freqs) return CLM_E_ARG;
    const int L = n_theta / 2;
    const size_t M = (size_t)n_r * L;
    const double omega0 = 2.0 * M_PI / (2.0 * n_r - 1.0);
    const double dtheta = 2.0 * M_PI / n_theta;
    for (int j = 0; j < L; j++)
        for (int k = 0; k < n_r; k++) {
            const size_t m = (size_t)j * n_r + k;
            freqs[m] = (k * omega0) * sin(j * dtheta);
            freqs[M + m] = (k * omega0) * cos(j * dtheta);
        }
    return CLM_OK;
}

int clm_pft_device(clm_ctx* c, const clm_pft_params* p, const void* d_img, void* d_pf, clm_pft_stats* st) {
    if (!c) return CLM_E_ARG;
    int rc = pft_check(c, p);
    if (rc) return rc;
    if (st) *st = clm_pft_stats{};
    if (p->K == 0) return CLM_OK;
    if (!d_img || !d_pf) return fail(c, CLM_E_ARG, "null image or pf pointer");
    HIPCHK(c, hipSetDevice(c->device));
    return pft_dispatch(c, p, d_img, d_pf, st);
}

int clm_pft(clm_ctx* c, const clm_pft_params* p, const void* img, void* pf, clm_pft_stats* st) {
    if (!c) return CLM_E_ARG;
    int rc = pft_check(c, p);
    if (rc) return rc;
    if (st) *st = clm_pft_stats{};
    if (p->K == 0) return CLM_OK;
    if (!img || !pf) return fail(c, CLM_E_ARG, "null image or pf pointer");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nimg = (size_t)p->n * p->n * p->K * (p->img_dtype == CLM_R64 ? 8 : 4);
    const size_t npf = (size_t)p->n_r * p->n_theta * p->K * (p->pf_dtype == CLM_C128 ? 16 : 8);
    if ((rc = ensure(c, c->pft_img, nimg)) || (rc = ensure(c, c->pft_out, npf))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->pft_img.p, img, nimg, hipMemcpyHostToDevice, c->stream));
    if ((rc = pft_dispatch(c, p, c->pft_img.p, c->pft_out.p, st))) return rc;
    HIPCHK(c, hipMemcpyAsync(pf, c->pft_out.p, npf, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

}  // extern "C"
