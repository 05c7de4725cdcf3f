// shifts.hip -- 2-D shift least squares after the common-lines search, on
// the device: the step after cryo_clmatrix.
//
// Replaces  est_shifts = lsqr(A, b, tol, maxit)  of
// sinograms/cryo_estimate_shifts.m:229-230 applied to the shift equations
// that cryo_clmatrix returns (A = shift_equations(:,1:end-1), b = its last
// column, rows from cryo_clmatrix_cpu.m:722-741).  A is never formed: every
// product reads the per-pair common lines and evaluates the row
//   alpha = (cl1-1)*pi/L, beta = (cl2-1)*pi/L
//   beta <  pi: [ sin(alpha),  cos(alpha), -sin(beta),      -cos(beta)     ]
//   beta >= pi: [-sin(alpha), -cos(alpha), -sin(beta - pi), -cos(beta - pi)]
// at columns [2k1, 2k1+1, 2k2, 2k2+1] from two small tables indexed by the
// line number.  LSQR is the published algorithm (Paige and Saunders, ACM
// TOMS 8(1), 1982) with its damping term, from x0 = 0, with MATLAB's two
// documented stopping tests; its scalar recurrences run on the device, so a
// solve is a fixed sequence of launches with no host round trip (after the
// tests pass, the remaining launches return at once).
//
// One pass over the pairs per iteration: u' = A v - (alpha/beta) u is formed
// and, in the same pass, A'u' (A' is linear, so the normalization by the new
// beta = |u'| is applied afterwards to the 2K-vector).  Memory per iteration
// (K = 5000, 12.5M pairs): the packed lines (4 B/pair) and u (8 B/pair) are
// read once and u written once, 20 B/pair, about 0.25 GB (stored tile by
// tile, with the empty slots of the diagonal tiles), plus the 2K-long
// partial sums of the 64 x 256-pair tiles; HBM-bound work with every
// reduction in a fixed order (results do not depend on scheduling).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "clm_ctx.h"

namespace lsq {

struct State {
    double alpha, beta, bnorm, phibar, rhobar, anorm, res2, rnorm, arnorm;
    double damp, tol;
    double t1, t2;      // the last scalar step's update of x and w
    double vscale;      // the current v is the stored v times vscale
    int itn, maxit, done, flag;
    int upd;            // 1: the last scalar step's update of x and w is pending
    int vset;           // 1: the first direction v was formed (b != 0)
};

__device__ inline long long first(long long k1, long long K) { return k1 * K - k1 * (k1 + 1) / 2; }


// ta[c] = (sin a, cos a), a = (c-1)*pi/L; tb[c] = (-sin b', -cos b', sign):
// b = (c-1)*pi/L, b' = b or b - pi, sign = +1 (b < pi) or -1 (b >= pi), for
// line numbers c = 0..2L (0 = no line; cryo_clmatrix_cpu.m:722-741).  Entry
// 2L+1 is all zeros: the row of a pair whose line numbers lie outside
// [0, 2L].
__global__ void tables_kernel(int L, int by_index, double2* ta, double* tb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > 2 * L + 1) return;
    if (c == 2 * L + 1) {
        ta[c] = make_double2(0.0, 0.0);
        tb[3 * c + 0] = tb[3 * c + 1] = tb[3 * c + 2] = 0.0;
        return;
    }
    const double dtheta = M_PI / (double)L;
    const double a = (double)(c - 1) * dtheta;
    double b = (double)(c - 1) * dtheta;
    ta[c] = make_double2(sin(a), cos(a));
    double sg = 1.0;
    // the second line's orientation: the angle test of cryo_clmatrix_cpu.m:733
    // (beta < pi), or with by_index the line index of
    // cryo_estimate_shifts.m:171-176,215-222 (c > L); they differ only at c = L + 1
    // for the L where L * RN(pi/L) < RN(pi)
    if (by_index ? c > L : !(b < M_PI)) {
        b = b - M_PI;
        sg = -1.0;
    }
    tb[3 * c + 0] = -sin(b);
    tb[3 * c + 1] = -cos(b);
    tb[3 * c + 2] = sg;
}

// Block reduction of one double per thread (fixed order); result in thread 0.
template <int NT>
__device__ inline double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Sum of p[0..n) in one fixed order whatever the block size NT >= 256:
// thread t < 256 adds p[t], p[t + 256], ... in order, then a 256-wide tree.
// Every block that calls it gets the same bits.
template <int NT>
__device__ inline double fixed_sum(const double* p, int n, double* sh) {
    const int t = threadIdx.x;
    double s = 0.0;
    if (t < 256)
        for (int i = t; i < n; i += 256) s += p[i];
    if (t < 256) sh[t] = s;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Pair tiles: rows k1 in [TR*kb, TR*kb + TR), columns k2 in [TC*cb, TC*cb + TC);
// a tile exists when it holds a pair k1 < k2 < K.
constexpr int TR = 64, TC = 256;

__host__ __device__ inline bool tile_exists(int K, int kb, int cb) {
    const int kmax = min(TC * cb + TC, K) - 1;  // largest k2 of the column block
    return TR * kb < kmax;
}

// x from the lane that DPP control CTRL names (both 32-bit halves moved
// alike); every lane of the wave must be active.
template <int CTRL>
__device__ inline double dpp_d(double x) {
    const unsigned long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffu), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// dpp_d without the old value: for the in-row patterns below every lane has
// a source lane, so nothing of the old value survives
template <int CTRL>
__device__ inline double mov_dpp_d(double x) {
    const unsigned long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffu), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double readlane_d(double x, int l) {
    const unsigned long long u = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(u & 0xffffffffu), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes in a fixed order, the same in every lane: DPP adds
// within each row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], row_ror 4,
// row_ror 8), then the four row sums in row order.  Every lane of the wave
// must be active.
__device__ inline double wave_sum(double x) {
    x += dpp_d<0xb1>(x);
    x += dpp_d<0x4e>(x);
    x += dpp_d<0x124>(x);
    x += dpp_d<0x128>(x);
    return ((readlane_d(x, 0) + readlane_d(x, 16)) + readlane_d(x, 32)) + readlane_d(x, 48);
}
// wave_sum of two values at once (the same order and bits as two calls),
// the steps interleaved so each fills the other's DPP wait
__device__ inline void wave_sum2(double& x, double& y) {
    x += mov_dpp_d<0xb1>(x);
    y += mov_dpp_d<0xb1>(y);
    x += mov_dpp_d<0x4e>(x);
    y += mov_dpp_d<0x4e>(y);
    x += mov_dpp_d<0x124>(x);
    y += mov_dpp_d<0x124>(y);
    x += mov_dpp_d<0x128>(x);
    y += mov_dpp_d<0x128>(y);
    x = ((readlane_d(x, 0) + readlane_d(x, 16)) + readlane_d(x, 32)) + readlane_d(x, 48);
    y = ((readlane_d(y, 0) + readlane_d(y, 16)) + readlane_d(y, 32)) + readlane_d(y, 48);
}
// Sum of p[0..n) in one fixed order by one whole wave: lane l adds p[l],
// p[l + 64], ... in order, then wave_sum.  Every caller gets the same bits.
__device__ inline double wave_fixed_sum(const double* p, int n) {
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    int i = lane;
    // eight loads in flight, then their adds in order
    for (; i + 7 * 64 < n; i += 8 * 64) {
        double a[8];
#pragma unroll
        for (int q = 0; q < 8; q++) a[q] = p[i + q * 64];
#pragma unroll
        for (int q = 0; q < 8; q++) s += a[q];
    }
    for (; i < n; i += 64) s += p[i];
    return wave_sum(s);
}

// One pass over tile t's pairs: pair (k1, k2) at row r = k1 - TR*kb and
// column c = k2 - TC*cb of the tile is equation e = first(k1) + k2 - k1 - 1
// and slot q = (t*TR + r)*TC + c of the tile-major pack and u:
//   INIT: pack[q] = cl1[e] | cl2[e] << 16 (both 2L+1 when either is outside
//         [0, 2L]) and u[q] = b[e];
//   else: u[q] <- (A v)[e] - (alpha/beta) u[q]  (u holds the previous
//         unnormalized vector, whose norm is beta);
// and, of the new u: normp[t] = sum of squares; rowp[cb][2k1 + c] = the
// tile's row part of A'u (the k1 coefficients), colp[kb][2k2 + c] = its
// column part (the k2 coefficients).  A thread owns one column and walks the
// tile's rows in order; row sums go through fixed wave butterflies and then
// the four waves in order.
// LT: the two coefficient tables (5 (2L+2) doubles, ta then tb, contiguous)
// are first copied to LDS, so the loop's table reads wait on the LDS counter
// and not behind the next rows' HBM loads on the vector-memory counter.
template <bool INIT, bool LT>
__global__ __launch_bounds__(256) void pass_kernel(int K, int L, const int2* __restrict__ tiles,
                                                   const int* __restrict__ cl1, const int* __restrict__ cl2,
                                                   const double* __restrict__ b, unsigned* __restrict__ pack,
                                                   const double2* __restrict__ ta, const double* __restrict__ tb,
                                                   const double* __restrict__ v, const State* S, double* u,
                                                   double* normp, double* rowp, double* colp) {
    __shared__ double s_row[TR][4][2];
    __shared__ double s_sq[4];
    extern __shared__ double s_tab[];
    if (!INIT && S->done) return;
    // the tile's rows of v (2 TR entries, zero past the last image)
    __shared__ double s_v[2 * TR];
    if (!INIT && threadIdx.x < 2 * TR) {
        const int i = 2 * TR * tiles[blockIdx.x].x + threadIdx.x;
        s_v[threadIdx.x] = i < 2 * K ? v[i] : 0.0;
    }
    if constexpr (LT) {
        const double* src = (const double*)ta;
        for (int i = threadIdx.x; i < 5 * (2 * L + 2); i += blockDim.x) s_tab[i] = src[i];
    }
    if (LT || !INIT) __syncthreads();
    const int2 tl = tiles[blockIdx.x];
    const int kb = tl.x, cb = tl.y;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int k2 = TC * cb + t;
    const double ca = INIT ? 0.0 : S->alpha / S->beta;
    const double vs = INIT ? 0.0 : S->vscale;  // stored v times vscale is the current v
    double va0 = 0.0, va1 = 0.0;
    if (!INIT && k2 < K) {
        va0 = v[2 * k2] * vs;
        va1 = v[2 * k2 + 1] * vs;
    }
    double c0 = 0.0, c1 = 0.0, sq = 0.0;
    const unsigned z = (unsigned)(2 * L + 1);
    // rows in groups of UR, double-buffered in registers: the next group's
    // loads are issued before this group's stores (u is read and written in
    // place, so the compiler cannot move a later row's load above an
    // earlier row's store); the rows are finished in order
    constexpr int UR = 4;
    // pack and u are stored tile by tile (TR x TC slots per tile, row-major):
    // a workgroup streams one contiguous block instead of TR rows that lie
    // K entries apart in equation order
    const size_t tbase = (size_t)blockIdx.x * TR * TC + t;
    unsigned pk[2][UR];
    double ub[2][UR];
    // The first pass writes every slot of a tile: a pair's packed lines and
    // right-hand side, or the no-line code (zero coefficients) and u = 0
    // where the tile holds no pair.  The later passes then load, compute
    // and store every slot without a branch (a slot without a pair stays 0
    // and adds +0 to every sum), which lets the compiler count the loads in
    // flight and wait for exactly the row it finishes.
    auto load_group = [&](int r0, unsigned* gp, double* gu) {
#pragma unroll
        for (int i = 0; i < UR; i++) {
            if constexpr (INIT) {
                const int k1 = TR * kb + r0 + i;
                gp[i] = z | (z << 16);
                gu[i] = 0.0;
                if (k1 < K && k2 > k1 && k2 < K) {
                    const long long e = first(k1, K) + (k2 - k1 - 1);
                    const int a = cl1[e], c2 = cl2[e];
                    if (a >= 0 && a <= 2 * L && c2 >= 0 && c2 <= 2 * L) gp[i] = (unsigned)a | ((unsigned)c2 << 16);
                    gu[i] = b[e];
                }
            } else {
                const size_t q = tbase + (size_t)(r0 + i) * TC;
                gp[i] = pack[q];
                gu[i] = u[q];
            }
        }
    };
    auto finish_group = [&](int r0, const unsigned* gp, const double* gu) {
#pragma unroll
        for (int i = 0; i < UR; i++) {
            const int r = r0 + i;
            const size_t q = tbase + (size_t)r * TC;
            double2 A;
            double B0, B1, sg;
            if constexpr (LT) {
                const int ia = gp[i] & 0xffff, ib = 2 * (2 * L + 2) + 3 * (gp[i] >> 16);
                A = make_double2(s_tab[2 * ia], s_tab[2 * ia + 1]);
                B0 = s_tab[ib];
                B1 = s_tab[ib + 1];
                sg = s_tab[ib + 2];
            } else {
                A = ta[gp[i] & 0xffff];
                const double* B = tb + 3 * (gp[i] >> 16);
                B0 = B[0];
                B1 = B[1];
                sg = B[2];
            }
            double un;
            if constexpr (INIT) {
                pack[q] = gp[i];
                un = gu[i];
            } else {
                const double av = sg * A.x * (s_v[2 * r] * vs) + sg * A.y * (s_v[2 * r + 1] * vs) + B0 * va0 +
                                  B1 * va1;
                un = av - ca * gu[i];
            }
            u[q] = un;
            sq += un * un;
            double r0v = sg * A.x * un;
            double r1v = sg * A.y * un;
            c0 += B0 * un;
            c1 += B1 * un;
            wave_sum2(r0v, r1v);
            if (lane == 0) {
                s_row[r][wave][0] = r0v;
                s_row[r][wave][1] = r1v;
            }
        }
    };
    static_assert((TR / UR) % 2 == 0, "groups come in pairs");
    load_group(0, pk[0], ub[0]);
#pragma unroll 1
    for (int r0 = 0; r0 < TR; r0 += 2 * UR) {
        load_group(r0 + UR, pk[1], ub[1]);
        finish_group(r0, pk[0], ub[0]);
        if (r0 + 2 * UR < TR) load_group(r0 + 2 * UR, pk[0], ub[0]);
        finish_group(r0 + UR, pk[1], ub[1]);
    }
    sq = wave_sum(sq);
    if (lane == 0) s_sq[wave] = sq;
    if (k2 < K) {
        colp[(size_t)kb * 2 * K + 2 * k2] = c0;
        colp[(size_t)kb * 2 * K + 2 * k2 + 1] = c1;
    }
    __syncthreads();
    if (t < 2 * TR) {
        const int r = t >> 1, c = t & 1, k1 = TR * kb + r;
        if (k1 < K)
            rowp[(size_t)cb * 2 * K + 2 * k1 + c] = ((s_row[r][0][c] + s_row[r][1][c]) + s_row[r][2][c]) + s_row[r][3][c];
    }
    if (t == 0) normp[blockIdx.x] = ((s_sq[0] + s_sq[1]) + s_sq[2]) + s_sq[3];
}

// Before the first iteration: beta = bnorm = |b| from the first pass's
// tiles' sums of squares.
__global__ __launch_bounds__(1024) void beta_kernel(int n, const double* part, State* S) {
    __shared__ double sh[1024];
    if (S->done) return;
    const double beta = sqrt(fixed_sum<1024>(part, n, sh));
    if (threadIdx.x == 0) {
        S->beta = beta;
        S->bnorm = beta;
        if (beta == 0.0) {  // b = 0: x = 0 is the solution
            S->done = 1;
            S->flag = 0;
        }
    }
}

// Entries of the 2K-vector per workgroup of the element-parallel steps: 16
// entries x 16 partial groups (256 threads); the first 16 threads finish the
// entries.
constexpr int VE = 16;
constexpr int NG = 256 / VE;  // partial groups per entry

// Group g's share (g = 0..NG-1) of (A'u)[i], entry i of the 2K-vector: the
// row parts of the tiles in image i/2's row block, then the column parts of
// the tiles in its column block, each in block order over the g-th of NG
// slices of the blocks.  (A'u)[i] = (..((g0 + g1) + g2) ..) + g(NG-1).
__device__ inline double atu_part(int K, const double* rowp, const double* colp, int i, int g) {
    const int k = i >> 1, n2 = 2 * K;
    const int kb = k / TR, cbk = k / TC;
    const int ncb = (K + TC - 1) / TC, nkb = (K + TR - 1) / TR;
    // the partials of tiles that do not exist are zero (cleared before the
    // solve, never written), so every block is added: +0 changes no sum
    // that starts at +0, and the loads need no branch
    (void)kb;
    (void)cbk;
    double t = 0.0;
    const double* rp = rowp + (size_t)(g * ncb / NG) * n2 + i;
#pragma unroll 4
    for (int cb = g * ncb / NG; cb < (g + 1) * ncb / NG; cb++, rp += n2) t += *rp;
    const double* cp = colp + (size_t)(g * nkb / NG) * n2 + i;
#pragma unroll 4
    for (int b = g * nkb / NG; b < (g + 1) * nkb / NG; b++, cp += n2) t += *cp;
    return t;
}


// One LSQR iteration after the pass is two launches:
//  stepa_kernel (NG threads per element): the previous step's update of x
//    and w, v <- A'u / beta - beta v, per-workgroup partial norms;
//  stepb_kernel, one wave (scalar_step): the running
//    Frobenius estimate of A, alpha = |v|, the damping and plane rotations,
//    the stopping tests; the update's coefficients and the normalization
//    1/alpha of v go to the state, applied by the next pass and stepa (or
//    finish_kernel after the last iteration).
// beta = 0 (b in the range of A): v and alpha stay, the rotation makes the
// last update of x and phibar becomes 0.
// Run by one whole wave (lane 0 updates the state).
__device__ inline void scalar_step(double beta, int nb, const double* part, State* S) {
    double alpha = S->alpha;
    if (beta > 0.0) alpha = sqrt(wave_fixed_sum(part, nb));
    if (threadIdx.x == 0) {
        S->beta = beta;
        if (beta > 0.0) S->anorm = sqrt(S->anorm * S->anorm + S->alpha * S->alpha + beta * beta + S->damp * S->damp);
        const double dampsq = S->damp * S->damp;
        const double rhobar1 = sqrt(S->rhobar * S->rhobar + dampsq);
        const double cs1 = S->rhobar / rhobar1, sn1 = S->damp / rhobar1;
        const double psi = sn1 * S->phibar;
        const double phibar1 = cs1 * S->phibar;
        const double rho = sqrt(rhobar1 * rhobar1 + beta * beta);
        const double cs = rhobar1 / rho, sn = beta / rho;
        const double theta = sn * alpha;
        S->rhobar = -cs * alpha;
        const double phi = cs * phibar1;
        S->phibar = sn * phibar1;
        const double tau = sn * phi;
        S->t1 = phi / rho;
        S->t2 = -theta / rho;
        S->vscale = beta > 0.0 ? (alpha > 0.0 ? 1.0 / alpha : 0.0) : 1.0;
        S->upd = 1;
        S->alpha = alpha;
        S->res2 += psi * psi;
        S->rnorm = sqrt(S->phibar * S->phibar + S->res2);
        S->arnorm = alpha * fabs(tau);
        S->itn += 1;
        if (S->rnorm <= S->tol * S->bnorm || S->arnorm <= S->tol * S->anorm * S->rnorm) {
            S->done = 1;
            S->flag = 0;
        } else if (S->itn >= S->maxit) {
            S->done = 1;
            S->flag = 1;
        }
    }
}

// INIT (first direction, u = b/beta, beta from beta_kernel): v <- A'u / beta.
// Else, with beta = |u| from the tiles' sums of squares (wave_fixed_sum,
// the same bits in every workgroup) and the current v (stored v times
// vscale):
// the last scalar step's pending update of x and w, then v <- A'u / beta -
// beta v (v unchanged when beta = 0).  Both with per-workgroup partial sums
// of the squares of the new v (fixed order).
template <bool INIT>
__global__ __launch_bounds__(256) void stepa_kernel(int K, int ntiles, const double* normp, const double* rowp,
                                                    const double* colp, const State* S, double* v, double* w,
                                                    double* x, double* part) {
    __shared__ double sh[NG][VE];
    __shared__ double s_beta;
    const int n2 = 2 * K;
    if (S->done) return;
    if (!INIT && threadIdx.x < 64) {
        const double bb = sqrt(wave_fixed_sum(normp, ntiles));  // the bits stepb_kernel gets
        if (threadIdx.x == 0) s_beta = bb;
    }
    const int j = threadIdx.x % VE, g = threadIdx.x / VE;
    const int i = blockIdx.x * VE + j;
    sh[g][j] = i < n2 ? atu_part(K, rowp, colp, i, g) : 0.0;  // used only when beta > 0
    __syncthreads();
    const double beta = INIT ? S->beta : s_beta;
    if (threadIdx.x < 64) {
        double t = 0.0;
        if (threadIdx.x < VE && i < n2) {
            double a = sh[0][j];
#pragma unroll
            for (int q = 1; q < NG; q++) a += sh[q][j];
            if (INIT) {
                t = a * (1.0 / beta);
            } else {
                const double vn = v[i] * S->vscale;
                if (S->upd) {
                    const double wi = w[i];
                    x[i] += S->t1 * wi;
                    w[i] = vn + S->t2 * wi;
                }
                t = beta > 0.0 ? a * (1.0 / beta) - beta * vn : vn;
            }
            v[i] = t;
        }
        const double sq = wave_sum(t * t);
        if (threadIdx.x == 0) part[blockIdx.x] = sq;
    }
}

// First direction, scalar part (one workgroup): alpha = |A'u / beta| and the
// state of iteration 0.
__global__ __launch_bounds__(1024) void vinit_b_kernel(int nb, const double* part, State* S) {
    __shared__ double sh[1024];
    if (S->done) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 1024) s += part[i];
    s = block_sum<1024>(s, sh);
    if (threadIdx.x == 0) {
        const double alpha = sqrt(s);
        S->alpha = alpha;
        S->phibar = S->beta;
        S->rhobar = alpha;
        S->anorm = 0.0;
        S->res2 = 0.0;
        S->rnorm = S->beta;
        S->arnorm = alpha * S->beta;
        S->itn = 0;
        S->vset = 1;
        S->vscale = 1.0;
        S->upd = 0;
        if (alpha == 0.0) {  // A'b = 0: x = 0 is a least-squares solution
            S->done = 1;
            S->flag = 0;
        }
    }
}

// First direction, element part: x = 0; v <- v / alpha (0 when alpha = 0)
// and w = v when v was formed.
__global__ __launch_bounds__(256) void vinit_c_kernel(int n2, const State* S, double* v, double* w, double* x) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n2) return;
    x[i] = 0.0;
    if (S->vset) {
        const double ia = S->alpha > 0.0 ? 1.0 / S->alpha : 0.0;
        const double t = v[i] * ia;
        v[i] = t;
        w[i] = t;
    }
}

// The scalar step, one wave: beta = |u| (the same sum as stepa's), then
// scalar_step on stepa's per-workgroup sums of squares.
__global__ __launch_bounds__(64) void stepb_kernel(int ntiles, const double* normp, int nb, const double* part,
                                                   State* S) {
    if (S->done) return;
    scalar_step(sqrt(wave_fixed_sum(normp, ntiles)), nb, part, S);
}

// After the last iteration: its pending update of x (x += t1 w).
__global__ __launch_bounds__(256) void finish_kernel(int n2, const State* S, const double* w, double* x) {
    if (!S->upd) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n2) x[i] += S->t1 * w[i];
}

}  // namespace lsq

using namespace clmh;

extern "C" int clm_estimate_shifts_ex(clm_ctx* c, int32_t K, int32_t n_theta, const int32_t* d_cl1,
                                      const int32_t* d_cl2, const double* d_shift1d, double tol, int32_t maxit,
                                      double damp, int32_t flags, double* d_x, clm_lsqr_stats* st) {
    using namespace lsq;
    if (!c) return CLM_E_ARG;
    if (flags & ~CLM_FLIP_BY_INDEX) return fail(c, CLM_E_ARG, "unknown flags 0x%x", flags);
    if (st) *st = clm_lsqr_stats{};
    if (K < 2) return fail(c, CLM_E_ARG, "K = %d: at least two images are needed", K);
    if (K > 65536) return fail(c, CLM_E_UNSUP, "K = %d exceeds 65536", K);
    if (n_theta <= 0 || n_theta % 2 || n_theta > 65534) return fail(c, CLM_E_ARG, "invalid n_theta %d", n_theta);
    if (maxit < 0 || !(tol >= 0.0) || !(damp >= 0.0)) return fail(c, CLM_E_ARG, "invalid tol, maxit or damp");
    if (!d_cl1 || !d_cl2 || !d_shift1d || !d_x) return fail(c, CLM_E_ARG, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    const long long P = (long long)K * (K - 1) / 2;
    const int L = n_theta / 2, n2 = 2 * K;
    const int ncb = (K + TC - 1) / TC, nkb = (K + TR - 1) / TR;
    int rc;
    if (c->ls_tiles_K != K) {
        // the pair tiles, row blocks in order (a row block's tiles share its v entries)
        std::vector<int2> h;
        for (int kb = 0; kb < nkb; kb++)
            for (int cb = 0; cb < ncb; cb++)
                if (tile_exists(K, kb, cb)) h.push_back(make_int2(kb, cb));
        if ((rc = ensure(c, c->ls_tiles, h.size() * sizeof(int2)))) return rc;
        HIPCHK(c, hipMemcpyAsync(c->ls_tiles.p, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));  // h is released on return
        c->ls_tiles_K = K;
        c->ls_ntiles = (int)h.size();
    }
    const int ntiles = c->ls_ntiles;
    const int nstep = (n2 + VE - 1) / VE, nelem = (n2 + 255) / 256;
    const size_t nslot = (size_t)ntiles * TR * TC;  // pack and u, tile by tile
    if ((rc = ensure(c, c->ls_pack, nslot * 4)) || (rc = ensure(c, c->ls_u, nslot * 8)) ||
        (rc = ensure(c, c->ls_part, (size_t)(ntiles + nstep) * 8)) ||
        (rc = ensure(c, c->ls_vec, (size_t)n2 * (2 + ncb + nkb) * 8)) ||
        (rc = ensure(c, c->ls_tab, (size_t)(2 * L + 2) * 5 * 8)) || (rc = ensure(c, c->ls_state, sizeof(State))))
        return rc;
    double* vec = (double*)c->ls_vec.p;
    double *v = vec, *w = vec + n2, *rowp = vec + 2 * n2, *colp = rowp + (size_t)ncb * n2;  // [ncb][2K], [nkb][2K]
    double2* ta = (double2*)c->ls_tab.p;
    double* tb = (double*)c->ls_tab.p + 2 * (2 * L + 2);
    State* S = (State*)c->ls_state.p;
    unsigned* pack = (unsigned*)c->ls_pack.p;
    double* u = (double*)c->ls_u.p;
    double* part = (double*)c->ls_part.p;  // the tiles' sums of squares of u
    double* part2 = part + ntiles;         // stepa's per-workgroup sums of squares of v
    const int2* tiles = (const int2*)c->ls_tiles.p;
    State h{};
    h.damp = damp;
    h.tol = tol;
    h.maxit = maxit;
    h.done = maxit == 0 ? 1 : 0;
    h.flag = maxit == 0 ? 1 : 0;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, hipMemcpyAsync(S, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    tables_kernel<<<(2 * L + 2 + 255) / 256, 256, 0, c->stream>>>(L, (flags & CLM_FLIP_BY_INDEX) ? 1 : 0, ta, tb);
    // coefficient tables in LDS when they are small (n_theta up to 816)
    const size_t tab_bytes = (size_t)5 * (2 * L + 2) * 8;
    const bool lt = tab_bytes <= 32768;
    auto pass = [&](bool init) {
        if (init) {
            if (lt)
                pass_kernel<true, true><<<ntiles, 256, tab_bytes, c->stream>>>(K, L, tiles, d_cl1, d_cl2, d_shift1d,
                                                                               pack, ta, tb, v, S, u, part, rowp, colp);
            else
                pass_kernel<true, false><<<ntiles, 256, 0, c->stream>>>(K, L, tiles, d_cl1, d_cl2, d_shift1d, pack, ta,
                                                                        tb, v, S, u, part, rowp, colp);
        } else {
            if (lt)
                pass_kernel<false, true><<<ntiles, 256, tab_bytes, c->stream>>>(K, L, tiles, nullptr, nullptr, nullptr,
                                                                                pack, ta, tb, v, S, u, part, rowp,
                                                                                colp);
            else
                pass_kernel<false, false><<<ntiles, 256, 0, c->stream>>>(K, L, tiles, nullptr, nullptr, nullptr, pack,
                                                                         ta, tb, v, S, u, part, rowp, colp);
        }
    };
    // partials of the tiles that do not exist stay zero (stepa adds every block)
    HIPCHK(c, hipMemsetAsync(rowp, 0, (size_t)(ncb + nkb) * n2 * 8, c->stream));
    pass(true);
    beta_kernel<<<1, 1024, 0, c->stream>>>(ntiles, part, S);
    stepa_kernel<true><<<nstep, 256, 0, c->stream>>>(K, ntiles, part, rowp, colp, S, v, w, d_x, part2);
    vinit_b_kernel<<<1, 1024, 0, c->stream>>>(nstep, part2, S);
    vinit_c_kernel<<<nelem, 256, 0, c->stream>>>(n2, S, v, w, d_x);
    HIPCHK(c, hipGetLastError());
    // launches after the stopping tests pass return at once; the done flag is
    // read back every 8 iterations so a converged solve stops issuing them
    for (int it = 0; it < maxit; it++) {
        pass(false);
        stepa_kernel<false><<<nstep, 256, 0, c->stream>>>(K, ntiles, part, rowp, colp, S, v, w, d_x, part2);
        stepb_kernel<<<1, 64, 0, c->stream>>>(ntiles, part, nstep, part2, S);
        if (it % 8 == 7 && it + 1 < maxit) {
            int done = 0;
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(&done, &S->done, sizeof done, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (done) break;
        }
    }
    finish_kernel<<<nelem, 256, 0, c->stream>>>(n2, S, w, d_x);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&h, S, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        st->iterations = h.itn;
        st->flag = h.flag;
        st->relres = h.bnorm > 0 ? h.rnorm / h.bnorm : 0.0;
        st->normar = h.anorm * h.rnorm > 0 ? h.arnorm / (h.anorm * h.rnorm) : 0.0;
        st->norm_a = h.anorm;
        st->ms = ms;
    }
    return CLM_OK;
}

extern "C" int clm_estimate_shifts(clm_ctx* c, int32_t K, int32_t n_theta, const int32_t* d_cl1,
                                   const int32_t* d_cl2, const double* d_shift1d, double tol, int32_t maxit,
                                   double damp, double* d_x, clm_lsqr_stats* st) {
    return clm_estimate_shifts_ex(c, K, n_theta, d_cl1, d_cl2, d_shift1d, tol, maxit, damp, 0, d_x, st);
}
