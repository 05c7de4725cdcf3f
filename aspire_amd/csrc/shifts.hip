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
// Memory per iteration (K = 5000, 12.5M pairs): the packed lines (4 B/pair)
// are read three times and u (8 B/pair) read three times and written once:
// about 0.5 GB, HBM-bound work with every reduction in a fixed order
// (results do not depend on scheduling).
#include <hip/hip_runtime.h>

#include <cmath>

#include "clm_ctx.h"

namespace lsq {

struct State {
    double alpha, beta, bnorm, phibar, rhobar, anorm, res2, rnorm, arnorm;
    double damp, tol;
    double t1, t2, ia;  // this iteration's update of x, w and v
    int itn, maxit, done, flag;
    int upd;            // 1: this iteration's scalar step ran, apply its update
};

__device__ inline long long first(long long k1, long long K) { return k1 * K - k1 * (k1 + 1) / 2; }

constexpr int NS = 16;  // k1 residue classes of the column part of A'u (partial sums)

// ta[c] = (sin a, cos a), a = (c-1)*pi/L; tb[c] = (-sin b', -cos b', sign):
// b = (c-1)*pi/L, b' = b or b - pi, sign = +1 (b < pi) or -1 (b >= pi), for
// line numbers c = 0..2L (0 = no line; cryo_clmatrix_cpu.m:722-741).  Entry
// 2L+1 is all zeros: the row of a pair whose line numbers lie outside
// [0, 2L].
__global__ void tables_kernel(int L, double2* ta, double* tb) {
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
    if (!(b < M_PI)) {
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

// pack[e] = cl1 | cl2 << 16 (both 2L+1 when either is outside [0, 2L]);
// u[e] = b[e];  part[k1] = sum over row k1 of b^2.
__global__ __launch_bounds__(256) void init_kernel(int K, int L, const int* cl1, const int* cl2, const double* b,
                                                   unsigned* pack, double* u, double* part) {
    __shared__ double sh[256];
    const int k1 = blockIdx.x;
    const long long e0 = first(k1, K);
    const int n = K - 1 - k1;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        const long long e = e0 + j;
        const int a = cl1[e], c2 = cl2[e];
        const bool ok = a >= 0 && a <= 2 * L && c2 >= 0 && c2 <= 2 * L;
        const unsigned z = (unsigned)(2 * L + 1);
        pack[e] = ok ? ((unsigned)a | ((unsigned)c2 << 16)) : (z | (z << 16));
        const double be = b[e];
        u[e] = be;
        s += be * be;
    }
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) part[k1] = s;
}

// u <- A v - alpha * u / beta_prev (u holds the previous unnormalized vector,
// whose norm is beta_prev); part[k1] = sum over row k1 of the new u^2.
__global__ __launch_bounds__(256) void av_kernel(int K, const unsigned* pack, const double2* ta, const double* tb,
                                                 const double* v, const State* S, double* u, double* part) {
    __shared__ double sh[256];
    if (S->done) return;
    const int k1 = blockIdx.x;
    const long long e0 = first(k1, K);
    const int n = K - 1 - k1;
    const double ca = S->alpha / S->beta;
    const double v0 = v[2 * k1], v1 = v[2 * k1 + 1];
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        const long long e = e0 + j;
        const int k2 = k1 + 1 + j;
        const unsigned pk = pack[e];
        const double2 A = ta[pk & 0xffff];
        const double* B = tb + 3 * (pk >> 16);
        const double sg = B[2];
        const double av = sg * A.x * v0 + sg * A.y * v1 + B[0] * v[2 * k2] + B[1] * v[2 * k2 + 1];
        const double un = av - ca * u[e];
        u[e] = un;
        s += un * un;
    }
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) part[k1] = s;
}

// beta = sqrt(sum of part); first call (init): bnorm = beta; later: the
// running Frobenius estimate of A.
__global__ __launch_bounds__(1024) void beta_kernel(int n, const double* part, State* S, int init) {
    __shared__ double sh[1024];
    if (S->done) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += part[i];
    s = block_sum<1024>(s, sh);
    if (threadIdx.x == 0) {
        const double beta = sqrt(s);
        S->beta = beta;
        if (init) {
            S->bnorm = beta;
            if (beta == 0.0) {  // b = 0: x = 0 is the solution
                S->done = 1;
                S->flag = 0;
            }
        } else if (beta > 0.0) {
            S->anorm = sqrt(S->anorm * S->anorm + S->alpha * S->alpha + beta * beta + S->damp * S->damp);
        }
    }
}

// A'u in one launch, in fixed-order partial sums:
//  blocks [0, K): vt_row[2k..2k+1] = sum over pairs (k, k2 > k) of the k1
//    coefficients * u (one image, its pairs contiguous);
//  blocks [K, K + ceil(K/16)*NS): vt_col[s][2k..2k+1] = sum over pairs
//    (k1 < k, k), k1 = s*16 + q (mod 16*NS), of the k2 coefficients * u: 16
//    consecutive images per block, so for one k1 the 16 pairs are
//    consecutive; lane q of image i walks its residue class.
__global__ __launch_bounds__(256) void atu_kernel(int K, const unsigned* pack, const double2* ta, const double* tb,
                                                  const double* u, const State* S, double* vt_row, double* vt_col) {
    __shared__ double sh0[256], sh1[256];
    if (S->done) return;
    double s0 = 0.0, s1 = 0.0;
    if ((int)blockIdx.x < K) {
        const int k = blockIdx.x;
        const long long e0 = first(k, K);
        const int n = K - 1 - k;
        for (int j = threadIdx.x; j < n; j += 256) {
            const unsigned pk = pack[e0 + j];
            const double2 A = ta[pk & 0xffff];
            const double sg = tb[3 * (pk >> 16) + 2];
            const double ue = u[e0 + j];
            s0 += sg * A.x * ue;
            s1 += sg * A.y * ue;
        }
        s0 = block_sum<256>(s0, sh0);
        s1 = block_sum<256>(s1, sh0);
        if (threadIdx.x == 0) {
            vt_row[2 * k] = s0;
            vt_row[2 * k + 1] = s1;
        }
        return;
    }
    const int cb = blockIdx.x - K;
    const int sp = cb % NS, kb = cb / NS;
    const int i = threadIdx.x & 15, q = threadIdx.x >> 4;
    const int k = kb * 16 + i;
    if (k < K) {
        for (int k1 = sp * 16 + q; k1 < k; k1 += 16 * NS) {
            const long long e = first(k1, K) + (k - k1 - 1);
            const unsigned pk = pack[e];
            const double* B = tb + 3 * (pk >> 16);
            const double ue = u[e];
            s0 += B[0] * ue;
            s1 += B[1] * ue;
        }
    }
    sh0[threadIdx.x] = s0;
    sh1[threadIdx.x] = s1;
    __syncthreads();
    if (q == 0 && k < K) {
        double a0 = 0.0, a1 = 0.0;
        for (int r = 0; r < 16; r++) {
            a0 += sh0[r * 16 + i];
            a1 += sh1[r * 16 + i];
        }
        vt_col[(size_t)sp * 2 * K + 2 * k] = a0;
        vt_col[(size_t)sp * 2 * K + 2 * k + 1] = a1;
    }
}

__device__ inline double atu_value(int n2, const double* vt_row, const double* vt_col, int i) {
    double t = vt_row[i];
#pragma unroll
    for (int sp = 0; sp < NS; sp++) t += vt_col[(size_t)sp * n2 + i];
    return t;
}

// First direction: v = A'u / |A'u| with u = b/beta; w = v; x = 0.
__global__ __launch_bounds__(1024) void vinit_kernel(int n2, const double* vt_row, const double* vt_col, State* S,
                                                     double* v, double* w, double* x) {
    __shared__ double sh[1024];
    if (S->done) {
        for (int i = threadIdx.x; i < n2; i += 1024) x[i] = 0.0;
        return;
    }
    const double ib = 1.0 / S->beta;
    double s = 0.0;
    for (int i = threadIdx.x; i < n2; i += 1024) {
        const double t = atu_value(n2, vt_row, vt_col, i) * ib;
        v[i] = t;
        s += t * t;
    }
    s = block_sum<1024>(s, sh);
    const double alpha = sqrt(s);
    const double ia = alpha > 0.0 ? 1.0 / alpha : 0.0;
    for (int i = threadIdx.x; i < n2; i += 1024) {
        const double t = v[i] * ia;
        v[i] = t;
        w[i] = t;
        x[i] = 0.0;
    }
    if (threadIdx.x == 0) {
        S->alpha = alpha;
        S->phibar = S->beta;
        S->rhobar = alpha;
        S->anorm = 0.0;
        S->res2 = 0.0;
        S->rnorm = S->beta;
        S->arnorm = alpha * S->beta;
        S->itn = 0;
        if (alpha == 0.0) {  // A'b = 0: x = 0 is a least-squares solution
            S->done = 1;
            S->flag = 0;
        }
    }
}

// One LSQR step after the new u and beta, in three launches:
//  stepa (one thread per element): v <- A'u / beta - beta v, and per-block
//    partial sums of its squares (fixed order);
//  stepb (one workgroup): alpha = |v|, the damping and plane rotations, the
//    stopping tests; the update's coefficients go to the state;
//  stepc (one thread per element): v <- v / alpha, x and w.
// beta = 0 (b in the range of A): v and alpha stay, the rotation below makes
// the last update of x and phibar becomes 0.  stepc applies the update of
// the iteration whose tests pass, as the last one.
__global__ __launch_bounds__(256) void stepa_kernel(int n2, const double* vt_row, const double* vt_col,
                                                    const State* S, double* v, double* part) {
    __shared__ double sh[256];
    if (S->done || !(S->beta > 0.0)) return;
    const double beta = S->beta, ib = 1.0 / beta;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double t = 0.0;
    if (i < n2) {
        t = atu_value(n2, vt_row, vt_col, i) * ib - beta * v[i];
        v[i] = t;
    }
    const double sq = block_sum<256>(t * t, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = sq;
}

__global__ __launch_bounds__(1024) void stepb_kernel(int nb, const double* part, State* S) {
    __shared__ double sh[1024];
    if (S->done) {
        if (threadIdx.x == 0) S->upd = 0;
        return;
    }
    const double beta = S->beta;
    double alpha = S->alpha;
    if (beta > 0.0) {
        double sq = 0.0;
        for (int i = threadIdx.x; i < nb; i += 1024) sq += part[i];
        sq = block_sum<1024>(sq, sh);
        alpha = sqrt(sq);
    }
    if (threadIdx.x == 0) {
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
        S->ia = beta > 0.0 ? (alpha > 0.0 ? 1.0 / alpha : 0.0) : 1.0;
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

__global__ __launch_bounds__(256) void stepc_kernel(int n2, const State* S, double* v, double* w, double* x) {
    if (!S->upd) return;
    const double t1 = S->t1, t2 = S->t2, ia = S->ia;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n2) {
        const double vi = v[i] * ia;
        v[i] = vi;
        const double wi = w[i];
        x[i] += t1 * wi;
        w[i] = vi + t2 * wi;
    }
}

}  // namespace lsq

using namespace clmh;

extern "C" int clm_estimate_shifts(clm_ctx* c, int32_t K, int32_t n_theta, const int32_t* d_cl1,
                                   const int32_t* d_cl2, const double* d_shift1d, double tol, int32_t maxit,
                                   double damp, double* d_x, clm_lsqr_stats* st) {
    using namespace lsq;
    if (!c) return CLM_E_ARG;
    if (st) *st = clm_lsqr_stats{};
    if (K < 2) return fail(c, CLM_E_ARG, "K = %d: at least two images are needed", K);
    if (K > 65536) return fail(c, CLM_E_UNSUP, "K = %d exceeds 65536", K);
    if (n_theta <= 0 || n_theta % 2 || n_theta > 65534) return fail(c, CLM_E_ARG, "invalid n_theta %d", n_theta);
    if (maxit < 0 || !(tol >= 0.0) || !(damp >= 0.0)) return fail(c, CLM_E_ARG, "invalid tol, maxit or damp");
    if (!d_cl1 || !d_cl2 || !d_shift1d || !d_x) return fail(c, CLM_E_ARG, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    const long long P = (long long)K * (K - 1) / 2;
    const int L = n_theta / 2, n2 = 2 * K;
    int rc;
    if ((rc = ensure(c, c->ls_pack, (size_t)P * 4)) || (rc = ensure(c, c->ls_u, (size_t)P * 8)) ||
        (rc = ensure(c, c->ls_part, (size_t)K * 8)) || (rc = ensure(c, c->ls_vec, (size_t)n2 * (3 + NS) * 8)) ||
        (rc = ensure(c, c->ls_tab, (size_t)(2 * L + 2) * 5 * 8)) || (rc = ensure(c, c->ls_state, sizeof(State))))
        return rc;
    double* vec = (double*)c->ls_vec.p;
    double *v = vec, *w = vec + n2, *vt_row = vec + 2 * n2, *vt_col = vec + 3 * n2;  // vt_col: NS x 2K
    double2* ta = (double2*)c->ls_tab.p;
    double* tb = (double*)c->ls_tab.p + 2 * (2 * L + 2);
    State* S = (State*)c->ls_state.p;
    unsigned* pack = (unsigned*)c->ls_pack.p;
    double* u = (double*)c->ls_u.p;
    double* part = (double*)c->ls_part.p;
    State h{};
    h.damp = damp;
    h.tol = tol;
    h.maxit = maxit;
    h.done = maxit == 0 ? 1 : 0;
    h.flag = maxit == 0 ? 1 : 0;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, hipMemcpyAsync(S, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    tables_kernel<<<(2 * L + 2 + 255) / 256, 256, 0, c->stream>>>(L, ta, tb);
    init_kernel<<<K, 256, 0, c->stream>>>(K, L, d_cl1, d_cl2, d_shift1d, pack, u, part);
    beta_kernel<<<1, 1024, 0, c->stream>>>(K, part, S, 1);
    const int natu = K + (K + 15) / 16 * NS;
    const int nstep = (n2 + 255) / 256;  // <= K: its partial sums fit in part
    atu_kernel<<<natu, 256, 0, c->stream>>>(K, pack, ta, tb, u, S, vt_row, vt_col);
    vinit_kernel<<<1, 1024, 0, c->stream>>>(n2, vt_row, vt_col, S, v, w, d_x);
    HIPCHK(c, hipGetLastError());
    // launches after the stopping tests pass return at once; the done flag is
    // read back every 8 iterations so a converged solve stops issuing them
    for (int it = 0; it < maxit; it++) {
        av_kernel<<<K, 256, 0, c->stream>>>(K, pack, ta, tb, v, S, u, part);
        beta_kernel<<<1, 1024, 0, c->stream>>>(K, part, S, 0);
        atu_kernel<<<natu, 256, 0, c->stream>>>(K, pack, ta, tb, u, S, vt_row, vt_col);
        stepa_kernel<<<nstep, 256, 0, c->stream>>>(n2, vt_row, vt_col, S, v, part);
        stepb_kernel<<<1, 1024, 0, c->stream>>>(nstep, part, S);
        stepc_kernel<<<nstep, 256, 0, c->stream>>>(n2, S, v, w, d_x);
        if (it % 8 == 7 && it + 1 < maxit) {
            int done = 0;
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(&done, &S->done, sizeof done, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (done) break;
        }
    }
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
