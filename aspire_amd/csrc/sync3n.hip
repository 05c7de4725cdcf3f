// sync3n.hip -- the sync3n triangle kernels on the device: the consumer of
// the common lines' relative rotations in the C1 ab-initio pipeline
// (abinitio/C1/cryo_abinitio_C1_worker.m:116-147).
//
// Replaces the reference's three pthread MEX kernels, which all walk every
// triangle i<j<k of the relative rotations R_pairs (3 x 3 x N-choose-2):
//  * triangles_scores_mex.c: per pair, the sum of its triangles' scores, and
//    the histogram of all scores;
//  * pairs_probabilities_mex.c: per pair, the log-likelihoods of its scores
//    under the "indicative" and "arbitrary" models;
//  * signs_times_v_mex.c: the J-synchronization matrix (entries +-1, scores,
//    or pair-score products, signed by the triangle's best configuration)
//    times n_eigs vectors, the step of the power method
//    (cryo_sync3n_Jsync_power_method.m:62).
// Per triangle the four J-configurations are matched,
// c = ||Rij Rjk - Rik||^2 (then with J Rij J, J Rjk J, J Rik J); the best one,
// and for each pair of sides the best configuration that differs from it in
// that pair's relative J, give the score 1 - sqrt(best / alternative).
//
// Formulation: one thread per pair (p, q), looping over the third image m in
// ascending order and adding its triangle's contribution as the reference's
// single-thread loop would, so every per-pair sum is bit-identical to a
// sequential restatement (3x3 products as (a + b) + c, squared norms summed
// in order, no fused multiply-add, IEEE division and square root).  Each
// triangle is therefore evaluated by its three pairs' threads (3x the
// reference's arithmetic) instead of once with atomics, which would make the
// sums depend on scheduling.  A workgroup owns a 16 x 16 block of pairs; the
// rotations of its 16 p-images and 16 q-images against 16 third images at a
// time are staged in LDS from a symmetric, plane-major copy of R
// (Rs[e][a][b] = R_{min(a,b),max(a,b)}[e]), so the loads are coalesced.
// Every score is counted once in the histogram: the (ij,jk) score by the ij
// pair's thread, (ik,jk) by jk's, (ij,ik) by ik's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "clm_ctx.h"

namespace s3 {

constexpr int TB = 16;  // pairs per tile side
constexpr int KC = 16;  // third images per LDS stage
constexpr int MAXE = 8; // vectors of the signs product

enum { SCORES = 0, PROBS = 1, SIGNS = 2 };

__device__ __host__ inline long long pidx(long long N, long long i, long long j) {
    return (2 * N - i - 1) * i / 2 + j - i - 1;
}

// Rs[e][a][b] = R[9*pidx(min,max) + e] (a != b); the diagonal is zero.
__global__ void sym_kernel(int N, const double* __restrict__ R, double* __restrict__ Rs) {
    const long long NN = (long long)N * N;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NN) return;
    const int a = (int)(t / N), b = (int)(t - (long long)a * N);
    if (a == b) {
        for (int e = 0; e < 9; e++) Rs[e * NN + t] = 0.0;
        return;
    }
    const long long pr = pidx(N, std::min(a, b), std::max(a, b));
    for (int e = 0; e < 9; e++) Rs[e * NN + t] = R[9 * pr + e];
}

// Vs[g][a][b] = v[pidx(min,max) + P*g] (one plane per vector; diagonal zero).
__global__ void symv_kernel(int N, int ne, const double* __restrict__ v, double* __restrict__ Vs) {
    const long long NN = (long long)N * N, P = (long long)N * (N - 1) / 2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NN) return;
    const int a = (int)(t / N), b = (int)(t - (long long)a * N);
    const long long pr = a == b ? -1 : pidx(N, std::min(a, b), std::max(a, b));
    for (int g = 0; g < ne; g++) Vs[g * NN + t] = pr < 0 ? 0.0 : v[pr + P * g];
}

__device__ inline void mult3(double* out, const double* A, const double* B) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double x = __dmul_rn(A[i], B[3 * j + 0]);
            const double y = __dmul_rn(A[3 + i], B[3 * j + 1]);
            const double z = __dmul_rn(A[6 + i], B[3 * j + 2]);
            out[3 * j + i] = __dadd_rn(__dadd_rn(x, y), z);
        }
}

__device__ inline double dnorm(const double* A, const double* B) {
    double n = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const double d = __dsub_rn(A[i], B[i]);
        n = __dadd_rn(n, __dmul_rn(d, d));
    }
    return n;
}

// J X J: the entries with exactly one index 3 change sign
__device__ inline void jrj(const double* R, double* A) {
    A[0] = R[0]; A[1] = R[1]; A[2] = -R[2];
    A[3] = R[3]; A[4] = R[4]; A[5] = -R[5];
    A[6] = -R[6]; A[7] = -R[7]; A[8] = R[8];
}

// Configurations (signs of the side pairs (ij,jk), (ik,jk), (ij,ik)): 0 none
// flipped, 1 Rij flipped, 2 Rjk flipped, 3 Rik flipped.  For a best
// configuration b and side pair q, the alternatives are the two
// configurations whose sign at q differs from b's.
__constant__ int ALT0[4][3] = {{1, 2, 1}, {0, 2, 0}, {0, 0, 1}, {1, 0, 0}};
__constant__ int ALT1[4][3] = {{2, 3, 3}, {3, 3, 2}, {3, 1, 3}, {2, 1, 2}};
__constant__ double SGN[4][3] = {{1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, -1}};

// Scores of side pairs qa and qb of the triangle (Rij, Rjk, Rik); returns the
// best configuration.
__device__ inline int triangle2(const double* Rij, const double* Rjk, const double* Rik, int qa, int qb, double& sa,
                                double& sb) {
    double J[9], tmp[9], c0, c1, c2, c3;
    mult3(tmp, Rij, Rjk);
    c0 = dnorm(tmp, Rik);
    jrj(Rik, J);
    c3 = dnorm(tmp, J);
    jrj(Rij, J);
    mult3(tmp, J, Rjk);
    c1 = dnorm(tmp, Rik);
    jrj(Rjk, J);
    mult3(tmp, Rij, J);
    c2 = dnorm(tmp, Rik);
    int b = 0;
    double bv = c0;
    if (c1 < bv) { b = 1; bv = c1; }
    if (c2 < bv) { b = 2; bv = c2; }
    if (c3 < bv) { b = 3; bv = c3; }
    const double cc[4] = {c0, c1, c2, c3};
    double alt = cc[ALT0[b][qa]];
    if (cc[ALT1[b][qa]] < alt) alt = cc[ALT1[b][qa]];
    sa = __dsub_rn(1.0, __dsqrt_rn(__ddiv_rn(bv, alt)));
    alt = cc[ALT0[b][qb]];
    if (cc[ALT1[b][qb]] < alt) alt = cc[ALT1[b][qb]];
    sb = __dsub_rn(1.0, __dsqrt_rn(__ddiv_rn(bv, alt)));
    return b;
}

struct TriArgs {
    int N;
    long long P;
    const double* Rs;           // [9][N][N]
    const int2* tiles;          // (bp, bq), bp <= bq
    int ntiles;
    // SCORES
    int n_intervals;
    const double* thr;          // [n_intervals-1] thresholds (repeated additions of 1/n)
    int thr_lds;                // 1: the thresholds are copied to LDS after the counts
    double* cum;                // [P] or null
    unsigned long long* hist;   // [n_intervals] counts
    // PROBS: P2, A, a, B, b, x0; and, when lform, log A, log(P2 B),
    // log((1 - P2) A), -b/(1 - x0)
    double prm[6];
    double lg[4];
    int lform;
    double* ind;
    double* arb;
    // SIGNS
    const double* Vs;           // [ne][N][N]
    const double* Ss;           // [N][N] pair scores (entries == 2)
    int ne;
    int entries;                // 0 signs, 1 scores, 2 pair-score products
    double* v2;                 // [ne][P]
};

__device__ inline int bin_of(double s, const double* thr, int n) {
    // first l in [0, n-2] with s < thr[l], else n-1 (NaN: n-1)
    if (!(s == s)) return n - 1;
    int l = (int)(s * n);
    l = l < 0 ? 0 : (l > n - 1 ? n - 1 : l);
    while (l > 0 && s < thr[l - 1]) l--;
    while (l < n - 1 && !(s < thr[l])) l++;
    return l;
}

__device__ inline double f_ind(double x, const double* prm) {
    // log(P2*(B*(1-x)^b*exp(-b/(1-x0)*(1-x))) + (1-P2)*A*(1-x)^a)  (pairs_probabilities_mex.c:188)
    const double P2 = prm[0], A = prm[1], a = prm[2], B = prm[3], b = prm[4], x0 = prm[5];
    const double u = 1 - x;
    return log(P2 * (B * pow(u, b) * exp(-b / (1 - x0) * u)) + (1 - P2) * A * pow(u, a));
}

__device__ inline double f_arb(double x, const double* prm) { return log(prm[1] * pow(1 - x, prm[2])); }

// Both log-likelihoods of one score from one logarithm, lu = log(1 - x), for
// A, B > 0, a, b > 0 and 0 <= P2 <= 1 (the same functions; the direct form
// above spends three powers, an exponential and two logarithms):
//   arb = log A + a lu
//   ind = log(e^t1 + e^t2),  t1 = log(P2 B) + b lu - b/(1-x0) (1-x),
//                            t2 = log((1-P2) A) + a lu
// with log(e^t1 + e^t2) = hi + log1p(exp(lo - hi)).
__device__ inline void f_both(double x, const double* prm, const double* lg, double& ind, double& arb) {
    const double u = 1 - x;
    const double lu = log(u);
    const double t1 = lg[1] + prm[4] * lu + lg[3] * u;
    const double t2 = lg[2] + prm[2] * lu;
    arb = lg[0] + prm[2] * lu;
    if (t1 != t1 || t2 != t2) {
        ind = t1 + t2;  // NaN
        return;
    }
    const double hi = t1 > t2 ? t1 : t2, lo = t1 > t2 ? t2 : t1;
    ind = hi == -INFINITY ? hi : hi + log1p(exp(lo - hi));
}

// NEC: vectors the signs product stages (a.ne <= NEC); its third images are
// staged 12 at a time (43 KB of LDS at NEC = 4: three workgroups per CU),
// the other modes' 16 at a time.  The stage size does not change the order
// of any sum.
template <int MODE, int NEC = 1, bool TL = false>
__device__ __forceinline__ void tri_body(TriArgs a) {
    constexpr int KCm = MODE == SIGNS ? 12 : KC;
    __shared__ double Lp[9][KCm][TB];
    __shared__ double Lq[9][KCm][TB];
    __shared__ double Vp[MODE == SIGNS ? NEC : 1][KCm][TB];
    __shared__ double Vq[MODE == SIGNS ? NEC : 1][KCm][TB];
    __shared__ double Sp[MODE == SIGNS ? KCm : 1][TB];
    __shared__ double Sq[MODE == SIGNS ? KCm : 1][TB];
    extern __shared__ unsigned int s_hist[];
    // the thresholds after the counts, 8-byte aligned (the first stage's
    // barrier orders these writes before any read)
    double* s_thr = (double*)(s_hist + ((a.n_intervals + 1) & ~1));
    const int N = a.N;
    const long long NN = (long long)N * N;
    const int tid = threadIdx.x;
    if (MODE == SCORES) {
        for (int l = tid; l < a.n_intervals; l += blockDim.x) s_hist[l] = 0;
        if (TL)
            for (int l = tid; l < a.n_intervals - 1; l += blockDim.x) s_thr[l] = a.thr[l];
    }
    for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        const int2 bt = a.tiles[t];
        const int p0 = bt.x * TB, q0 = bt.y * TB;
        const int tp = tid / TB, tq = tid % TB;
        const int p = p0 + tp, q = q0 + tq;
        const bool act = p < q && q < N;
        double Rpq[9];
#pragma unroll
        for (int e = 0; e < 9; e++) Rpq[e] = act ? a.Rs[e * NN + (long long)p * N + q] : 0.0;
        double acc0 = 0.0, acc1 = 0.0;
        double vacc[MODE == SIGNS ? NEC : 1];
#pragma unroll
        for (int g = 0; g < (MODE == SIGNS ? NEC : 1); g++) vacc[g] = 0.0;
        for (int k0 = 0; k0 < N; k0 += KCm) {
            __syncthreads();
            // stage Rs[e][k][p0..p0+15] and Rs[e][k][q0..q0+15] for k in [k0, k0+KCm)
            for (int x = tid; x < 9 * KCm * TB; x += blockDim.x) {
                const int e = x / (KCm * TB), kk = (x / TB) % KCm, j = x % TB;
                const int k = k0 + kk;
                const bool okp = k < N && p0 + j < N, okq = k < N && q0 + j < N;
                Lp[e][kk][j] = okp ? a.Rs[e * NN + (long long)k * N + p0 + j] : 0.0;
                Lq[e][kk][j] = okq ? a.Rs[e * NN + (long long)k * N + q0 + j] : 0.0;
            }
            if constexpr (MODE == SIGNS) {
                for (int x = tid; x < a.ne * KCm * TB; x += blockDim.x) {
                    const int g = x / (KCm * TB), kk = (x / TB) % KCm, j = x % TB;
                    const int k = k0 + kk;
                    Vp[g][kk][j] = (k < N && p0 + j < N) ? a.Vs[g * NN + (long long)k * N + p0 + j] : 0.0;
                    Vq[g][kk][j] = (k < N && q0 + j < N) ? a.Vs[g * NN + (long long)k * N + q0 + j] : 0.0;
                }
                if (a.entries == 2)
                    for (int x = tid; x < KCm * TB; x += blockDim.x) {
                        const int kk = x / TB, j = x % TB, k = k0 + kk;
                        Sp[kk][j] = (k < N && p0 + j < N) ? a.Ss[(long long)k * N + p0 + j] : 0.0;
                        Sq[kk][j] = (k < N && q0 + j < N) ? a.Ss[(long long)k * N + q0 + j] : 0.0;
                    }
            }
            __syncthreads();
            if (!act) continue;
            const int kend = min(KCm, N - k0);
            for (int kk = 0; kk < kend; kk++) {
                const int m = k0 + kk;
                if (m == p || m == q) continue;
                double Rpm[9], Rqm[9];
#pragma unroll
                for (int e = 0; e < 9; e++) {
                    Rpm[e] = Lp[e][kk][tp];
                    Rqm[e] = Lq[e][kk][tq];
                }
                // the triangle in the reference's (i, j, k) order and this
                // pair's role in it: m > q: pair is ij, sides (p,q),(q,m),(p,m);
                // p < m < q: pair is ik, sides (p,m),(m,q),(p,q);
                // m < p: pair is jk, sides (m,p),(p,q),(m,q)
                const int role = m > q ? 0 : (m > p ? 2 : 1);
                const double* Rij = role == 0 ? Rpq : (role == 2 ? Rpm : Rpm);
                const double* Rjk = role == 0 ? Rqm : (role == 2 ? Rqm : Rpq);
                const double* Rik = role == 0 ? Rpm : (role == 2 ? Rpq : Rqm);
                // side pairs touching this pair: ij -> (0, 2), jk -> (0, 1), ik -> (1, 2);
                // the histogram takes (ij,jk) from ij, (ik,jk) from jk, (ij,ik) from ik
                const int qa = role == 2 ? 1 : 0;
                const int qb = role == 1 ? 1 : 2;
                double sa, sb;
                const int best = triangle2(Rij, Rjk, Rik, qa, qb, sa, sb);
                if constexpr (MODE == SCORES) {
                    acc0 += sa + sb;
                    const double hs = role == 0 ? sa : (role == 1 ? sb : sb);
                    const int bin = TL ? bin_of(hs, s_thr, a.n_intervals) : bin_of(hs, a.thr, a.n_intervals);
                    atomicAdd(&s_hist[bin], 1u);
                } else if constexpr (MODE == PROBS) {
                    if (a.lform) {
                        double ia, aa, ib, ab;
                        f_both(sa, a.prm, a.lg, ia, aa);
                        f_both(sb, a.prm, a.lg, ib, ab);
                        acc0 += ia + ib;
                        acc1 += aa + ab;
                    } else {
                        acc0 += f_ind(sa, a.prm) + f_ind(sb, a.prm);
                        acc1 += f_arb(sa, a.prm) + f_arb(sb, a.prm);
                    }
                } else {
                    double ea, eb;
                    const double sga = SGN[best][qa], sgb = SGN[best][qb];
                    if (a.entries == 1) {
                        ea = sga * sa;
                        eb = sgb * sb;
                    } else if (a.entries == 2) {
                        // pair_scores[ij]*pair_scores[jk]*pair_scores[ik] in that order
                        const double spq = a.Ss[(long long)p * N + q], spm = Sp[kk][tp], sqm = Sq[kk][tq];
                        const double sij = role == 0 ? spq : spm, sjk = role == 0 ? sqm : (role == 2 ? sqm : spq);
                        const double sik = role == 0 ? spm : (role == 2 ? spq : sqm);
                        const double tsc = sij * sjk * sik;
                        ea = sga * tsc;
                        eb = sgb * tsc;
                    } else {
                        ea = sga;
                        eb = sgb;
                    }
                    // ij: v2 += e(ij,jk) v[jk] + e(ij,ik) v[ik]; jk: e(ij,jk) v[ij] + e(ik,jk) v[ik];
                    // ik: e(ij,ik) v[ij] + e(ik,jk) v[jk]
#pragma unroll
                    for (int g = 0; g < NEC; g++) {
                        if (g >= a.ne) break;
                        const double vpm = Vp[g][kk][tp], vqm = Vq[g][kk][tq];
                        const double x = role == 0 ? ea * vqm : (role == 1 ? ea * vpm : eb * vpm);
                        const double y = role == 0 ? eb * vpm : (role == 1 ? eb * vqm : ea * vqm);
                        vacc[g] += x + y;
                    }
                }
            }
        }
        if (act) {
            const long long e = pidx(N, p, q);
            if constexpr (MODE == SCORES) {
                if (a.cum) a.cum[e] = acc0;
            } else if constexpr (MODE == PROBS) {
                a.ind[e] = acc0;
                a.arb[e] = acc1;
            } else {
#pragma unroll
                for (int g = 0; g < NEC; g++)
                    if (g < a.ne) a.v2[e + a.P * g] = vacc[g];
            }
        }
    }
    if (MODE == SCORES) {
        __syncthreads();
        for (int l = tid; l < a.n_intervals; l += blockDim.x)
            if (s_hist[l]) atomicAdd(&a.hist[l], (unsigned long long)s_hist[l]);
    }
}

__global__ void hist_out_kernel(int n, const unsigned long long* h, double* out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) out[l] = (double)h[l];
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) p[l] = 0ull;
}

// One kernel per mode, for their own occupancy: the scores fit 128 VGPRs
// (four waves per SIMD); the probabilities (log, exp, log1p) need more and
// run at two; the signs product runs at three for up to four vectors (its
// LDS and registers sized for them), at two for up to eight.
template <bool TL>
__global__ __launch_bounds__(256, 4) void tri_scores_kernel(TriArgs a) { tri_body<SCORES, 1, TL>(a); }
__global__ __launch_bounds__(256) void tri_probs_kernel(TriArgs a) { tri_body<PROBS>(a); }
__global__ __launch_bounds__(256, 3) void tri_signs4_kernel(TriArgs a) { tri_body<SIGNS, 4>(a); }
__global__ __launch_bounds__(256) void tri_signs8_kernel(TriArgs a) { tri_body<SIGNS, MAXE>(a); }

}  // namespace s3

using namespace clmh;

namespace {

// Shared set-up: argument checks, the symmetric copy of R and the pair tiles.
int s3_prepare(clm_ctx* c, int32_t N, const double* d_R, s3::TriArgs& a, clm_sync3n_stats* st) {
    using namespace s3;
    if (!c) return CLM_E_ARG;
    if (st) *st = clm_sync3n_stats{};
    if (N < 3 || N > 46340) return fail(c, CLM_E_ARG, "N = %d outside [3, 46340]", N);
    if (!d_R) return fail(c, CLM_E_ARG, "null R_pairs");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    const long long NN = (long long)N * N;
    const int nb = (N + TB - 1) / TB;
    const int nt = nb * (nb + 1) / 2;
    if ((rc = ensure(c, c->s3_rs, (size_t)9 * NN * sizeof(double)))) return rc;
    if (c->s3_nb != nb) {
        // tiles in row order of the upper triangle of blocks: consecutive
        // workgroups share the p-block (its staged rows meet in L2)
        std::vector<int2> h;
        h.reserve(nt);
        for (int bp = 0; bp < nb; bp++)
            for (int bq = bp; bq < nb; bq++) h.push_back(make_int2(bp, bq));
        if ((rc = ensure(c, c->s3_tiles, h.size() * sizeof(int2)))) return rc;
        HIPCHK(c, hipMemcpyAsync(c->s3_tiles.p, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));  // h is released on return
        c->s3_nb = nb;
    }
    a = TriArgs{};
    a.N = N;
    a.P = (long long)N * (N - 1) / 2;
    a.Rs = (const double*)c->s3_rs.p;
    a.tiles = (const int2*)c->s3_tiles.p;
    a.ntiles = nt;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    sym_kernel<<<(unsigned)((NN + 255) / 256), 256, 0, c->stream>>>(N, d_R, (double*)c->s3_rs.p);
    HIPCHK(c, hipGetLastError());
    return CLM_OK;
}

int s3_finish(clm_ctx* c, const s3::TriArgs& a, clm_sync3n_stats* st) {
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st) {
        float x = 0.f, y = 0.f;
        HIPCHK(c, hipEventElapsedTime(&x, c->ev[0], c->ev[1]));
        HIPCHK(c, hipEventElapsedTime(&y, c->ev[1], c->ev[2]));
        st->ms_layout = x;
        st->ms_kernel = y;
        st->triangles = (long long)a.N * (a.N - 1) * (a.N - 2) / 6;
    }
    return CLM_OK;
}

template <int MODE>
int s3_launch(clm_ctx* c, const s3::TriArgs& a, size_t dyn_lds) {
    const unsigned grid = (unsigned)std::min<long long>(a.ntiles, (long long)c->n_cu * 8);
    if constexpr (MODE == s3::SCORES)
        if (a.thr_lds)
            s3::tri_scores_kernel<true><<<grid, 256, dyn_lds, c->stream>>>(a);
        else
            s3::tri_scores_kernel<false><<<grid, 256, dyn_lds, c->stream>>>(a);
    else if constexpr (MODE == s3::PROBS)
        s3::tri_probs_kernel<<<grid, 256, dyn_lds, c->stream>>>(a);
    else if (a.ne <= 4)
        s3::tri_signs4_kernel<<<grid, 256, dyn_lds, c->stream>>>(a);
    else
        s3::tri_signs8_kernel<<<grid, 256, dyn_lds, c->stream>>>(a);
    HIPCHK(c, hipGetLastError());
    return CLM_OK;
}

}  // namespace

extern "C" int clm_sync3n_triangles_scores(clm_ctx* c, int32_t N, const double* d_R, int32_t n_intervals,
                                           double* d_cum, double* d_hist, clm_sync3n_stats* st) {
    using namespace s3;
    if (c && (n_intervals < 1 || n_intervals > 8192))
        return fail(c, CLM_E_ARG, "n_intervals = %d outside [1, 8192]", n_intervals);
    if (c && !d_hist) return fail(c, CLM_E_ARG, "null histogram");
    TriArgs a;
    int rc = s3_prepare(c, N, d_R, a, st);
    if (rc) return rc;
    // thresholds as the reference forms them: threshold += h, n-1 times
    std::vector<double> thr(std::max(n_intervals - 1, 1));
    const double h = 1. / n_intervals;
    double t = 0;
    for (int l = 0; l < n_intervals - 1; l++) {
        t += h;
        thr[l] = t;
    }
    if ((rc = ensure(c, c->s3_thr, thr.size() * sizeof(double))) ||
        (rc = ensure(c, c->s3_hist, (size_t)n_intervals * sizeof(unsigned long long))))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->s3_thr.p, thr.data(), thr.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    zero_u64_kernel<<<(n_intervals + 255) / 256, 256, 0, c->stream>>>((unsigned long long*)c->s3_hist.p, n_intervals);
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    a.n_intervals = n_intervals;
    a.thr = (const double*)c->s3_thr.p;
    a.cum = d_cum;
    a.hist = (unsigned long long*)c->s3_hist.p;
    // thresholds in LDS for up to 1024 intervals (8 KB)
    a.thr_lds = n_intervals <= 1024;
    const size_t lds = (size_t)((n_intervals + 1) & ~1) * sizeof(unsigned int) +
                       (a.thr_lds ? (size_t)(n_intervals - 1) * sizeof(double) : 0);
    if ((rc = s3_launch<SCORES>(c, a, lds))) return rc;
    hist_out_kernel<<<(n_intervals + 255) / 256, 256, 0, c->stream>>>(n_intervals, a.hist, d_hist);
    HIPCHK(c, hipGetLastError());
    rc = s3_finish(c, a, st);
    HIPCHK(c, hipStreamSynchronize(c->stream));  // thr is released on return
    return rc;
}

extern "C" int clm_sync3n_pairs_probabilities(clm_ctx* c, int32_t N, const double* d_R, const double* prm,
                                              double* d_ln_f_ind, double* d_ln_f_arb, clm_sync3n_stats* st) {
    using namespace s3;
    if (c && (!prm || !d_ln_f_ind || !d_ln_f_arb)) return fail(c, CLM_E_ARG, "null pointer");
    TriArgs a;
    int rc = s3_prepare(c, N, d_R, a, st);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    for (int i = 0; i < 6; i++) a.prm[i] = prm[i];
    {
        const double P2 = prm[0], A = prm[1], ea = prm[2], B = prm[3], eb = prm[4], x0 = prm[5];
        a.lform = A > 0 && B > 0 && ea > 0 && eb > 0 && P2 >= 0 && P2 <= 1 && x0 != 1;
        a.lg[0] = std::log(A);
        a.lg[1] = std::log(P2 * B);
        a.lg[2] = std::log((1 - P2) * A);
        a.lg[3] = -eb / (1 - x0);
    }
    a.ind = d_ln_f_ind;
    a.arb = d_ln_f_arb;
    if ((rc = s3_launch<PROBS>(c, a, 0))) return rc;
    return s3_finish(c, a, st);
}

extern "C" int clm_sync3n_signs_times_v(clm_ctx* c, int32_t N, const double* d_R, const double* d_v, int32_t n_eigs,
                                        int32_t scores_as_entries, const double* d_pair_scores, double* d_v2,
                                        clm_sync3n_stats* st) {
    using namespace s3;
    if (c && (n_eigs < 1 || n_eigs > MAXE)) return fail(c, CLM_E_ARG, "n_eigs = %d outside [1, %d]", n_eigs, MAXE);
    if (c && (scores_as_entries < 0 || scores_as_entries > 2))
        return fail(c, CLM_E_ARG, "scores_as_entries = %d outside {0, 1, 2}", scores_as_entries);
    if (c && (!d_v || !d_v2 || (scores_as_entries == 2 && !d_pair_scores))) return fail(c, CLM_E_ARG, "null pointer");
    TriArgs a;
    int rc = s3_prepare(c, N, d_R, a, st);
    if (rc) return rc;
    const long long NN = (long long)N * N;
    if ((rc = ensure(c, c->s3_vs, (size_t)n_eigs * NN * sizeof(double)))) return rc;
    symv_kernel<<<(unsigned)((NN + 255) / 256), 256, 0, c->stream>>>(N, n_eigs, d_v, (double*)c->s3_vs.p);
    if (scores_as_entries == 2) {
        if ((rc = ensure(c, c->s3_ss, (size_t)NN * sizeof(double)))) return rc;
        symv_kernel<<<(unsigned)((NN + 255) / 256), 256, 0, c->stream>>>(N, 1, d_pair_scores, (double*)c->s3_ss.p);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    a.Vs = (const double*)c->s3_vs.p;
    a.Ss = scores_as_entries == 2 ? (const double*)c->s3_ss.p : nullptr;
    a.ne = n_eigs;
    a.entries = scores_as_entries;
    a.v2 = d_v2;
    if ((rc = s3_launch<SIGNS>(c, a, 0))) return rc;
    return s3_finish(c, a, st);
}
