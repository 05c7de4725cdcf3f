// vote.hip -- the common-lines voting synchronization matrix on the device:
// the consumer of cryo_clmatrix's clstack in the C1 ab-initio pipeline.
//
// Replaces S = cryo_syncmatrix_vote(clstack, n_theta)
// (abinitio/C1/sync2n/cryo_syncmatrix_vote.m:38-57, called from
// abinitio/C1/cryo_abinitio_C1_worker.m:160) without reference rotations.
// For every pair (k1, k2) with a common line:
//  * every third image k3 with lines to both gives a triangle; its angle
//    phi = acos((c3 - c1 c2) / (sin t1 sin t2)) in degrees, t1 = (l13-l12)2pi/L,
//    t2 = (l21-l23)2pi/L, t3 = (l32-l31)2pi/L, kept when
//    1 + 2c1c2c3 - (c1^2+c2^2+c3^2) > 1e-5            (cryo_voteIJ.m:40-63)
//  * the angles vote in a Gaussian histogram (sigma 3) on 60 points of
//    [0, 180]; the first peak wins; triangles within 6 degrees of it are kept
//    (cryo_voteIJ.m:90-107)
//  * each kept triangle gives a relative rotation (rotratio_eulerangle_vec.m);
//    their top-left 2x2 blocks are averaged into S's block (k1, k2) and its
//    transpose into (k2, k1) (cryo_syncmatrixIJ_vote.m:125-164).
// A pair without a common line, or without voted triangles, gets a zero
// block; the diagonal blocks are the identity.  cos_phi2 and c_alpha are
// clamped to [-1, 1] before acos.
//
// Two persistent kernels, one workgroup per pair.  peak_kernel: each wave
// takes chunks of 64 third images -- one angle per lane, passed to every lane
// through scalar registers -- then each lane owns one of the 60 histogram
// points and adds the chunk's 64 Gaussians in k3 order; the four waves'
// histograms are added in a fixed order and the first peak is stored per
// pair.  rot_kernel: selection and rotations, one thread per third image,
// fixed-order block sums.  The cosines and sines of the integer line
// differences come from a table built with the formulas' own expression.
// The work is FP64 transcendental evaluations (60 exponentials per
// triangle); memory is the clstack in two 16-bit orientations (row k1/k2
// and column k1/k2 reads are contiguous), the per-pair peaks and S.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "clm_ctx.h"

namespace vote {

constexpr int NT = 60;  // histogram points (ntics)

// M[i*K + j] = clstack(i, j), MT[j*K + i] = clstack(i, j) (1-based lines, 0 = none)
__global__ void build_kernel(int K, const int* cl1, const int* cl2, unsigned short* M, unsigned short* MT) {
    const int k1 = blockIdx.x;
    const long long e0 = (long long)k1 * K - (long long)k1 * (k1 + 1) / 2;
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        if (j <= k1) {
            if (j == k1) {
                M[(size_t)k1 * K + k1] = 0;
                MT[(size_t)k1 * K + k1] = 0;
            }
            continue;
        }
        const long long e = e0 + (j - k1 - 1);
        const unsigned short a = (unsigned short)cl1[e], b = (unsigned short)cl2[e];
        M[(size_t)k1 * K + j] = a;   // clstack(k1, k2)
        M[(size_t)j * K + k1] = b;   // clstack(k2, k1)
        MT[(size_t)j * K + k1] = a;
        MT[(size_t)k1 * K + j] = b;
    }
}

__global__ void diag_kernel(int K, double* S) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const size_t n2 = 2 * (size_t)K;
    S[(2 * (size_t)k) * n2 + 2 * k] = 1.0;
    S[(2 * (size_t)k + 1) * n2 + 2 * k] = 0.0;
    S[(2 * (size_t)k) * n2 + 2 * k + 1] = 0.0;
    S[(2 * (size_t)k + 1) * n2 + 2 * k + 1] = 1.0;
}

__device__ inline double tick(int d) { return d == NT - 1 ? 180.0 : d * (180.0 / (NT - 1)); }

// cs[d + n_theta] = (cos, sin) of d * 2 * pi / n_theta for the integer line
// differences d in [-n_theta, n_theta]: the same expression, evaluated
// once, as the angle of a line difference in the triangle and rotation
// formulas (2 * pi * d / L rounds the same way, the factor 2 being exact).
__global__ void trig_kernel(int n_theta, double2* cs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > 2 * n_theta) return;
    const double d = (double)(i - n_theta);
    const double th = d * 2 * M_PI / (double)n_theta;
    cs[i] = make_double2(cos(th), sin(th));
}

// Entry of line difference d (clamped to the table, so that a malformed
// clstack reads in bounds).
__device__ inline double2 trig(const double2* cs, int nt, int d) {
    const int i = d + nt;
    return cs[i < 0 ? 0 : (i > 2 * nt ? 2 * nt : i)];
}

// Cosine of the angle of triangle (k1, k2, k3), clamped to [-1, 1]; NaN
// when k3 takes no part.  The clstack rows and columns of k1 and k2 are read
// at index i (i = k3, or k3's place in a staged chunk of them).
__device__ inline double vote_cos(const unsigned short* r1, const unsigned short* r2, const unsigned short* c1_,
                                  const unsigned short* c2_, int i, int k1, int k3, int l12, int l21,
                                  const double2* cs, int nt) {
    const int l13 = r1[i], l23 = r2[i];
    if (k3 == k1 || l13 == 0 || l23 == 0) return NAN;
    const int l31 = c1_[i], l32 = c2_[i];
    // theta1 = (l13 - l12) 2pi/L, theta2 = (l21 - l23) 2pi/L, theta3 = (l32 - l31) 2pi/L
    const double2 t1 = trig(cs, nt, l13 - l12), t2 = trig(cs, nt, l21 - l23), t3 = trig(cs, nt, l32 - l31);
    const double c1 = t1.x, c2 = t2.x, c3 = t3.x;
    const double cond = 1 + 2 * c1 * c2 * c3 - (c1 * c1 + c2 * c2 + c3 * c3);
    if (!(cond > 1.0e-5)) return NAN;
    double cp = (c3 - c1 * c2) / (t1.y * t2.y);
    cp = cp > 1 ? 1 : (cp < -1 ? -1 : cp);
    return cp;
}

// Angle (degrees) of triangle (k1, k2, k3), NaN when k3 takes no part.
__device__ inline double vote_angle(const unsigned short* r1, const unsigned short* r2, const unsigned short* c1_,
                                    const unsigned short* c2_, int k1, int k3, int l12, int l21,
                                    const double2* cs, int nt) {
    const double cp = vote_cos(r1, r2, c1_, c2_, k3, k1, k3, l12, l21, cs, nt);
    return isnan(cp) ? cp : acos(cp) * 180 / M_PI;
}

// The (cos, sin) table in LDS when it fits (dynamic LDS of the launch), the
// same values read from global memory otherwise.
__device__ inline const double2* stage_trig(const double2* cs, int nt, int use_lds) {
    extern __shared__ double2 s_cs[];
    if (!use_lds) return cs;
    for (int i = threadIdx.x; i <= 2 * nt; i += blockDim.x) s_cs[i] = cs[i];
    __syncthreads();
    return s_cs;
}

// Pair (k1, k2) of flat index e (k1 < k2, row-major over k1).
__device__ inline void pair_of(long long e, int K, int& k1, int& k2) {
    int a = (int)floor(((2.0 * K - 1.0) - sqrt((2.0 * K - 1.0) * (2.0 * K - 1.0) - 8.0 * (double)e)) / 2.0);
    if (a < 0) a = 0;
    auto first = [&](long long x) { return x * K - x * (x + 1) / 2; };
    while (a > 0 && first(a) > e) a--;
    while (first(a + 1) <= e) a++;
    k1 = a;
    k2 = (int)(e - first(a)) + a + 1;
}

// Histogram of the triangle angles and its first peak, one workgroup per
// pair; peak[e] = the peak's point, -1 when the pair has no common line or
// no triangle voted.  A kernel of its own, so that its few registers allow
// many waves per SIMD: the loop is a chain of dependent FP64 operations per
// angle, and occupancy hides its latency.
// t / 18 correctly rounded, without the general division sequence.  Let
// q = t/18 lie in [2^e, 2^(e+1)) and u = 2^(e-52) (its ulp).  R = RN(1/18)
// has |R - 1/18| <= 2^-58, so q0 = RN(t R) is within 1.6 u of q; the
// remainder r = t - 18 q0 is a multiple of u (t is a multiple of 16 u, 18 q0
// of 9 u) below 29 u in magnitude, so the FMA computes it exactly; and
// q0 + r R, rounded once by the second FMA, differs from q by
// |r| |R - 1/18| < 29 u 2^-58.  q itself lies at least u/18 from every
// rounding midpoint (t - 18 (2k+1) u/2 is an odd multiple of u), so both
// round to the same double.  t here is 0 or a normal number (a combination
// of angles in degrees), far from the subnormal range.
__device__ inline double div18(double t) {
    constexpr double R = 1.0 / 18.0;
    const double q0 = t * R;
    const double r = fma(-q0, 18.0, t);
    return fma(r, R, q0);
}

__global__ __launch_bounds__(256) void peak_kernel(int K, int n_theta, const unsigned short* M,
                                                   const unsigned short* MT, const double2* cs_g, int* peak,
                                                   int cs_lds) {
    __shared__ double s_h[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double2* cs = stage_trig(cs_g, n_theta, cs_lds);
    const long long P = (long long)K * (K - 1) / 2;
    for (long long e = blockIdx.x; e < P; e += gridDim.x) {
        int k1, k2;
        pair_of(e, K, k1, k2);
        const unsigned short* r1 = M + (size_t)k1 * K;   // clstack(k1, :)
        const unsigned short* r2 = M + (size_t)k2 * K;   // clstack(k2, :)
        const unsigned short* c1 = MT + (size_t)k1 * K;  // clstack(:, k1)
        const unsigned short* c2 = MT + (size_t)k2 * K;  // clstack(:, k2)
        const int l12 = r1[k2], l21 = r2[k1];
        if (l12 == 0) {
            if (tid == 0) peak[e] = -1;
            continue;
        }
        double h = 0.0;
        const double x = lane < NT ? tick(lane) : 0.0;
        // lane q's angle to every lane through scalar registers
        auto angle_of = [](int alo, int ahi, int q) {
            const unsigned lo = (unsigned)__builtin_amdgcn_readlane(alo, q);
            const long long hi = (long long)__builtin_amdgcn_readlane(ahi, q);
            return __longlong_as_double((hi << 32) | lo);
        };
        auto gauss = [&](double a) {
            double t = 2 * a * x;
            t = t - a * a;
            t = t - x * x;
            return exp(div18(t));
        };
        for (int base = 64 * wave; base < K; base += 256) {
            const int k3 = base + lane;
            const double ang = k3 < K ? vote_angle(r1, r2, c1, c2, k1, k3, l12, l21, cs, n_theta) : NAN;
            const long long ab = __double_as_longlong(ang);
            const int alo = (int)ab, ahi = (int)(ab >> 32);
            // the lanes with an angle (others take no part), in lane order;
            // UNR independent exponentials per step, added in that order
            constexpr int UNR = 4;
            unsigned long long m = __ballot(!isnan(ang));
            while (__popcll(m) >= UNR) {
                double e[UNR];
#pragma unroll
                for (int u = 0; u < UNR; u++) {
                    const int q = __builtin_ctzll(m);
                    m &= m - 1;
                    e[u] = gauss(angle_of(alo, ahi, q));
                }
#pragma unroll
                for (int u = 0; u < UNR; u++) h = h + e[u];
            }
            while (m) {
                const int q = __builtin_ctzll(m);
                m &= m - 1;
                h = h + gauss(angle_of(alo, ahi, q));
            }
        }
        s_h[wave][lane] = h;
        __syncthreads();
        if (tid == 0) {
            int pk = -1;
            double best = 0.0;
            for (int d = 0; d < NT; d++) {
                const double v = ((s_h[0][d] + s_h[1][d]) + s_h[2][d]) + s_h[3][d];
                if (pk < 0 || v > best) {
                    best = v;
                    pk = d;
                }
            }
            peak[e] = best > 0.0 ? pk : -1;  // no triangle voted
        }
        __syncthreads();
    }
}

// Sum over the 64 lanes in a fixed order, the same in every lane: DPP adds
// within each row of 16 lanes, then the four row sums in row order.  Every
// lane of the wave must be active.
template <int CTRL>
__device__ inline double dpp_d(double x) {
    const unsigned long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffu), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double readlane_d(double x, int l) {
    const unsigned long long u = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(u & 0xffffffffu), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double wave_sum(double x) {
    x += dpp_d<0xb1>(x);
    x += dpp_d<0x4e>(x);
    x += dpp_d<0x124>(x);
    x += dpp_d<0x128>(x);
    return ((readlane_d(x, 0) + readlane_d(x, 16)) + readlane_d(x, 32)) + readlane_d(x, 48);
}

// Third images per staged chunk of the rotation kernel (4 x 8 KB of LDS).
constexpr int RCH = 4096;

// Triangles within 6 degrees of each pair's peak: their relative rotations,
// averaged into the pair's 2x2 block of S and its transpose.  The pair's
// clstack rows and columns are staged in LDS a chunk of third images at a
// time (independent loads of the whole chunk, then LDS reads in the loop).
__global__ __launch_bounds__(256) void rot_kernel(int K, int n_theta, const unsigned short* M,
                                                  const unsigned short* MT, const double2* cs_g, const int* peak,
                                                  double* S, int cs_lds) {
    __shared__ double s_red[4][4];
    __shared__ int s_cnt[4];
    __shared__ unsigned short s_rc[4][RCH];  // r1, r2, c1, c2 of the chunk
    const int tid = threadIdx.x;
    const double2* cs = stage_trig(cs_g, n_theta, cs_lds);
    const long long P = (long long)K * (K - 1) / 2;
    const double L = (double)n_theta;
    const size_t n2 = 2 * (size_t)K;
    for (long long e = blockIdx.x; e < P; e += gridDim.x) {
        int k1, k2;
        pair_of(e, K, k1, k2);
        const unsigned short* r1 = M + (size_t)k1 * K;
        const unsigned short* r2 = M + (size_t)k2 * K;
        const unsigned short* c1 = MT + (size_t)k1 * K;
        const unsigned short* c2 = MT + (size_t)k2 * K;
        const int il12 = r1[k2], il21 = r2[k1];
        const double l12 = il12, l21 = il21;
        const int pk = peak[e];  // -1: no common line or no vote
        double r00 = 0, r01 = 0, r10 = 0, r11 = 0;
        if (pk >= 0) {
            int cnt = 0;
            const double xp = tick(pk);
            const double aa = (l12 - 1) * 2 * M_PI / L, bb = (l21 - 1) * 2 * M_PI / L;
            const double ang1 = M_PI - bb, ang3 = aa - M_PI;
            const double sa = sin(ang1), ca = cos(ang1), sc = sin(ang3), cc = cos(ang3);
            // cosines of angles certainly farther than 360/NT degrees (by
            // 1e-4 degree, far beyond the rounding of acos and cos) from the
            // peak: the triangle is out without its acos; the rest take the
            // exact test
            const double w = 360.0 / NT + 1e-4;
            const double c_hi = xp - w > 0 ? cos((xp - w) * M_PI / 180) : INFINITY;
            const double c_lo = xp + w < 180 ? cos((xp + w) * M_PI / 180) : -INFINITY;
            for (int k0 = 0; k0 < K; k0 += RCH) {
                const int nk = K - k0 < RCH ? K - k0 : RCH;
                __syncthreads();  // the previous chunk's reads are done
                for (int i = tid; i < nk; i += 256) {
                    s_rc[0][i] = r1[k0 + i];
                    s_rc[1][i] = r2[k0 + i];
                    s_rc[2][i] = c1[k0 + i];
                    s_rc[3][i] = c2[k0 + i];
                }
                __syncthreads();
                for (int i = tid; i < nk; i += 256) {
                    const int k3 = k0 + i;
                    const double cp = vote_cos(s_rc[0], s_rc[1], s_rc[2], s_rc[3], i, k1, k3, il12, il21, cs, n_theta);
                    if (isnan(cp) || cp >= c_hi || cp <= c_lo) continue;
                    const double ang = acos(cp) * 180 / M_PI;
                    if (!(fabs(ang - xp) < 360.0 / NT)) continue;
                    cnt++;
                    const int d1 = (int)s_rc[3][i] - (int)s_rc[2][i];  // cl(k3,k2) - cl(k3,k1)
                    const int d2 = (int)s_rc[1][i] - il21;             // cl(k2,k3) - cl(k2,k1)
                    const int d3 = (int)s_rc[0][i] - il12;             // cl(k1,k3) - cl(k1,k2)
                    const double idx2 = d2, idx3 = d3;
                    // cos(2 pi idx / L) of the three differences
                    const double a = trig(cs, n_theta, d1).x;
                    const double b = trig(cs, n_theta, d2).x;
                    const double c = trig(cs, n_theta, d3).x;
                    const double cond = 1 + 2 * a * b * c - (a * a + b * b + c * c);
                    if (!(cond > 1.0e-5)) continue;
                    double ca_ = (a - b * c) / sqrt(1 - b * b) / sqrt(1 - c * c);
                    const double TOL_idx = 1e-12;
                    const bool ind1 = (idx3 > L / 2 + TOL_idx) || (idx3 < -TOL_idx && idx3 > -L / 2);
                    const bool ind2 = (idx2 > L / 2 + TOL_idx) || (idx2 < -TOL_idx && idx2 > -L / 2);
                    if (ind1 != ind2) ca_ = -ca_;
                    ca_ = ca_ > 1 ? 1 : (ca_ < -1 ? -1 : ca_);
                    const double cb = cos(acos(ca_));
                    r00 = r00 + (cc * ca - sc * cb * sa);
                    r01 = r01 + (-cc * sa - sc * cb * ca);
                    r10 = r10 + (sc * ca + cc * cb * sa);
                    r11 = r11 + (-sa * sc + cc * cb * ca);
                }
            }
            // sums over the workgroup: each wave's (DPP within rows of 16,
            // then the rows in order), then the four waves in order
            const int lane = tid & 63, wave = tid >> 6;
            r00 = wave_sum(r00);
            r01 = wave_sum(r01);
            r10 = wave_sum(r10);
            r11 = wave_sum(r11);
            const double wc = wave_sum((double)cnt);
            if (lane == 0) {
                s_red[0][wave] = r00;
                s_red[1][wave] = r01;
                s_red[2][wave] = r10;
                s_red[3][wave] = r11;
                s_cnt[wave] = (int)wc;
            }
            __syncthreads();
            if (tid == 0) {
#pragma unroll
                for (int m = 0; m < 4; m++) s_red[m][0] = ((s_red[m][0] + s_red[m][1]) + s_red[m][2]) + s_red[m][3];
                s_cnt[0] = ((s_cnt[0] + s_cnt[1]) + s_cnt[2]) + s_cnt[3];
            }
            __syncthreads();
            const int n = s_cnt[0];
            if (n > 0) {
                r00 = s_red[0][0] / n;
                r01 = s_red[1][0] / n;
                r10 = s_red[2][0] / n;
                r11 = s_red[3][0] / n;
            } else {
                r00 = r01 = r10 = r11 = 0.0;
            }
            __syncthreads();
        }
        if (tid < 4) {
            const int a = tid >> 1, b = tid & 1;
            const double v = tid == 0 ? r00 : tid == 1 ? r01 : tid == 2 ? r10 : r11;
            S[(2 * (size_t)k2 + b) * n2 + 2 * (size_t)k1 + a] = v;  // S(2k1+a, 2k2+b) = R22(a,b)
            S[(2 * (size_t)k1 + a) * n2 + 2 * (size_t)k2 + b] = v;  // S(2k2+b, 2k1+a)
        }
    }
}

}  // namespace vote

using namespace clmh;

extern "C" int clm_syncmatrix_vote(clm_ctx* c, int32_t K, int32_t n_theta, const int32_t* d_cl1,
                                   const int32_t* d_cl2, double* d_S, clm_vote_stats* st) {
    using namespace vote;
    if (!c) return CLM_E_ARG;
    if (st) *st = clm_vote_stats{};
    if (K < 1 || K > 65536) return fail(c, CLM_E_ARG, "K = %d outside [1, 65536]", K);
    if (n_theta <= 0 || n_theta > 65535) return fail(c, CLM_E_ARG, "invalid n_theta %d", n_theta);
    if (!d_S || (K > 1 && (!d_cl1 || !d_cl2))) return fail(c, CLM_E_ARG, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    const size_t kk = (size_t)K * K;
    const long long P = (long long)K * (K - 1) / 2;
    if ((rc = ensure(c, c->vt_m, kk * 2)) || (rc = ensure(c, c->vt_mt, kk * 2)) ||
        (rc = ensure(c, c->vt_pk, (size_t)std::max<long long>(P, 1) * 4)) ||
        (rc = ensure(c, c->vt_cs, (size_t)(2 * n_theta + 1) * sizeof(double2))))
        return rc;
    unsigned short* M = (unsigned short*)c->vt_m.p;
    unsigned short* MT = (unsigned short*)c->vt_mt.p;
    int* pk = (int*)c->vt_pk.p;
    double2* cs = (double2*)c->vt_cs.p;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    build_kernel<<<K, 256, 0, c->stream>>>(K, d_cl1, d_cl2, M, MT);
    trig_kernel<<<(2 * n_theta + 1 + 255) / 256, 256, 0, c->stream>>>(n_theta, cs);
    diag_kernel<<<(K + 255) / 256, 256, 0, c->stream>>>(K, d_S);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    if (P > 0) {
        const long long nb = std::min<long long>(P, (long long)c->n_cu * 16);
        // the (cos, sin) table in LDS up to 16 KB (n_theta <= 511)
        const size_t cs_bytes = (size_t)(2 * n_theta + 1) * sizeof(double2);
        const int cs_lds = cs_bytes <= 16384;
        peak_kernel<<<(unsigned)nb, 256, cs_lds ? cs_bytes : 0, c->stream>>>(K, n_theta, M, MT, cs, pk, cs_lds);
        rot_kernel<<<(unsigned)nb, 256, cs_lds ? cs_bytes : 0, c->stream>>>(K, n_theta, M, MT, cs, pk, d_S, cs_lds);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st) {
        float a = 0.f, b = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        HIPCHK(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        st->ms_build = a;
        st->ms_vote = b;
        st->pairs = P;
    }
    return CLM_OK;
}
