// shiftnet.hip -- the shift-proposal network of the learned shift search.
//
// The network reads the 16 lowest radial samples of all 360 rays of one
// image (32 channels: re, im per radius; positions = ray angle, circular)
// and proposes the image's in-plane 2-D shift (image axes x, y in pixels):
//   h1 = relu(conv5(x;  32 -> 32) + b1)     circular, kernel 5
//   h2 = relu(conv5(h1; 32 -> 8)  + b2)
//   z  = relu(F1 . flatten(h2) + c1)        2880 -> 16
//   y  = F2 . z + c2                        16 -> 2;  shift = 8 * y
// 52,570 parameters.  Weights are BF16 in every kernel, activations are
// stored as BF16 after the ReLU, every sum accumulates in FP32; gradients and
// the optimizer state are FP32.  Training minimizes
//   loss = 1/(2B) * sum_b sum_o (y_bo - t_bo / 8)^2.
// Kernels: features, conv1 forward, conv2 + dense forward, per-image
// backward (dense layers and conv2 data gradient), weight gradients (dense
// and conv), Adam.  The C-ABI entry points are declared in include/clm.h.
#include <hip/hip_runtime.h>

#include <cmath>

#include "clm.h"
#include "clm_ctx.h"

using namespace clmh;

namespace sn {

constexpr int NR = CLM_SN_NR;      // radial samples read
constexpr int CIN = 2 * NR;        // input channels
constexpr int T = CLM_SN_T;        // rays (circular positions)
constexpr int C1 = 32, C2 = 8, KW = 5, HALF = KW / 2, H = 16, NO = 2;
constexpr int FLAT = C2 * T;
constexpr int OW1 = 0, OB1 = OW1 + C1 * CIN * KW, OW2 = OB1 + C1, OB2 = OW2 + C2 * C1 * KW, OF1 = OB2 + C2,
              OC1 = OF1 + H * FLAT, OF2 = OC1 + H, OC2 = OF2 + NO * H, NPARAM = OC2 + NO;
static_assert(NPARAM == CLM_SN_NPARAM, "parameter count");
constexpr float SCALE = 8.0f;       // output units: pixels / SCALE

typedef __bf16 bf16;

__device__ inline float bf(bf16 v) { return (float)v; }
__device__ inline bf16 tobf(float v) { return (bf16)v; }  // round to nearest even

__device__ inline float block_sum256(float v, float* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    const float s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

// Input features: x[k][2r + {0,1}][t] = (re, im) of pf(r, t, k) / rms_k,
// r < 16, with rms_k the root mean square of those 5760 values.
template <typename InT>
__global__ __launch_bounds__(256) void features_kernel(const InT* __restrict__ pf, int n_r, bf16* __restrict__ x) {
    __shared__ float red[4];
    const int k = blockIdx.x;
    const InT* base = pf + 2 * (size_t)k * T * n_r;
    float ss = 0.f;
    for (int i = threadIdx.x; i < NR * T; i += blockDim.x) {
        const int t = i / NR, r = i % NR;
        const InT* z = base + 2 * ((size_t)t * n_r + r);
        const float re = (float)z[0], im = (float)z[1];
        ss = fmaf(re, re, fmaf(im, im, ss));
    }
    const float tot = block_sum256(ss, red);
    const float sc = 1.0f / sqrtf(tot / (float)(NR * T) + 1e-30f);
    bf16* xo = x + (size_t)k * CIN * T;
    for (int i = threadIdx.x; i < NR * T; i += blockDim.x) {
        const int t = i / NR, r = i % NR;
        const InT* z = base + 2 * ((size_t)t * n_r + r);
        xo[(2 * r) * T + t] = tobf((float)z[0] * sc);
        xo[(2 * r + 1) * T + t] = tobf((float)z[1] * sc);
    }
}

// h1 = relu(conv(x, W1) + b1), one workgroup per image.
__global__ __launch_bounds__(256) void conv1_kernel(const bf16* __restrict__ w, const bf16* __restrict__ x,
                                                    bf16* __restrict__ h1) {
    __shared__ bf16 sx[CIN * T];
    __shared__ float sw[CIN * KW][C1];  // [ci*KW + j][co]
    __shared__ float sb[C1];
    const int b = blockIdx.x, tid = threadIdx.x;
    const bf16* xb = x + (size_t)b * CIN * T;
    for (int i = tid; i < CIN * T; i += blockDim.x) sx[i] = xb[i];
    for (int i = tid; i < C1 * CIN * KW; i += blockDim.x) sw[i % (CIN * KW)][i / (CIN * KW)] = bf(w[OW1 + i]);
    if (tid < C1) sb[tid] = bf(w[OB1 + tid]);
    __syncthreads();
    for (int t = tid; t < T; t += blockDim.x) {
        int idx[KW];
#pragma unroll
        for (int j = 0; j < KW; j++) idx[j] = (t + j - HALF + T) % T;
        float acc[C1];
#pragma unroll
        for (int co = 0; co < C1; co++) acc[co] = sb[co];
        for (int ci = 0; ci < CIN; ci++) {
#pragma unroll
            for (int j = 0; j < KW; j++) {
                const float xv = bf(sx[ci * T + idx[j]]);
                const float* wr = sw[ci * KW + j];
#pragma unroll
                for (int co = 0; co < C1; co++) acc[co] = fmaf(wr[co], xv, acc[co]);
            }
        }
        bf16* hb = h1 + (size_t)b * C1 * T + t;
#pragma unroll
        for (int co = 0; co < C1; co++) hb[co * T] = tobf(fmaxf(acc[co], 0.f));
    }
}

// h2 = relu(conv(h1, W2) + b2), z = relu(F1 h2 + c1), y = F2 z + c2; one
// workgroup per image.  h2, z, y are stored when their pointers are set;
// out_px (if set) receives SCALE * y.
__global__ __launch_bounds__(256) void head_kernel(const bf16* __restrict__ w, const bf16* __restrict__ h1,
                                                   bf16* __restrict__ h2, float* __restrict__ z,
                                                   float* __restrict__ y, float* __restrict__ out_px) {
    __shared__ bf16 sh1[C1 * T];
    __shared__ float sw[C1 * KW][C2];
    __shared__ float sh2[FLAT];
    __shared__ float sred[4][H];
    __shared__ float sz[H];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bf16* hb = h1 + (size_t)b * C1 * T;
    for (int i = tid; i < C1 * T; i += blockDim.x) sh1[i] = hb[i];
    for (int i = tid; i < C2 * C1 * KW; i += blockDim.x) sw[i % (C1 * KW)][i / (C1 * KW)] = bf(w[OW2 + i]);
    __syncthreads();
    for (int t = tid; t < T; t += blockDim.x) {
        int idx[KW];
#pragma unroll
        for (int j = 0; j < KW; j++) idx[j] = (t + j - HALF + T) % T;
        float acc[C2];
#pragma unroll
        for (int co = 0; co < C2; co++) acc[co] = bf(w[OB2 + co]);
        for (int ci = 0; ci < C1; ci++) {
#pragma unroll
            for (int j = 0; j < KW; j++) {
                const float xv = bf(sh1[ci * T + idx[j]]);
#pragma unroll
                for (int co = 0; co < C2; co++) acc[co] = fmaf(sw[ci * KW + j][co], xv, acc[co]);
            }
        }
#pragma unroll
        for (int co = 0; co < C2; co++) {
            const bf16 v = tobf(fmaxf(acc[co], 0.f));
            sh2[co * T + t] = bf(v);
            if (h2) h2[(size_t)b * FLAT + co * T + t] = v;
        }
    }
    __syncthreads();
    float part[H];
#pragma unroll
    for (int h = 0; h < H; h++) part[h] = 0.f;
    for (int i = tid; i < FLAT; i += blockDim.x) {
        const float hv = sh2[i];
#pragma unroll
        for (int h = 0; h < H; h++) part[h] = fmaf(bf(w[OF1 + h * FLAT + i]), hv, part[h]);
    }
#pragma unroll
    for (int h = 0; h < H; h++) {
        float v = part[h];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) sred[wave][h] = v;
    }
    __syncthreads();
    if (tid < H) {
        const float v = bf(w[OC1 + tid]) + ((sred[0][tid] + sred[1][tid]) + (sred[2][tid] + sred[3][tid]));
        sz[tid] = fmaxf(v, 0.f);
        if (z) z[(size_t)b * H + tid] = sz[tid];
    }
    __syncthreads();
    if (tid < NO) {
        float v = bf(w[OC2 + tid]);
        for (int h = 0; h < H; h++) v = fmaf(bf(w[OF2 + tid * H + h]), sz[h], v);
        if (y) y[(size_t)b * NO + tid] = v;
        if (out_px) out_px[(size_t)b * NO + tid] = SCALE * v;
    }
}

// Per-image backward through the loss, the dense layers and conv2's input:
// dy, dz, dz2 = d loss / d (conv2 pre-activation), dz1 = d loss / d (conv1
// pre-activation).  loss accumulates 1/(2B) * |y - t/SCALE|^2.
__global__ __launch_bounds__(256) void backward_kernel(const bf16* __restrict__ w, const bf16* __restrict__ h1,
                                                       const bf16* __restrict__ h2, const float* __restrict__ z,
                                                       const float* __restrict__ y, const float* __restrict__ tgt,
                                                       int B, float* __restrict__ dy, float* __restrict__ dz,
                                                       float* __restrict__ dz2, float* __restrict__ dz1,
                                                       float* __restrict__ loss) {
    __shared__ float s_dy[NO], s_dz[H];
    __shared__ float s_dz2[FLAT];
    __shared__ float sw[C2][C1 * KW];  // W2 as [co][ci*KW + j]
    const int b = blockIdx.x, tid = threadIdx.x;
    if (tid == 0) {
        float l = 0.f;
        for (int o = 0; o < NO; o++) {
            const float d = y[(size_t)b * NO + o] - tgt[(size_t)b * NO + o] / SCALE;
            s_dy[o] = d / (float)B;
            dy[(size_t)b * NO + o] = d / (float)B;
            l = fmaf(d, d, l);
        }
        atomicAdd(loss, 0.5f * l / (float)B);
    }
    for (int i = tid; i < C2 * C1 * KW; i += blockDim.x) sw[i / (C1 * KW)][i % (C1 * KW)] = bf(w[OW2 + i]);
    __syncthreads();
    if (tid < H) {
        float g = 0.f;
        for (int o = 0; o < NO; o++) g = fmaf(bf(w[OF2 + o * H + tid]), s_dy[o], g);
        const float v = z[(size_t)b * H + tid] > 0.f ? g : 0.f;
        s_dz[tid] = v;
        dz[(size_t)b * H + tid] = v;
    }
    __syncthreads();
    for (int i = tid; i < FLAT; i += blockDim.x) {
        float g = 0.f;
#pragma unroll
        for (int h = 0; h < H; h++) g = fmaf(bf(w[OF1 + h * FLAT + i]), s_dz[h], g);
        const float v = bf(h2[(size_t)b * FLAT + i]) > 0.f ? g : 0.f;
        s_dz2[i] = v;
        dz2[(size_t)b * FLAT + i] = v;
    }
    __syncthreads();
    // conv2 data gradient: d in[ci][u] = sum_co sum_j W2[co][ci][j] dz2[co][(u - j + HALF) mod T]
    for (int i = tid; i < C1 * T; i += blockDim.x) {
        const int ci = i / T, u = i % T;
        float g = 0.f;
        for (int co = 0; co < C2; co++) {
#pragma unroll
            for (int j = 0; j < KW; j++) g = fmaf(sw[co][ci * KW + j], s_dz2[co * T + (u - j + HALF + T) % T], g);
        }
        const float v = bf(h1[(size_t)b * C1 * T + i]) > 0.f ? g : 0.f;
        dz1[(size_t)b * C1 * T + i] = v;
    }
}

// Gradients of F1, c1, F2, c2: one thread per parameter, sums over the batch.
__global__ __launch_bounds__(256) void grad_dense_kernel(const bf16* __restrict__ h2, const float* __restrict__ z,
                                                         const float* __restrict__ dy, const float* __restrict__ dz,
                                                         int B, float* __restrict__ grad) {
    const int p = OF1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0.f;
    if (p < OC1) {
        const int h = (p - OF1) / FLAT, i = (p - OF1) % FLAT;
        for (int b = 0; b < B; b++) g = fmaf(dz[(size_t)b * H + h], bf(h2[(size_t)b * FLAT + i]), g);
    } else if (p < OF2) {
        const int h = p - OC1;
        for (int b = 0; b < B; b++) g += dz[(size_t)b * H + h];
    } else if (p < OC2) {
        const int o = (p - OF2) / H, h = (p - OF2) % H;
        for (int b = 0; b < B; b++) g = fmaf(dy[(size_t)b * NO + o], z[(size_t)b * H + h], g);
    } else {
        const int o = p - OC2;
        for (int b = 0; b < B; b++) g += dy[(size_t)b * NO + o];
    }
    grad[p] = g;
}

// Gradients of a conv layer (CIN_ -> COUT): dW[co][ci][j] = sum_b sum_t
// dzo[b][co][t] in[b][ci][(t + j - HALF) mod T], db[co] = sum_b sum_t
// dzo[b][co][t].  Workgroup (co group of 8, batch chunk); partial sums are
// added to grad (which the caller cleared).
template <int CIN_, int COUT>
__global__ __launch_bounds__(256) void grad_conv_kernel(const bf16* __restrict__ in, const float* __restrict__ dzo,
                                                        int B, int ipb, int ow, int ob, float* __restrict__ grad) {
    constexpr int NW = 8 * CIN_ * KW;               // weights of the co group
    constexpr int PER = (NW + 255) / 256;
    __shared__ bf16 sin_[CIN_ * T];
    __shared__ float sdz[8 * T];
    const int cog = blockIdx.x, tid = threadIdx.x;
    const int b0 = blockIdx.y * ipb, b1 = min(B, b0 + ipb);
    float acc[PER], accb = 0.f;
    int wco[PER], wci[PER], wj[PER];
#pragma unroll
    for (int m = 0; m < PER; m++) {
        acc[m] = 0.f;
        const int wi = tid + 256 * m;
        wco[m] = wi / (CIN_ * KW);
        wci[m] = (wi % (CIN_ * KW)) / KW;
        wj[m] = wi % KW;
    }
    for (int b = b0; b < b1; b++) {
        for (int i = tid; i < CIN_ * T; i += blockDim.x) sin_[i] = in[(size_t)b * CIN_ * T + i];
        for (int i = tid; i < 8 * T; i += blockDim.x) sdz[i] = dzo[((size_t)b * COUT + cog * 8) * T + i];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < PER; m++) {
            if (tid + 256 * m < NW) {
                const float* d = sdz + wco[m] * T;
                const bf16* xi = sin_ + wci[m] * T;
                const int sh = wj[m] - HALF;
                float a = acc[m];
                for (int t = 0; t < T; t++) {
                    int u = t + sh;
                    u = u < 0 ? u + T : (u >= T ? u - T : u);
                    a = fmaf(d[t], bf(xi[u]), a);
                }
                acc[m] = a;
            }
        }
        if (tid < 8)
            for (int t = 0; t < T; t++) accb += sdz[tid * T + t];
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const int wi = tid + 256 * m;
        if (wi < NW) atomicAdd(grad + ow + (cog * 8 + wco[m]) * CIN_ * KW + wci[m] * KW + wj[m], acc[m]);
    }
    if (tid < 8) atomicAdd(grad + ob + cog * 8 + tid, accb);
}

__global__ void adam_kernel(float* __restrict__ master, bf16* __restrict__ w, float* __restrict__ m,
                            float* __restrict__ v, const float* __restrict__ g, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    const float wi = master[i] - upd;
    master[i] = wi;
    w[i] = tobf(wi);
}

__global__ void zero_f32(float* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace sn

namespace {

int check_batch(clm_ctx* c, int B) {
    if (B <= 0 || B > (1 << 20)) return fail(c, CLM_E_ARG, "batch size %d out of range", B);
    return CLM_OK;
}

// Forward pass into the context's activation buffers.
int forward(clm_ctx* c, const __bf16* w, const __bf16* x, int B, bool keep, float* out_px) {
    int rc;
    if ((rc = ensure(c, c->sn_h1, (size_t)B * sn::C1 * sn::T * 2))) return rc;
    if (keep) {
        if ((rc = ensure(c, c->sn_h2, (size_t)B * sn::FLAT * 2)) || (rc = ensure(c, c->sn_z, (size_t)B * sn::H * 4)) ||
            (rc = ensure(c, c->sn_y, (size_t)B * sn::NO * 4)))
            return rc;
    }
    hipLaunchKernelGGL(sn::conv1_kernel, dim3(B), dim3(256), 0, c->stream, w, x, (__bf16*)c->sn_h1.p);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(sn::head_kernel, dim3(B), dim3(256), 0, c->stream, w, (const __bf16*)c->sn_h1.p,
                       keep ? (__bf16*)c->sn_h2.p : nullptr, keep ? (float*)c->sn_z.p : nullptr,
                       keep ? (float*)c->sn_y.p : nullptr, out_px);
    HIPCHK(c, hipGetLastError());
    return CLM_OK;
}

}  // namespace

extern "C" {

int clm_shiftnet_nparam(void) { return sn::NPARAM; }

int clm_shiftnet_features(clm_ctx* c, const void* d_pf, int32_t n_r, int32_t n_theta, int32_t K, int32_t pf_dtype,
                          uint16_t* d_x) {
    if (!c) return CLM_E_ARG;
    if (n_theta != sn::T) return fail(c, CLM_E_UNSUP, "the shift network reads n_theta = %d rays (got %d)", sn::T, n_theta);
    if (n_r < sn::NR) return fail(c, CLM_E_UNSUP, "the shift network reads %d radial samples (n_r = %d)", sn::NR, n_r);
    if (K < 0 || (K > 0 && (!d_pf || !d_x))) return fail(c, CLM_E_ARG, "invalid features arguments");
    if (pf_dtype != CLM_C128 && pf_dtype != CLM_C64) return fail(c, CLM_E_ARG, "unknown pf dtype %d", pf_dtype);
    HIPCHK(c, hipSetDevice(c->device));
    if (K == 0) return CLM_OK;
    if (pf_dtype == CLM_C128)
        hipLaunchKernelGGL(sn::features_kernel<double>, dim3(K), dim3(256), 0, c->stream, (const double*)d_pf, n_r,
                           (__bf16*)d_x);
    else
        hipLaunchKernelGGL(sn::features_kernel<float>, dim3(K), dim3(256), 0, c->stream, (const float*)d_pf, n_r,
                           (__bf16*)d_x);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_shiftnet_forward(clm_ctx* c, const uint16_t* d_w, const uint16_t* d_x, int32_t B, float* d_out_px) {
    if (!c) return CLM_E_ARG;
    int rc;
    if ((rc = check_batch(c, B))) return rc;
    if (!d_w || !d_x || !d_out_px) return fail(c, CLM_E_ARG, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    if ((rc = forward(c, (const __bf16*)d_w, (const __bf16*)d_x, B, false, d_out_px))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_shiftnet_grad(clm_ctx* c, const uint16_t* d_w, const uint16_t* d_x, const float* d_target_px, int32_t B,
                      float* d_grad, float* d_loss) {
    if (!c) return CLM_E_ARG;
    int rc;
    if ((rc = check_batch(c, B))) return rc;
    if (!d_w || !d_x || !d_target_px || !d_grad || !d_loss) return fail(c, CLM_E_ARG, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    const __bf16* w = (const __bf16*)d_w;
    const __bf16* x = (const __bf16*)d_x;
    if ((rc = forward(c, w, x, B, true, nullptr))) return rc;
    if ((rc = ensure(c, c->sn_dy, (size_t)B * sn::NO * 4)) || (rc = ensure(c, c->sn_dz, (size_t)B * sn::H * 4)) ||
        (rc = ensure(c, c->sn_dz2, (size_t)B * sn::FLAT * 4)) || (rc = ensure(c, c->sn_dz1, (size_t)B * sn::C1 * sn::T * 4)))
        return rc;
    hipLaunchKernelGGL(sn::zero_f32, dim3((sn::NPARAM + 255) / 256), dim3(256), 0, c->stream, d_grad, sn::NPARAM);
    hipLaunchKernelGGL(sn::zero_f32, dim3(1), dim3(64), 0, c->stream, d_loss, 1);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(sn::backward_kernel, dim3(B), dim3(256), 0, c->stream, w, (const __bf16*)c->sn_h1.p,
                       (const __bf16*)c->sn_h2.p, (const float*)c->sn_z.p, (const float*)c->sn_y.p, d_target_px, B,
                       (float*)c->sn_dy.p, (float*)c->sn_dz.p, (float*)c->sn_dz2.p, (float*)c->sn_dz1.p, d_loss);
    HIPCHK(c, hipGetLastError());
    const int nd = sn::NPARAM - sn::OF1;
    hipLaunchKernelGGL(sn::grad_dense_kernel, dim3((nd + 255) / 256), dim3(256), 0, c->stream,
                       (const __bf16*)c->sn_h2.p, (const float*)c->sn_z.p, (const float*)c->sn_dy.p,
                       (const float*)c->sn_dz.p, B, d_grad);
    HIPCHK(c, hipGetLastError());
    const int ipb = 8;
    const unsigned nchunk = (unsigned)((B + ipb - 1) / ipb);
    hipLaunchKernelGGL((sn::grad_conv_kernel<sn::C1, sn::C2>), dim3(sn::C2 / 8, nchunk), dim3(256), 0, c->stream,
                       (const __bf16*)c->sn_h1.p, (const float*)c->sn_dz2.p, B, ipb, sn::OW2, sn::OB2, d_grad);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL((sn::grad_conv_kernel<sn::CIN, sn::C1>), dim3(sn::C1 / 8, nchunk), dim3(256), 0, c->stream, x,
                       (const float*)c->sn_dz1.p, B, ipb, sn::OW1, sn::OB1, d_grad);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_shiftnet_adam(clm_ctx* c, float* d_master, uint16_t* d_w, float* d_m, float* d_v, const float* d_grad,
                      int32_t step, float lr, float beta1, float beta2, float eps) {
    if (!c) return CLM_E_ARG;
    if (!d_master || !d_w || !d_m || !d_v || !d_grad || step < 1) return fail(c, CLM_E_ARG, "invalid Adam arguments");
    HIPCHK(c, hipSetDevice(c->device));
    const float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(sn::adam_kernel, dim3((sn::NPARAM + 255) / 256), dim3(256), 0, c->stream, d_master,
                       (__bf16*)d_w, d_m, d_v, d_grad, sn::NPARAM, lr, beta1, beta2, eps, c1, c2);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

}  // extern "C"
