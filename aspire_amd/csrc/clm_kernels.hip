// clm_kernels.hip -- gfx950 kernels of the common-lines engine.
//
// The hot path of ASPIRE's cryo_clmatrix_cpu (sinograms/cryo_clmatrix_cpu.m:
// 296-399) for every pair (k1,k2), every shift s_j and every pair of lines:
//   C1 = 2 Re((P1 .* ph_j)^H P2),  C2 = 2 Re((conj(P1) .* ph_j)^H P2)
// followed by a first-max argmax.  Here it is computed as
//   X_e = Re(P1.*ph_e)^T Re(P2),   Y_e = Im(P1.*ph_e)^T Im(P2)
//   C1(s_e) = 2 (X_e + Y_e),       C2(-s_e) = 2 (X_e - Y_e)
// (exact identity: conj(P1).*ph_{-s} = conj(P1.*ph_s)), so one MFMA pass over
// the evaluation shifts e yields both maps; on a symmetric shift grid the
// best of the two at every position is t = X + |Y|.
//
// Kernels:
//   prep_kernel      line assembly, filter, band zeroing, normalization
//                    (:96, :277-293, cryo_raynormalize.m:17-24) -> FP64 and
//                    FP32 half lines + MFMA operand planes.
//   pair_kernel      persistent all-pairs MFMA contraction with a fused
//                    max epilogue (pass 1: best cells; pass 2: every
//                    candidate above a certified threshold).
//   rescore_kernel   FP32 re-score of the surviving candidates, FP64
//                    re-score of near ties, reference tie order.
//   exact64_kernel   FP64 exhaustive search of one pair (fallback and the
//                    CLM_FP64_REF path); the reference's own arithmetic.
#include "clm_device.h"

namespace clm {

// ---------------------------------------------------------------- operands
template <typename OpT> struct Op;

template <> struct Op<__bf16> {
    using frag = bf16x8;
    static constexpr int KSTEP = 16;  // k per MFMA
    __device__ static inline f32x16 mma(frag a, frag b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct Op<_Float16> {
    using frag = f16x8;
    static constexpr int KSTEP = 16;
    __device__ static inline f32x16 mma(frag a, frag b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct Op<float> {
    using frag = float;
    static constexpr int KSTEP = 2;
    __device__ static inline f32x16 mma(frag a, frag b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
};

__device__ inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// -------------------------------------------------------------- preprocess
// One thread per line (k,l).  Sequential FP64 arithmetic in the oracle's
// order so P64 is bit-identical to the CPU restatement.
template <typename InT>
__global__ void prep_kernel(const InT* __restrict__ pf, Geom g,
                            const double* __restrict__ H,  // 2n_r-1, zeroed band
                            double2* __restrict__ P64,     // [K][L][R] or null
                            float2* __restrict__ P32,      // [K][L][Rp]
                            void* __restrict__ Bop,        // [K][2][Lp][Rp]
                            int op_kind,                   // 0 f32, 1 bf16, 2 f16
                            unsigned* __restrict__ nrm_bits,  // [K] float bits
                            unsigned long long* __restrict__ n_small) {
    // lines run over the padded range l < Lp so padding rows of the operand
    // planes are written (zeros) by the same launch
    const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= (long)g.K * g.Lp) return;
    const int k = (int)(line / g.Lp), l = (int)(line % g.Lp);
    const int n_r = g.n_r, R = g.R, L = g.L, T = g.T, Rp = g.Rp, Lp = g.Lp;
    auto put = [&](size_t idx, float v) {
        if (op_kind == 0) ((float*)Bop)[idx] = v;
        else if (op_kind == 1) ((__bf16*)Bop)[idx] = (__bf16)v;
        else ((_Float16*)Bop)[idx] = (_Float16)v;
    };
    const size_t pre0 = (((size_t)k * 2 + 0) * Lp + l) * Rp;
    const size_t pim0 = (((size_t)k * 2 + 1) * Lp + l) * Rp;
    if (l >= L) {
        for (int r = 0; r < Rp; r++) { put(pre0 + r, 0.f); put(pim0 + r, 0.f); }
        return;
    }
    const InT* base = pf + 2 * (size_t)k * T * n_r;
    double ss = 0.0;
    for (int i = 0; i < 2 * n_r - 1; i++) {
        int rr, tt;
        if (i < n_r - 1) { rr = n_r - 1 - i; tt = L + l; }
        else             { rr = i - (n_r - 1); tt = l; }
        const InT* z = base + 2 * ((size_t)tt * n_r + rr);
        double re = (double)z[0] * H[i], im = (double)z[1] * H[i];
        ss += re * re + im * im;
    }
    const double nr = sqrt(ss);
    if (nr < 1.0e-13) atomicAdd(n_small, 1ull);
    const InT* zc = base + 2 * ((size_t)(L + l) * n_r);
    double2* o64 = P64 ? P64 + ((size_t)k * L + l) * R : nullptr;
    float2* o32 = P32 + ((size_t)k * L + l) * Rp;
    double hn = 0.0;
    for (int r = 0; r < R; r++) {
        const int rr = n_r - 1 - r;
        double re = ((double)zc[2 * rr] * H[r]) / nr;
        double im = ((double)zc[2 * rr + 1] * H[r]) / nr;
        if (o64) o64[r] = make_double2(re, im);
        float fr = (float)re, fi = (float)im;
        o32[r] = make_float2(fr, fi);
        hn += re * re + im * im;
        put(pre0 + r, fr);
        put(pim0 + r, fi);
    }
    for (int r = R; r < Rp; r++) {
        o32[r] = make_float2(0.f, 0.f);
        put(pre0 + r, 0.f);
        put(pim0 + r, 0.f);
    }
    // max half-line norm of the image, rounded up (bound input of the
    // certified threshold).  Non-finite norms are left out: their lines
    // give NaN correlations, which every max below ignores.
    float hf = (float)(sqrt(hn) * (1.0 + 1.0e-6));
    if (isfinite(hf)) atomicMax(nrm_bits + k, __float_as_uint(hf));
}

// -------------------------------------------------------------- pair kernel
struct PairArgs {
    Geom g;
    const float2* P32;     // [K][L][Rp]
    const void* Bop;       // [K][2][Lp][Rp] operand planes
    const float2* ph32;    // [NE + 1][Rp] (cos, sin) of the evaluation shifts; row NE zero
    const float* nrm;      // [K]
    const int4* tiles;     // {k1, k2_first | list offset, count, slot | -1}
    int ntiles;
    int* counter;          // next tile (persistent work queue)
    const int4* plist;     // explicit pairs {k2, slot, u, 0}
    PairRec* recs;         // indexed by slot
    EmitRec* erecs;        // indexed by u (second pass)
    int CPP;               // waves per pair (column chunks)
    float coef;            // delta = coef * n1 * n2 + sub * (n1 + n2)
    float sub;
    float tau_t;           // tie margin in t units
    int stage;             // 1: image k1's rays and the phase table live in LDS
};

constexpr int WCAP = 64;   // per-wave list of candidate cells (one per lane)

__device__ inline unsigned long long lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Maximum over the 64 lanes with DPP row operations and 4 lane reads.
__device__ inline float wave_max_dpp(float v) {
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0xb1, 0xf, 0xf, false)));  // quad_perm [1,0,3,2]
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x4e, 0xf, 0xf, false)));  // quad_perm [2,3,0,1]
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x124, 0xf, 0xf, false))); // row_ror:4
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x128, 0xf, 0xf, false))); // row_ror:8
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// Dynamic LDS bytes of pair_kernel for a launch shape.
__host__ __device__ inline size_t pair_lds_bytes(int Rp, int osz, int W, int NE, int L, int stage) {
    size_t b = 64 * (size_t)(Rp * 2 * osz + 16) + 384 + (size_t)W * WCAP * 8;
    if (stage) b += (size_t)(NE + 1 + L) * Rp * sizeof(float2);
    return b;
}

// Persistent all-pairs kernel.  A workgroup takes a tile (one image k1 and up
// to W/CPP partner images k2), stages k1's rays in LDS, and sweeps the rows
// (e, l1) of k1's shift-expanded rays in A-tiles of 32 rows.  Each A-tile is
// generated once in LDS (rays x phase, rounded to the operand type) and
// multiplied by every wave against its k2's columns, which stay in
// registers.  The epilogue reduces each lane's 32x32 output tile to one
// "cell" maximum of t = X + |Y|.
//   EMIT = false (pass 1): each wave keeps every cell that comes within the
//     certified margin of the pair's running maximum; at the end the cells
//     within the margin of the pair's final maximum go to the record.
//   EMIT = true (pass 2): every candidate at or above the pair's threshold T
//     from pass 1 is listed.
template <typename OpT, int CT, int NK, bool SYM, bool EMIT>
__global__ __launch_bounds__(512) void pair_kernel(PairArgs a) {
    extern __shared__ __align__(16) char smem[];
    using O = Op<OpT>;
    using frag = typename O::frag;
    const Geom g = a.g;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = blockDim.x >> 6;
    const int h = lane >> 5, lc = lane & 31;
    constexpr int Rp = NK * O::KSTEP;                   // padded ray length (== g.Rp)
    constexpr int RS = Rp * 2 * (int)sizeof(OpT) + 16;  // LDS row stride (bytes)
    char* abuf0 = smem;
    char* abuf1 = smem + 32 * RS;
    int* s_i = (int*)(smem + 64 * RS);        // [0] tile, [1..] per-pair counters
    float* s_pm = (float*)(s_i + 32);         // per-wave running maxima
    int* s_unc = (int*)(s_pm + 32);           // per-pair flags
    float2* s_list = (float2*)(s_unc + 32);   // [W][WCAP] (value, id bits)
    float2* s_ph = s_list + W * WCAP;         // [NE + 1][Rp], row NE zero
    float2* s_p1 = s_ph + (size_t)(g.NE + 1) * Rp;  // [L][Rp]
    float2* wlist = s_list + wave * WCAP;
    const int pi = wave / a.CPP, chunk = wave % a.CPP;

    if (a.stage) {
        const int n4 = (g.NE + 1) * Rp / 2;
        const float4* src = (const float4*)a.ph32;
        for (int i = tid; i < n4; i += blockDim.x) ((float4*)s_ph)[i] = src[i];
    }

    for (;;) {
        if (tid == 0) s_i[0] = atomicAdd(a.counter, 1);
        if (tid >= 1 && tid < 32) s_i[tid] = 0;
        if (tid < 32) { s_unc[tid] = 0; s_pm[tid] = -INFINITY; }
        __syncthreads();
        const int t = s_i[0];
        if (t >= a.ntiles) break;
        const int4 tile = a.tiles[t];
        const int k1 = tile.x;
        const bool active = pi < tile.z;
        int k2 = 0, slot = 0, uidx = 0;
        if (active) {
            if (tile.w >= 0) { k2 = tile.y + pi; slot = tile.w + pi; }
            else { int4 pe = a.plist[tile.y + pi]; k2 = pe.x; slot = pe.y; uidx = pe.z; }
        }
        if (a.stage) {
            const int n4 = g.L * Rp / 2;
            const float4* src = (const float4*)(a.P32 + (size_t)k1 * g.L * Rp);
            for (int i = tid; i < n4; i += blockDim.x) ((float4*)s_p1)[i] = src[i];
        }
        // ---- B operand (image k2, this wave's CT column tiles) in registers
        frag Bre[CT][NK], Bim[CT][NK];
        {
            const OpT* B = (const OpT*)a.Bop;
            const size_t plane = (size_t)g.Lp * Rp;
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                const int col = (chunk * CT + ct) * 32 + lc;
                const OpT* br = B + ((size_t)k2 * 2 + 0) * plane + (size_t)col * Rp;
                const OpT* bi = br + plane;
#pragma unroll
                for (int ks = 0; ks < NK; ks++) {
                    if constexpr (O::KSTEP == 16) {
                        Bre[ct][ks] = *(const frag*)(br + 16 * ks + 8 * h);
                        Bim[ct][ks] = *(const frag*)(bi + 16 * ks + 8 * h);
                    } else {
                        Bre[ct][ks] = br[h * (Rp / 2) + ks];
                        Bim[ct][ks] = bi[h * (Rp / 2) + ks];
                    }
                }
            }
        }
        const float n1 = a.nrm[k1], n2 = a.nrm[k2];
        const float margin = 2.0f * (a.coef * n1 * n2 + a.sub * (n1 + n2)) + a.tau_t;
        float thr = -INFINITY;
        if (EMIT && active) thr = a.recs[slot].T;
        float b = -INFINITY;   // lane's running cell maximum
        float pm = -INFINITY;  // pair's running maximum (all waves, previous A-tile)
        int wc = 0;            // entries in this wave's cell list
        bool lost = false;     // a cell within the margin could not be kept

        // ---- A-tile generation: rows (e,l1) of image k1 times the shift
        // phase, rounded once to the operand type.  One item = 4 complex
        // values of one row (two float4 of rays, two of phases); consecutive
        // threads take consecutive items.  (g_e0, g_l10) is the (e, l1) of
        // row 0 of the next tile to generate, advanced by 32 rows per tile.
        constexpr int QC = Rp / 4;        // items per row
        constexpr int NC = 32 * QC;       // items per tile
        int g_e0 = 0, g_l10 = 0;
        const float2* p1_glb = a.P32 + (size_t)k1 * g.L * Rp;
        auto gen = [&](int at, char* buf) {
            for (int it = tid; it < NC; it += blockDim.x) {
                const int row = it / QC, q = it - row * QC;
                int e = g_e0, l1 = g_l10 + row;
                if (g.L >= 32) {
                    if (l1 >= g.L) { l1 -= g.L; e++; }
                } else {
                    e += l1 / g.L;
                    l1 -= (l1 / g.L) * g.L;
                }
                // rows past the last shift read the zero phase row NE, so
                // padding rows of the last tile come out as exact zeros
                const int ep = min(e, g.NE);
                float4 p0, p1v, q0, q1;
                if (a.stage) {
                    const float4* pp = (const float4*)(s_p1 + l1 * Rp + 4 * q);
                    const float4* qp = (const float4*)(s_ph + ep * Rp + 4 * q);
                    p0 = pp[0]; p1v = pp[1]; q0 = qp[0]; q1 = qp[1];
                } else {
                    const float4* pp = (const float4*)(p1_glb + (size_t)l1 * Rp + 4 * q);
                    const float4* qp = (const float4*)(a.ph32 + (size_t)ep * Rp + 4 * q);
                    p0 = pp[0]; p1v = pp[1]; q0 = qp[0]; q1 = qp[1];
                }
                float re[4], im[4];
                re[0] = fmaf(p0.x, q0.x, -p0.y * q0.y);
                im[0] = fmaf(p0.x, q0.y, p0.y * q0.x);
                re[1] = fmaf(p0.z, q0.z, -p0.w * q0.w);
                im[1] = fmaf(p0.z, q0.w, p0.w * q0.z);
                re[2] = fmaf(p1v.x, q1.x, -p1v.y * q1.y);
                im[2] = fmaf(p1v.x, q1.y, p1v.y * q1.x);
                re[3] = fmaf(p1v.z, q1.z, -p1v.w * q1.w);
                im[3] = fmaf(p1v.z, q1.w, p1v.w * q1.z);
                if (!SYM && e >= g.S) {
#pragma unroll
                    for (int u = 0; u < 4; u++) im[u] = -im[u];
                }
                OpT* dre = (OpT*)(buf + row * RS) + 4 * q;
                OpT* dim = dre + Rp;
                if constexpr (O::KSTEP == 16) {
                    typedef OpT op4 __attribute__((ext_vector_type(4)));
                    op4 vr, vi;
#pragma unroll
                    for (int u = 0; u < 4; u++) { vr[u] = (OpT)re[u]; vi[u] = (OpT)im[u]; }
                    *(op4*)dre = vr;
                    *(op4*)dim = vi;
                } else {
                    *(float4*)dre = make_float4(re[0], re[1], re[2], re[3]);
                    *(float4*)dim = make_float4(im[0], im[1], im[2], im[3]);
                }
            }
            g_l10 += 32;
            while (g_l10 >= g.L) { g_l10 -= g.L; g_e0++; }
        };

        if (a.stage) __syncthreads();  // k1's rays staged
        gen(0, abuf0);
        __syncthreads();
        for (int at = 0; at < g.NAT; at++) {
            char* cur = (at & 1) ? abuf1 : abuf0;
            char* nxt = (at & 1) ? abuf0 : abuf1;
            const char* arow = cur + lc * RS;
            float m[CT];
            f32x16 X[CT], Y[CT];
            frag fa[O::KSTEP == 16 ? NK : 1], fb[O::KSTEP == 16 ? NK : 1];
            // best of both maps over a cell: t = X + |Y|; pass 2 also lists
            // every candidate at or above the threshold
            auto epi = [&](int ct) {
                float mv = -INFINITY;
#pragma unroll
                for (int v = 0; v < 16; v++)
                    mv = fmaxf(mv, SYM ? X[ct][v] + fabsf(Y[ct][v]) : X[ct][v] + Y[ct][v]);
                m[ct] = mv;
                if constexpr (EMIT) {
                    if (mv >= thr) {
                        const int col = (chunk * CT + ct) * 32 + lc;
#pragma unroll
                        for (int v = 0; v < 16; v++) {
                            const float tv = SYM ? X[ct][v] + fabsf(Y[ct][v]) : X[ct][v] + Y[ct][v];
                            const int row = at * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
                            if (tv >= thr && row < g.NROWS && col < g.L) {
                                const int pos = atomicAdd(&s_i[1 + pi], 1);
                                if (pos < MAXE) a.erecs[uidx].cand[pos] = (uint32_t)(row * g.Lp + col);
                            }
                        }
                    }
                }
            };
            auto mma16 = [&](int ct) {
                f32x16 z = {0.f};
                X[ct] = O::mma(fa[0], Bre[ct][0], z);
                Y[ct] = O::mma(fb[0], Bim[ct][0], z);
#pragma unroll
                for (int ks = 1; ks < NK; ks++) {
                    X[ct] = O::mma(fa[ks], Bre[ct][ks], X[ct]);
                    Y[ct] = O::mma(fb[ks], Bim[ct][ks], Y[ct]);
                }
            };
            if constexpr (O::KSTEP == 16) {
                // A fragments of this tile, then the first column tile's MFMAs,
                // then the next tile's generation while they run
                if (active) {
#pragma unroll
                    for (int ks = 0; ks < NK; ks++) {
                        fa[ks] = *(const frag*)((const OpT*)arow + 16 * ks + 8 * h);
                        fb[ks] = *(const frag*)((const OpT*)arow + Rp + 16 * ks + 8 * h);
                    }
                    mma16(0);
                }
                if (at + 1 < g.NAT) gen(at + 1, nxt);
                if (active) {
#pragma unroll
                    for (int ct = 0; ct < CT; ct++) {
                        if (ct + 1 < CT) mma16(ct + 1);
                        epi(ct);
                    }
                }
            } else {
                if (at + 1 < g.NAT) gen(at + 1, nxt);
                if (active) {
                    // FP32 operands: one k-step at a time over every column tile
#pragma unroll
                    for (int ks = 0; ks < NK; ks++) {
                        const frag a0 = ((const float*)arow)[h * (Rp / 2) + ks];
                        const frag b0 = ((const float*)arow)[Rp + h * (Rp / 2) + ks];
#pragma unroll
                        for (int ct = 0; ct < CT; ct++) {
                            if (ks == 0) {
                                f32x16 z = {0.f};
                                X[ct] = O::mma(a0, Bre[ct][0], z);
                                Y[ct] = O::mma(b0, Bim[ct][0], z);
                            } else {
                                X[ct] = O::mma(a0, Bre[ct][ks], X[ct]);
                                Y[ct] = O::mma(b0, Bim[ct][ks], Y[ct]);
                            }
                        }
                    }
#pragma unroll
                    for (int ct = 0; ct < CT; ct++) epi(ct);
                }
            }
            if (active) {
                if constexpr (!EMIT) {
                    float mm = m[0];
#pragma unroll
                    for (int ct = 1; ct < CT; ct++) mm = fmaxf(mm, m[ct]);
                    b = fmaxf(b, mm);
                    // wave-uniform threshold from the running maxima seen so
                    // far (all <= the pair's final maximum)
                    const float wm = wave_max_dpp(b);
                    if (lane == 0) s_pm[wave] = wm;
                    float th = fmaxf(wm, pm) - margin;
                    if (__ballot(mm >= th && mm > -INFINITY)) {
                        // rare once the pair's maximum has been seen
#pragma unroll
                        for (int ct = 0; ct < CT; ct++) {
                            bool e = (m[ct] >= th) && (m[ct] > -INFINITY);
                            unsigned long long mask = __ballot(e);
                            if (!mask) continue;
                            int cnt = __popcll(mask);
                            if (wc + cnt > WCAP) {
                                // compact: drop entries below the current threshold
                                const float2 ent = lane < wc ? wlist[lane] : make_float2(-INFINITY, 0.f);
                                const bool keep = lane < wc && ent.x >= th;
                                const unsigned long long km = __ballot(keep);
                                __builtin_amdgcn_wave_barrier();
                                if (keep) wlist[__popcll(km & lanemask_lt())] = ent;
                                wc = __popcll(km);
                                __builtin_amdgcn_wave_barrier();
                            }
                            if (e) {
                                const int pos = wc + __popcll(mask & lanemask_lt());
                                const uint32_t id = ((uint32_t)(at * g.NCT + chunk * CT + ct) << 6) | (uint32_t)lane;
                                if (pos < WCAP) wlist[pos] = make_float2(m[ct], __uint_as_float(id));
                            }
                            if (wc + cnt > WCAP) lost = true;
                            wc = min(wc + cnt, WCAP);
                            __builtin_amdgcn_wave_barrier();
                        }
                    }
                }
            }
            __syncthreads();
            if (!EMIT && active) {
                for (int c = 0; c < a.CPP; c++) pm = fmaxf(pm, s_pm[pi * a.CPP + c]);
            }
        }

        // ---- per-pair finalisation
        if constexpr (!EMIT) {
            // s_pm holds every wave's final maximum (written before the last barrier)
            if (active) {
                float M = -INFINITY;
                for (int c = 0; c < a.CPP; c++) M = fmaxf(M, s_pm[pi * a.CPP + c]);
                const float T = M - margin;
                const float2 ent = lane < wc ? wlist[lane] : make_float2(-INFINITY, 0.f);
                if (lane < wc && ent.x >= T) {
                    const int pos = atomicAdd(&s_i[1 + pi], 1);
                    if (pos < MAXC) a.recs[slot].cell[pos] = __float_as_uint(ent.y);
                }
                if (lost) atomicOr(&s_unc[pi], 1);
            }
            __syncthreads();
            if (active && chunk == 0 && lane == 0) {
                float M = -INFINITY;
                for (int c = 0; c < a.CPP; c++) M = fmaxf(M, s_pm[pi * a.CPP + c]);
                const int n = s_i[1 + pi];
                PairRec* r = a.recs + slot;
                r->M = M;
                r->T = M - margin;
                r->k1 = k1;
                r->k2 = k2;
                r->ncells = n < MAXC ? n : MAXC;
                r->flags = (n > MAXC ? F_OVERFLOW : 0) | (s_unc[pi] ? F_UNCERTAIN : 0) |
                           (M == -INFINITY ? F_EMPTY : 0);
            }
        } else {
            __syncthreads();
            if (active && chunk == 0 && lane == 0) {
                const int n = s_i[1 + pi];
                a.erecs[uidx].n = n < MAXE ? n : MAXE;
                a.erecs[uidx].overflow = n > MAXE;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ re-scoring
struct RescoreArgs {
    Geom g;
    const float2* P32;        // [K][L][Rp]
    const double2* P64;       // [K][L][R] (null for plain FP32)
    const float2* ph32;       // [NE][Rp]
    const double2* ph64;      // [S][R] phases of the reference grid
    const double* shiftv;     // [S] s_j
    const PairRec* recs;      // pass 1 records (by slot)
    const EmitRec* erecs;     // pass 2 candidate lists (by u)
    const int* ulist;         // pass 2: u -> slot
    int n_items;              // pairs (pass 1) or uncertain pairs (pass 2)
    int mode;                 // 0: cells of pass 1, 1: emitted candidates
    int fp64;                 // re-score near ties in FP64
    float tau;                // near-tie margin (correlation units)
    int* unc_count;           // pass 1: uncertain pairs appended here
    int* unc_list;
    int* fb_count;            // pass 2: overflowing pairs appended here
    int* fb_list;
    unsigned long long* n64;  // FP64 re-scored candidates (stats)
    unsigned long long* ncell;  // cells re-scored (stats)
    unsigned long long* nemit;  // second-pass candidates (stats)
    int32_t* cl1; int32_t* cl2; int32_t* jidx; double* corr; double* shift1d;
};

// Order key of the reference's tie rule (shift index first, then the column
// of [C1 C2], then the row): smaller key wins among equal values.
__device__ inline long long tie_key(int j, int l2p, int l1, int L) {
    return ((long long)j * (2 * L) + l2p) * L + l1;
}

__device__ inline bool better(float v, long long k, float bv, long long bk) {
    return v > bv || (v == bv && k < bk);
}
__device__ inline bool better64(double v, long long k, double bv, long long bk) {
    return v > bv || (v == bv && k < bk);
}

// C(l1, l2p) at shift j in FP64, the oracle's arithmetic and order.
__device__ inline double entry64(const double2* P1, const double2* P2, const double2* ph, int R, bool flip) {
    double acc = 0.0;
    for (int r = 0; r < R; r++) {
        const double pr = P1[r].x, pi = P1[r].y, qr = ph[r].x, qi = ph[r].y;
        double xr, xi;
        if (!flip) { xr = pr * qr - pi * qi; xi = pr * qi + pi * qr; }
        else       { xr = pr * qr + pi * qi; xi = pr * qi - pi * qr; }
        const double t = xr * P2[r].x, u = xi * P2[r].y;
        acc = acc + (t + u);
    }
    return 2.0 * acc;
}

// Decode candidate q of a pair: pass-1 cells expand to 16 rows each,
// pass-2 lists hold (row, col) directly.  row = -1 marks padding.
__device__ inline void cand_rc(const RescoreArgs& a, const PairRec& rec, const EmitRec* er,
                               int q, int& row, int& col) {
    const Geom& g = a.g;
    if (a.mode == 0) {
        const uint32_t c = rec.cell[q >> 4];
        const int cl = c & 63, cid = c >> 6;
        const int at = cid / g.NCT, ct = cid % g.NCT;
        const int ri = q & 15;
        row = at * 32 + 8 * (ri >> 2) + 4 * (cl >> 5) + (ri & 3);
        col = ct * 32 + (cl & 31);
    } else {
        const uint32_t c = er->cand[q];
        row = (int)(c / g.Lp);
        col = (int)(c % g.Lp);
    }
    if (row >= g.NROWS || col >= g.L) row = -1;
}

__device__ inline double eval64(const RescoreArgs& a, int k1, int k2, long long key) {
    const Geom& g = a.g;
    const int l1 = (int)(key % g.L);
    const int l2p = (int)((key / g.L) % (2 * g.L));
    const int j = (int)(key / ((long long)g.L * 2 * g.L));
    const double2* P1 = a.P64 + ((size_t)k1 * g.L + l1) * g.R;
    const double2* P2 = a.P64 + ((size_t)k2 * g.L + (l2p % g.L)) * g.R;
    return entry64(P1, P2, a.ph64 + (size_t)j * g.R, g.R, l2p >= g.L);
}

__device__ inline void write_result(const RescoreArgs& a, int slot, double v, long long key) {
    const Geom& g = a.g;
    if (v > 0.0) {
        const int l1 = (int)(key % g.L);
        const int l2p = (int)((key / g.L) % (2 * g.L));
        const int j = (int)(key / ((long long)g.L * 2 * g.L));
        if (a.cl1) a.cl1[slot] = l1 + 1;
        if (a.cl2) a.cl2[slot] = l2p + 1;
        if (a.jidx) a.jidx[slot] = j;
        if (a.corr) a.corr[slot] = v;
        if (a.shift1d) a.shift1d[slot] = a.shiftv[j];
    } else {
        if (a.cl1) a.cl1[slot] = 0;
        if (a.cl2) a.cl2[slot] = 0;
        if (a.jidx) a.jidx[slot] = -1;
        if (a.corr) a.corr[slot] = 0.0;
        if (a.shift1d) a.shift1d[slot] = 0.0;
    }
}

constexpr int RS_MAXV = 2 * MAXE;  // FP32 values kept per pair (2 per row)

// One wave per pair.  FP32 re-score with 4 lanes per candidate row (16 rows
// per step, contiguous 16-byte loads), values kept in LDS; then FP64 re-score
// of every candidate within tau of the FP32 best, one lane per candidate with
// the oracle's sequential arithmetic.
__global__ __launch_bounds__(256) void rescore_kernel(RescoreArgs a) {
    __shared__ float s_val[4][RS_MAXV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int W = blockDim.x >> 6;
    float* sval = s_val[wave];
    const Geom& g = a.g;
    const int q = lane & 3, rl = lane >> 2;  // k-quarter, row within the step
    const int kq = g.Rp / 4;                 // multiple of 4
    for (int item = blockIdx.x * W + wave; item < a.n_items; item += gridDim.x * W) {
        const int slot = a.mode == 0 ? item : a.ulist[item];
        const PairRec rec = a.recs[slot];
        const int k1 = rec.k1, k2 = rec.k2;
        const EmitRec* er = a.mode ? a.erecs + item : nullptr;
        int nrows;
        if (a.mode == 0) {
            if (rec.flags & F_EMPTY) {
                if (lane == 0) write_result(a, slot, 0.0, 0);
                continue;
            }
            if (rec.flags & (F_UNCERTAIN | F_OVERFLOW)) {
                if (lane == 0) { const int u = atomicAdd(a.unc_count, 1); a.unc_list[u] = slot; }
                continue;
            }
            nrows = 16 * rec.ncells;
            if (lane == 0 && a.ncell) atomicAdd(a.ncell, (unsigned long long)rec.ncells);
        } else {
            if (er->overflow) {
                if (lane == 0) { const int u = atomicAdd(a.fb_count, 1); a.fb_list[u] = slot; }
                continue;
            }
            nrows = er->n;
            if (lane == 0 && a.nemit) atomicAdd(a.nemit, (unsigned long long)nrows);
        }
        // ---- FP32 re-score
        float bv = -INFINITY;
        long long bk = 0x7fffffffffffffffLL;
        for (int base = 0; base < nrows; base += 16) {
            const int r = base + rl;
            int row = -1, col = 0;
            if (r < nrows) cand_rc(a, rec, er, r, row, col);
            float X = 0.f, Y = 0.f;
            int e = 0, l1 = 0;
            if (row >= 0) {
                e = row / g.L;
                l1 = row - e * g.L;
                const float4* p = (const float4*)(a.P32 + ((size_t)k1 * g.L + l1) * g.Rp + q * kq);
                const float4* bb = (const float4*)(a.P32 + ((size_t)k2 * g.L + col) * g.Rp + q * kq);
                const float4* ph = (const float4*)(a.ph32 + (size_t)e * g.Rp + q * kq);
                for (int i = 0; i < kq / 2; i++) {
                    const float4 pv = p[i], qv = ph[i], b4 = bb[i];
                    const float ar0 = fmaf(pv.x, qv.x, -pv.y * qv.y);
                    const float ai0 = fmaf(pv.x, qv.y, pv.y * qv.x);
                    const float ar1 = fmaf(pv.z, qv.z, -pv.w * qv.w);
                    const float ai1 = fmaf(pv.z, qv.w, pv.w * qv.z);
                    X = fmaf(ar0, b4.x, X);
                    Y = fmaf(ai0, b4.y, Y);
                    X = fmaf(ar1, b4.z, X);
                    Y = fmaf(ai1, b4.w, Y);
                }
            }
            X += __shfl_xor(X, 1, 64);
            Y += __shfl_xor(Y, 1, 64);
            X += __shfl_xor(X, 2, 64);
            Y += __shfl_xor(Y, 2, 64);
            float v1 = -INFINITY, v2 = -INFINITY;
            long long q1 = 0x7fffffffffffffffLL, q2 = 0x7fffffffffffffffLL;
            if (row >= 0) {
                if (g.sym) {
                    v1 = 2.f * (X + Y); q1 = tie_key(e, col, l1, g.L);
                    v2 = 2.f * (X - Y); q2 = tie_key(g.S - 1 - e, g.L + col, l1, g.L);
                } else if (e < g.S) {
                    v1 = 2.f * (X + Y); q1 = tie_key(e, col, l1, g.L);
                } else {
                    // rows e >= S carry the phase of -s_{e-S}: C2(s_{e-S}) = 2(X - Y)
                    v2 = 2.f * (X - Y); q2 = tie_key(e - g.S, g.L + col, l1, g.L);
                }
            }
            if (q == 0 && r < nrows && 2 * r + 1 < RS_MAXV) {
                sval[2 * r] = v1;
                sval[2 * r + 1] = v2;
            }
            if (better(v1, q1, bv, bk)) { bv = v1; bk = q1; }
            if (better(v2, q2, bv, bk)) { bv = v2; bk = q2; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(bv, o, 64);
            const long long ok = __shfl_xor(bk, o, 64);
            if (better(ov, ok, bv, bk)) { bv = ov; bk = ok; }
        }
        double fv = (double)bv;
        long long fk = bk;
        if (a.fp64 && bv > -INFINITY) {
            __builtin_amdgcn_wave_barrier();
            // ---- FP64 re-score of every candidate within tau of the FP32 best
            const float lim = bv - a.tau;
            double b64 = -INFINITY;
            long long k64 = 0x7fffffffffffffffLL;
            int n64 = 0;
            const int nv = min(2 * nrows, RS_MAXV);
            for (int base = 0; base < nv; base += 64) {
                const int iv = base + lane;
                if (iv < nv && sval[iv] >= lim) {
                    int row, col;
                    cand_rc(a, rec, er, iv >> 1, row, col);
                    const int e = row / g.L, l1 = row - e * g.L;
                    long long key;
                    if (g.sym) key = (iv & 1) ? tie_key(g.S - 1 - e, g.L + col, l1, g.L) : tie_key(e, col, l1, g.L);
                    else key = e < g.S ? tie_key(e, col, l1, g.L) : tie_key(e - g.S, g.L + col, l1, g.L);
                    const double v = eval64(a, k1, k2, key);
                    n64++;
                    if (!isnan(v) && better64(v, key, b64, k64)) { b64 = v; k64 = key; }
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(b64, o, 64);
                const long long ok = __shfl_xor(k64, o, 64);
                if (better64(ov, ok, b64, k64)) { b64 = ov; k64 = ok; }
                n64 += __shfl_xor(n64, o, 64);
            }
            if (lane == 0 && a.n64) atomicAdd(a.n64, (unsigned long long)n64);
            fv = b64;
            fk = k64;
        }
        if (lane == 0) write_result(a, slot, fv, fk);
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------ FP64 exhaustive
// One workgroup per pair.  Candidates are enumerated in the reference's tie
// order q = (j*2L + l2p)*L + l1; the best (value, q) wins.
struct Exact64Args {
    Geom g;
    const double2* P64;
    const double2* ph64;   // [S][R]
    const double* shiftv;  // [S]
    const int* list;       // slots to compute (null: slot = blockIdx-strided 0..n-1)
    int n;
    int K;
    int32_t* cl1; int32_t* cl2; int32_t* jidx; double* corr; double* shift1d;
};

__device__ inline void slot_to_pair(long long s, int K, int& k1, int& k2) {
    // e(k1,k2) = k1*K - k1*(k1+1)/2 + (k2-k1-1)
    double Kd = (double)K;
    double d = (2.0 * Kd - 1.0) * (2.0 * Kd - 1.0) - 8.0 * (double)s;
    int a = (int)floor(((2.0 * Kd - 1.0) - sqrt(d > 0 ? d : 0)) / 2.0);
    if (a < 0) a = 0;
    auto first = [&](long long x) { return x * K - x * (x + 1) / 2; };
    while (a > 0 && first(a) > s) a--;
    while (first(a + 1) <= s) a++;
    k1 = a;
    k2 = (int)(s - first(a)) + a + 1;
}

__global__ __launch_bounds__(256) void exact64_kernel(Exact64Args a) {
    __shared__ double s_v[256];
    __shared__ long long s_k[256];
    const Geom g = a.g;
    const long long NQ = (long long)g.S * 2 * g.L * g.L;
    for (int it = blockIdx.x; it < a.n; it += gridDim.x) {
        const int slot = a.list ? a.list[it] : it;
        int k1, k2;
        slot_to_pair(slot, a.K, k1, k2);
        double bv = -INFINITY;
        long long bk = 0x7fffffffffffffffLL;
        for (long long q = threadIdx.x; q < NQ; q += blockDim.x) {
            const int l1 = (int)(q % g.L);
            const int l2p = (int)((q / g.L) % (2 * g.L));
            const int j = (int)(q / ((long long)2 * g.L * g.L));
            const double2* P1 = a.P64 + ((size_t)k1 * g.L + l1) * g.R;
            const double2* P2 = a.P64 + ((size_t)k2 * g.L + (l2p % g.L)) * g.R;
            const double v = entry64(P1, P2, a.ph64 + (size_t)j * g.R, g.R, l2p >= g.L);
            if (!isnan(v) && better64(v, q, bv, bk)) { bv = v; bk = q; }
        }
        s_v[threadIdx.x] = bv;
        s_k[threadIdx.x] = bk;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                if (better64(s_v[threadIdx.x + o], s_k[threadIdx.x + o], s_v[threadIdx.x], s_k[threadIdx.x])) {
                    s_v[threadIdx.x] = s_v[threadIdx.x + o];
                    s_k[threadIdx.x] = s_k[threadIdx.x + o];
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double v = s_v[0];
            const long long q = s_k[0];
            if (v > 0.0) {
                const int l1 = (int)(q % g.L);
                const int l2p = (int)((q / g.L) % (2 * g.L));
                const int j = (int)(q / ((long long)2 * g.L * g.L));
                if (a.cl1) a.cl1[slot] = l1 + 1;
                if (a.cl2) a.cl2[slot] = l2p + 1;
                if (a.jidx) a.jidx[slot] = j;
                if (a.corr) a.corr[slot] = v;
                if (a.shift1d) a.shift1d[slot] = a.shiftv[j];
            } else {
                if (a.cl1) a.cl1[slot] = 0;
                if (a.cl2) a.cl2[slot] = 0;
                if (a.jidx) a.jidx[slot] = -1;
                if (a.corr) a.corr[slot] = 0.0;
                if (a.shift1d) a.shift1d[slot] = 0.0;
            }
        }
        __syncthreads();
    }
}

__global__ void fill_i32(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace clm
