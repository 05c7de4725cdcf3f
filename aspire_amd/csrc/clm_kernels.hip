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

// ------------------------------------------------------------- A-stacks
// Image k's shift-expanded rays ("A-stack"): rows (e, l1) of P_k .* ph_e,
// rounded once to the operand type, 32 rows per A-tile, each row stored as
// CPR 16-byte chunks (re half, then im half).  Chunk c of row r is stored at
// position (c + r) mod CPR, so the MFMA fragment reads (32 rows at the same
// logical chunk) hit distinct LDS banks without row padding.  Rows past
// the last shift are zero.  An A-tile is one contiguous block that the pair
// kernel copies to LDS with direct global-to-LDS loads.
template <typename OpT, int NK> struct ATile {
    static constexpr int Rp = NK * Op<OpT>::KSTEP;
    static constexpr int CPR = 2 * Rp * (int)sizeof(OpT) / 16;  // chunks per row
    static constexpr int EPC = 16 / (int)sizeof(OpT);            // elements per chunk
    static constexpr int NCH = 32 * CPR;                         // chunks per tile
    static constexpr int BYTES = NCH * 16;
    __device__ static inline int phys(int row, int c) { return (c + row) % CPR; }
};

template <typename OpT, int NK, bool SYM>
__global__ void astack_kernel(Geom g, const float2* __restrict__ P32, const float2* __restrict__ ph32,
                              int4* __restrict__ astk) {
    using AT = ATile<OpT, NK>;
    constexpr int Rp = AT::Rp, CPR = AT::CPR, EPC = AT::EPC;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)g.K * g.NAT * AT::NCH) return;
    const int pch = (int)(q % CPR);
    long r = q / CPR;
    const int row = (int)(r % 32);
    r /= 32;
    const int at = (int)(r % g.NAT);
    const int k = (int)(r / g.NAT);
    const int c = (pch - row % CPR + CPR) % CPR;   // logical chunk stored here
    const bool imag = c >= CPR / 2;
    const int k0 = (c % (CPR / 2)) * EPC;
    const int gr = at * 32 + row;
    OpT v[EPC];
    if (gr < g.NROWS) {
        const int e = gr / g.L, l1 = gr - e * g.L;
        const float sg = (!SYM && e >= g.S) ? -1.0f : 1.0f;
        const float2* p = P32 + ((size_t)k * g.L + l1) * Rp + k0;
        const float2* ph = ph32 + (size_t)e * Rp + k0;
#pragma unroll
        for (int u = 0; u < EPC; u++) {
            const float2 pv = p[u], qv = ph[u];
            const float re = fmaf(pv.x, qv.x, -pv.y * qv.y);
            const float im = fmaf(pv.x, qv.y, pv.y * qv.x);
            v[u] = (OpT)(imag ? sg * im : re);
        }
    } else {
#pragma unroll
        for (int u = 0; u < EPC; u++) v[u] = (OpT)0.0f;
    }
    int4 out;
    __builtin_memcpy(&out, v, 16);
    astk[q] = out;
}

// -------------------------------------------------------------- pair kernel
struct PairArgs {
    const StackDev* stacks;
    const int4* tiles;     // {k1, k2_first | list offset, count | stack << 4, slot | -1}
    int ntiles;            // tile count, or
    const int* ntiles_dev; // (when set) tile count in device memory
    int* counter;          // next tile (persistent work queue)
    const int4* plist;     // explicit pairs {k2, slot, u, 0}
    EmitRec* erecs;        // indexed by u (second pass)
    int CPP;               // waves per pair (column chunks); same for all stacks
    int lag;               // 1: one wave per SIMD runs one A-tile behind
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

// Dynamic LDS bytes of pair_kernel (the A-tile ring is static).
__host__ __device__ inline size_t pair_lds_bytes(int W) { return 384 + (size_t)W * WCAP * 8; }

// Persistent all-pairs kernel.  A workgroup takes a tile (one image k1 and up
// to W/CPP partner images k2).  k1's A-tiles (32 rows (e, l1) of its
// shift-expanded rays each, precomputed by astack_kernel) stream through two
// LDS buffers, prefetched one tile ahead through registers; each wave
// multiplies every A-tile against its k2's columns, which stay in
// registers.  The epilogue reduces each lane's 32x32 output tile to
// one "cell" maximum of t = X + |Y|.
//   EMIT = false (pass 1): each wave keeps every cell that comes within the
//     certified margin of the pair's running maximum; at the end the cells
//     within the margin of the pair's final maximum go to the record.
//   EMIT = true (pass 2): every candidate at or above the pair's threshold T
//     from pass 1 is listed.
template <typename OpT, int CT, int NK, bool SYM, bool EMIT>
__global__ __launch_bounds__(512) void pair_kernel(PairArgs a) {
    using O = Op<OpT>;
    using frag = typename O::frag;
    using AT = ATile<OpT, NK>;
    constexpr int Rp = AT::Rp, CPR = AT::CPR, NCH = AT::NCH;
    constexpr int NPT = (NCH + 255) / 256;    // chunks per thread (W >= 4)
    __shared__ __align__(16) int4 sa0[NCH];
    __shared__ __align__(16) int4 sa1[NCH];
    extern __shared__ __align__(16) char smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = blockDim.x >> 6;
    const int h = lane >> 5, lc = lane & 31;
    int* s_i = (int*)smem;                    // [0] tile, [1..] per-pair counters
    float* s_pm = (float*)(s_i + 32);         // per-wave running maxima
    int* s_unc = (int*)(s_pm + 32);           // per-pair flags
    float2* s_list = (float2*)(s_unc + 32);   // [W][WCAP] (value, id bits)
    float2* wlist = s_list + wave * WCAP;
    const int pi = wave / a.CPP, chunk = wave % a.CPP;
    const int wv = __builtin_amdgcn_readfirstlane(wave);

    // Which waves lag: the second wave (by index) on each SIMD, from the SIMD
    // id in HW_REG_HW_ID (bits 5:4), shared through LDS.
    bool lag_wave = false;
    if (a.lag) {
        const unsigned hwid = __builtin_amdgcn_s_getreg((31 << 11) | (0 << 6) | 4);
        if (lane == 0) s_unc[wave] = (int)((hwid >> 4) & 3);
        __syncthreads();
        const int my = s_unc[wv];
        for (int w = 0; w < wv; w++)
            if (s_unc[w] == my) lag_wave = true;
        __syncthreads();
    }
    const bool lag = a.lag && lag_wave;
    const int ntiles = a.ntiles_dev ? *a.ntiles_dev : a.ntiles;

    for (;;) {
        if (tid == 0) s_i[0] = atomicAdd(a.counter, 1);
        if (tid >= 1 && tid < 32) s_i[tid] = 0;
        if (tid < 32) { s_unc[tid] = 0; s_pm[tid] = -INFINITY; }
        __syncthreads();
        const int t = s_i[0];
        if (t >= ntiles) break;
        const int4 tile = a.tiles[t];
        const int k1 = tile.x;
        const int cnt_t = tile.z & 15;
        const StackDev& sd = a.stacks[tile.z >> 4];
        const Geom g = sd.g;
        const bool active = pi < cnt_t;
        int k2 = 0, slot = 0, uidx = 0;
        if (active) {
            if (tile.w >= 0) { k2 = tile.y + pi; slot = tile.w + pi; }
            else { int4 pe = a.plist[tile.y + pi]; k2 = pe.x; slot = pe.y; uidx = pe.z; }
        }
        // A-tiles of image k1 stream through two LDS buffers: each thread
        // loads its 16-byte chunks of tile at+1 into registers one step
        // ahead and writes them to LDS at the start of the next step.
        const int4* asrc = (const int4*)((const char*)sd.astk + (size_t)k1 * g.NAT * AT::BYTES);
        int4 pre[NPT];
        auto load_pre = [&](int at) {
            const int4* src = asrc + (size_t)at * NCH;
#pragma unroll
            for (int j = 0; j < NPT; j++) {
                const int q = tid + j * (int)blockDim.x;
                if (q < NCH) pre[j] = src[q];
            }
        };
        auto store_pre = [&](int4* buf) {
#pragma unroll
            for (int j = 0; j < NPT; j++) {
                const int q = tid + j * (int)blockDim.x;
                if (q < NCH) buf[q] = pre[j];
            }
        };
        load_pre(0);
        // ---- B operand (image k2, this wave's CT column tiles) in registers
        frag Bre[CT][NK], Bim[CT][NK];
        {
            const OpT* B = (const OpT*)sd.Bop;
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
        const float n1 = sd.nrm[k1], n2 = sd.nrm[k2];
        const float margin = 2.0f * (sd.coef * n1 * n2 + sd.sub * (n1 + n2)) + sd.tau_t;
        float thr = -INFINITY;
        if (EMIT && active) thr = sd.recs[slot].T;
        float b = -INFINITY;   // lane's running cell maximum
        float pm = -INFINITY;  // pair's running maximum (all waves, previous A-tile)
        int wc = 0;            // entries in this wave's cell list
        bool lost = false;     // a cell within the margin could not be kept

        float m[CT];
        f32x16 X[CT], Y[CT];
        // best of both maps over a cell: t = X + |Y|; pass 2 also lists
        // every candidate of tile tat at or above the threshold
        auto epi = [&](int ct, int tat) {
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
                        const int row = tat * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
                        if (tv >= thr && row < g.NROWS && col < g.L) {
                            const int pos = atomicAdd(&s_i[1 + pi], 1);
                            if (pos < MAXE) a.erecs[uidx].cand[pos] = (uint32_t)(row * g.Lp + col);
                        }
                    }
                }
            }
        };
        // pass 1: keep the cells of tile tat that come within the margin
        // of the pair's running maximum; publish this wave's maximum
        auto keep_cells = [&](int tat) {
            if constexpr (!EMIT) {
                float mm = m[0];
#pragma unroll
                for (int ct = 1; ct < CT; ct++) mm = fmaxf(mm, m[ct]);
                b = fmaxf(b, mm);
                // wave-uniform threshold from the running maxima seen so
                // far (all <= the pair's final maximum)
                const float wm = wave_max_dpp(b);
                if (lane == 0) s_pm[wave] = wm;
                const float th = fmaxf(wm, pm) - margin;
                if (__ballot(mm >= th && mm > -INFINITY)) {
                    // rare once the pair's maximum has been seen
#pragma unroll
                    for (int ct = 0; ct < CT; ct++) {
                        const bool e = (m[ct] >= th) && (m[ct] > -INFINITY);
                        const unsigned long long mask = __ballot(e);
                        if (!mask) continue;
                        const int cnt = __popcll(mask);
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
                            const uint32_t id = ((uint32_t)(tat * g.NCT + chunk * CT + ct) << 6) | (uint32_t)lane;
                            if (pos < WCAP) wlist[pos] = make_float2(m[ct], __uint_as_float(id));
                        }
                        if (wc + cnt > WCAP) lost = true;
                        wc = min(wc + cnt, WCAP);
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
        };
        // one A-tile: write tile at+1 (loaded during the previous step) to
        // the other buffer and start loading tile at+2; reduce (lagging
        // waves: the previous tile), multiply, reduce (leading waves: this
        // tile)
        auto step = [&](int at, const int4* cur, int4* nxt) {
            if (at + 1 < g.NAT) {
                store_pre(nxt);
                if (at + 2 < g.NAT) load_pre(at + 2);
            }
            if (lag && active && at > 0) {
#pragma unroll
                for (int ct = 0; ct < CT; ct++) epi(ct, at - 1);
                keep_cells(at - 1);
            }
            if (active) {
                const char* arow = (const char*)cur + lc * CPR * 16;
#pragma unroll
                for (int ks = 0; ks < NK; ks++) {
                    frag a0, b0;
                    if constexpr (O::KSTEP == 16) {
                        a0 = *(const frag*)(arow + AT::phys(lc, 2 * ks + h) * 16);
                        b0 = *(const frag*)(arow + AT::phys(lc, CPR / 2 + 2 * ks + h) * 16);
                    } else {
                        const int kk = h * (Rp / 2) + ks;
                        a0 = *(const float*)(arow + AT::phys(lc, kk / 4) * 16 + (kk % 4) * 4);
                        b0 = *(const float*)(arow + AT::phys(lc, CPR / 2 + kk / 4) * 16 + (kk % 4) * 4);
                    }
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
            }
            if (!lag && active) {
#pragma unroll
                for (int ct = 0; ct < CT; ct++) epi(ct, at);
                keep_cells(at);
            }
            __syncthreads();
            if (!EMIT && active) {
                for (int c = 0; c < a.CPP; c++) pm = fmaxf(pm, s_pm[pi * a.CPP + c]);
            }
        };

        store_pre(sa0);
        if (g.NAT > 1) load_pre(1);
        __syncthreads();
        for (int at = 0; at < g.NAT; at += 2) {
            step(at, sa0, sa1);
            if (at + 1 < g.NAT) step(at + 1, sa1, sa0);
        }
        if (lag && active) {
#pragma unroll
            for (int ct = 0; ct < CT; ct++) epi(ct, g.NAT - 1);
            keep_cells(g.NAT - 1);
        }
        __syncthreads();

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
                    if (pos < MAXC) sd.recs[slot].cell[pos] = __float_as_uint(ent.y);
                }
                if (lost) atomicOr(&s_unc[pi], 1);
            }
            __syncthreads();
            if (active && chunk == 0 && lane == 0) {
                float M = -INFINITY;
                for (int c = 0; c < a.CPP; c++) M = fmaxf(M, s_pm[pi * a.CPP + c]);
                const int n = s_i[1 + pi];
                PairRec* r = sd.recs + slot;
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
    const StackDev* stacks;   // the whole table (SlotRef.stack indexes it)
    int stack0, nstacks;      // mode 0: stacks [stack0, stack0 + nstacks)
    int mode;                 // 0: every pair of those stacks (pass 1), 1: pass-2 list
    long long n_items;        // mode 0: total pairs of all stacks
    const int* n_items_dev;   // mode 1: number of list entries (device)
    int item_cap;             // mode 1: entries beyond this were not given a pass 2
    const SlotRef* ulist;     // mode 1: u -> (stack, slot)
    const EmitRec* erecs;     // mode 1: candidates of entry u
    int* unc_count;           // mode 0: uncertain pairs appended here
    SlotRef* unc_list;
    int* fb_count;            // both: pairs for the FP64 exhaustive search
    SlotRef* fb_list;
    unsigned long long* n64;  // FP64 re-scored candidates (stats)
    unsigned long long* ncell;  // cells re-scored (stats)
    unsigned long long* nemit;  // second-pass candidates (stats)
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

// Stack of flat pair index `item` (stacks sorted by pair_base).
__device__ inline int find_stack(const StackDev* st, int n, long long item) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (st[mid].pair_base <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Decode candidate q of a pair: pass-1 cells expand to 16 rows each,
// pass-2 lists hold (row, col) directly.  row = -1 marks padding.
__device__ inline void cand_rc(const Geom& g, int mode, const PairRec& rec, const EmitRec* er,
                               int q, int& row, int& col) {
    if (mode == 0) {
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

__device__ inline double eval64(const StackDev& sd, int k1, int k2, long long key) {
    const Geom& g = sd.g;
    const int l1 = (int)(key % g.L);
    const int l2p = (int)((key / g.L) % (2 * g.L));
    const int j = (int)(key / ((long long)g.L * 2 * g.L));
    const double2* P1 = sd.P64 + ((size_t)k1 * g.L + l1) * g.R;
    const double2* P2 = sd.P64 + ((size_t)k2 * g.L + (l2p % g.L)) * g.R;
    return entry64(P1, P2, sd.ph64 + (size_t)j * g.R, g.R, l2p >= g.L);
}

__device__ inline void write_result(const StackDev& sd, int slot, double v, long long key) {
    const Geom& g = sd.g;
    if (v > 0.0) {
        const int l1 = (int)(key % g.L);
        const int l2p = (int)((key / g.L) % (2 * g.L));
        const int j = (int)(key / ((long long)g.L * 2 * g.L));
        if (sd.cl1) sd.cl1[slot] = l1 + 1;
        if (sd.cl2) sd.cl2[slot] = l2p + 1;
        if (sd.jidx) sd.jidx[slot] = j;
        if (sd.corr) sd.corr[slot] = v;
        if (sd.shift1d) sd.shift1d[slot] = sd.shiftv[j];
    } else {
        if (sd.cl1) sd.cl1[slot] = 0;
        if (sd.cl2) sd.cl2[slot] = 0;
        if (sd.jidx) sd.jidx[slot] = -1;
        if (sd.corr) sd.corr[slot] = 0.0;
        if (sd.shift1d) sd.shift1d[slot] = 0.0;
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
    const int q = lane & 3, rl = lane >> 2;  // k-quarter, row within the step
    const long long n_items = a.mode == 0 ? a.n_items : (long long)min(*a.n_items_dev, a.item_cap);
    for (long long item = (long long)blockIdx.x * W + wave; item < n_items; item += (long long)gridDim.x * W) {
        int stk, slot;
        if (a.mode == 0) {
            stk = a.stack0 + (a.nstacks > 1 ? find_stack(a.stacks + a.stack0, a.nstacks, item) : 0);
            slot = (int)(item - a.stacks[stk].pair_base);
        } else {
            stk = a.ulist[item].stack;
            slot = a.ulist[item].slot;
        }
        const StackDev& sd = a.stacks[stk];
        const Geom& g = sd.g;
        const int kq = g.Rp / 4;  // multiple of 4
        const PairRec rec = sd.recs[slot];
        const int k1 = rec.k1, k2 = rec.k2;
        const EmitRec* er = a.mode ? a.erecs + item : nullptr;
        int nrows;
        if (a.mode == 0) {
            if (rec.flags & F_EMPTY) {
                if (lane == 0) write_result(sd, slot, 0.0, 0);
                continue;
            }
            if (rec.flags & (F_UNCERTAIN | F_OVERFLOW)) {
                if (lane == 0) { const int u = atomicAdd(a.unc_count, 1); a.unc_list[u] = SlotRef{stk, slot}; }
                continue;
            }
            nrows = 16 * rec.ncells;
            if (lane == 0 && a.ncell) atomicAdd(a.ncell, (unsigned long long)rec.ncells);
        } else {
            if (er->overflow) {
                if (lane == 0) { const int u = atomicAdd(a.fb_count, 1); a.fb_list[u] = SlotRef{stk, slot}; }
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
            if (r < nrows) cand_rc(g, a.mode, rec, er, r, row, col);
            float X = 0.f, Y = 0.f;
            int e = 0, l1 = 0;
            if (row >= 0) {
                e = row / g.L;
                l1 = row - e * g.L;
                const float4* p = (const float4*)(sd.P32 + ((size_t)k1 * g.L + l1) * g.Rp + q * kq);
                const float4* bb = (const float4*)(sd.P32 + ((size_t)k2 * g.L + col) * g.Rp + q * kq);
                const float4* ph = (const float4*)(sd.ph32 + (size_t)e * g.Rp + q * kq);
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
        if (sd.fp64 && bv > -INFINITY) {
            __builtin_amdgcn_wave_barrier();
            // ---- FP64 re-score of every candidate within tau of the FP32 best
            const float lim = bv - sd.tau;
            double b64 = -INFINITY;
            long long k64 = 0x7fffffffffffffffLL;
            int n64 = 0;
            const int nv = min(2 * nrows, RS_MAXV);
            for (int base = 0; base < nv; base += 64) {
                const int iv = base + lane;
                if (iv < nv && sval[iv] >= lim) {
                    int row, col;
                    cand_rc(g, a.mode, rec, er, iv >> 1, row, col);
                    const int e = row / g.L, l1 = row - e * g.L;
                    long long key;
                    if (g.sym) key = (iv & 1) ? tie_key(g.S - 1 - e, g.L + col, l1, g.L) : tie_key(e, col, l1, g.L);
                    else key = e < g.S ? tie_key(e, col, l1, g.L) : tie_key(e - g.S, g.L + col, l1, g.L);
                    const double v = eval64(sd, k1, k2, key);
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
        if (lane == 0) write_result(sd, slot, fv, fk);
        __builtin_amdgcn_wave_barrier();
    }
}

// Pass-2 tiles from the uncertain list, on the device: one pair per tile
// (entry u -> tile u and explicit pair u).  Entries past `cap` have no
// pass-2 buffers and go straight to the FP64 exhaustive list.
__global__ void build_pass2_kernel(const StackDev* stacks, const int* n_unc, const SlotRef* ulist, int cap,
                                   int4* tiles, int4* plist, int* ntiles, int* fb_count, SlotRef* fb_list) {
    const int n = *n_unc;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *ntiles = min(n, cap);
    for (int u = i; u < n; u += gridDim.x * blockDim.x) {
        const SlotRef r = ulist[u];
        if (u < cap) {
            const PairRec& rec = stacks[r.stack].recs[r.slot];
            tiles[u] = make_int4(rec.k1, u, 1 | (r.stack << 4), -1);
            plist[u] = make_int4(rec.k2, r.slot, u, 0);
        } else {
            const int f = atomicAdd(fb_count, 1);
            fb_list[f] = r;
        }
    }
}

// ------------------------------------------------------ FP64 exhaustive
// One workgroup per pair.  Candidates are enumerated in the reference's tie
// order q = (j*2L + l2p)*L + l1; the best (value, q) wins.
struct Exact64Args {
    const StackDev* stacks;   // the whole table (SlotRef.stack indexes it)
    int stack0, nstacks;      // without a list: stacks [stack0, stack0 + nstacks)
    const SlotRef* list;      // pairs to compute, or null: every pair of every stack
    const int* n_dev;         // list length (device)
    long long n_all;          // total pairs when list is null
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
    const long long n = a.list ? (long long)*a.n_dev : a.n_all;
    for (long long it = blockIdx.x; it < n; it += gridDim.x) {
        int stk, slot;
        if (a.list) { stk = a.list[it].stack; slot = a.list[it].slot; }
        else {
            stk = a.stack0 + (a.nstacks > 1 ? find_stack(a.stacks + a.stack0, a.nstacks, it) : 0);
            slot = (int)(it - a.stacks[stk].pair_base);
        }
        const StackDev& sd = a.stacks[stk];
        const Geom& g = sd.g;
        const long long NQ = (long long)g.S * 2 * g.L * g.L;
        int k1, k2;
        slot_to_pair(slot, g.K, k1, k2);
        double bv = -INFINITY;
        long long bk = 0x7fffffffffffffffLL;
        for (long long qq = threadIdx.x; qq < NQ; qq += blockDim.x) {
            const int l1 = (int)(qq % g.L);
            const int l2p = (int)((qq / g.L) % (2 * g.L));
            const int j = (int)(qq / ((long long)2 * g.L * g.L));
            const double2* P1 = sd.P64 + ((size_t)k1 * g.L + l1) * g.R;
            const double2* P2 = sd.P64 + ((size_t)k2 * g.L + (l2p % g.L)) * g.R;
            const double v = entry64(P1, P2, sd.ph64 + (size_t)j * g.R, g.R, l2p >= g.L);
            if (!isnan(v) && better64(v, qq, bv, bk)) { bv = v; bk = qq; }
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
        if (threadIdx.x == 0) write_result(sd, slot, s_v[0], s_k[0]);
        __syncthreads();
    }
}

__global__ void fill_i32(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace clm
