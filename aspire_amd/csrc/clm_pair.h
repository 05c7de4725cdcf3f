// clm_pair.h -- operand traits, A-stack and pair-kernel templates of the
// common-lines engine.  Instantiated per operand type and shift-grid symmetry
// in pair_variants.hip (one object per combination, compiled in parallel);
// the engine reaches them through the pick_* functions declared at the end.
#pragma once
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

// ------------------------------------------------------------- A-stacks
// Image k's shift-expanded rays ("A-stack"): rows (e, l1) of P_k .* ph_e,
// rounded once to the operand type, 32 rows per A-tile, each row stored as
// CPR 16-byte chunks (re half, then im half).  Chunk c of row r is stored at
// position (c + r) mod CPR, so the MFMA fragment reads (32 rows at the same
// logical chunk) hit distinct LDS banks without row padding.  Rows past
// the last shift are zero.  An A-tile is one contiguous block in this layout,
// so the pair kernel copies it to LDS chunk for chunk.
template <typename OpT, int NK> struct ATile {
    static constexpr int Rp = NK * Op<OpT>::KSTEP;
    static constexpr int CPR = 2 * Rp * (int)sizeof(OpT) / 16;  // chunks per row
    static constexpr int EPC = 16 / (int)sizeof(OpT);            // elements per chunk
    static constexpr int NCH = 32 * CPR;                         // chunks per tile
    static constexpr int BYTES = NCH * 16;
    __device__ static inline int phys(int row, int c) { return (c + row) % CPR; }
};

// imgs: null (A-stacks of images 0..nimg-1 in order) or the nimg images
// whose A-stacks are built, in this order.
template <typename OpT, int NK, bool SYM>
__global__ void astack_kernel(Geom g, const float2* __restrict__ P32, const float2* __restrict__ ph32,
                              int4* __restrict__ astk, const int* __restrict__ imgs, int nimg) {
    using AT = ATile<OpT, NK>;
    constexpr int Rp = AT::Rp, CPR = AT::CPR, EPC = AT::EPC;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)nimg * g.NAT * AT::NCH) return;
    const int pch = (int)(q % CPR);
    long r = q / CPR;
    const int row = (int)(r % 32);
    r /= 32;
    const int at = (int)(r % g.NAT);
    const int k = imgs ? imgs[r / g.NAT] : (int)(r / g.NAT);
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
    int* counter;          // next tile (persistent work queue); per queue when nq > 1
    const int* qbeg;       // nq > 1: queue q serves tiles [qbeg[q], qbeg[q+1])
    int nq;                // work queues: 1, or one per XCD (workgroup b starts on queue b % nq)
    const int4* plist;     // explicit pairs {k2, slot, u, 0}
    EmitRec* erecs;        // indexed by u (second pass)
    int CPP;               // waves per pair (column chunks); same for all stacks
    int lag;               // 1: one wave per SIMD runs one A-tile behind
    int prio;              // 1: raised wave priority while issuing MFMAs (the other
                           // wave on the SIMD runs its epilogue meanwhile)
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
    // XCD-aware queues: dispatch places workgroup b on XCD b % 8, and queue
    // q holds a contiguous run of k1-major tiles, so the tiles of one image
    // k1 (which all stream its A-stack) meet in one XCD's L2.  A workgroup
    // drains its own XCD's queue, then helps the next ones.
    int cq = a.nq > 1 ? (int)(blockIdx.x % (unsigned)a.nq) : 0;
    int left = a.nq > 1 ? a.nq : 1;

    for (;;) {
        if (tid == 0) {
            int t = ntiles;
            if (a.nq > 1) {
                while (left > 0) {
                    const int u = a.qbeg[cq] + atomicAdd(a.counter + cq, 1);
                    if (u < a.qbeg[cq + 1]) {
                        t = u;
                        break;
                    }
                    cq = cq + 1 == a.nq ? 0 : cq + 1;
                    left--;
                }
            } else {
                t = atomicAdd(a.counter, 1);
            }
            s_i[0] = t;
        }
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
        const int ka = sd.amap ? sd.amap[k1] : k1;
        const int4* asrc = (const int4*)((const char*)sd.astk + (size_t)ka * g.NAT * AT::BYTES);
        // native vectors: copies of HIP's int4 wrapper kept this buffer in
        // private memory
        typedef int v4i __attribute__((ext_vector_type(4)));
        v4i pre[NPT];
        // through a global-address-space pointer: flat loads would also
        // count against the LDS wait counter, so every wait for an A
        // fragment would wait for this prefetch as well
        typedef __attribute__((address_space(1))) const v4i gv4i;
        auto load_pre = [&](int at) {
            gv4i* src = (gv4i*)(asrc + (size_t)at * NCH);
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
                if (q < NCH) ((v4i*)buf)[q] = pre[j];
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
        // A fragments of k-step ks (real and imaginary halves of the row)
        auto frag_re = [&](const int4* buf, int ks) -> frag {
            const char* arow = (const char*)buf + lc * CPR * 16;
            if constexpr (O::KSTEP == 16) {
                return *(const frag*)(arow + AT::phys(lc, 2 * ks + h) * 16);
            } else {
                const int kk = h * (Rp / 2) + ks;
                return *(const float*)(arow + AT::phys(lc, kk / 4) * 16 + (kk % 4) * 4);
            }
        };
        auto frag_im = [&](const int4* buf, int ks) -> frag {
            const char* arow = (const char*)buf + lc * CPR * 16;
            if constexpr (O::KSTEP == 16) {
                return *(const frag*)(arow + AT::phys(lc, CPR / 2 + 2 * ks + h) * 16);
            } else {
                const int kk = h * (Rp / 2) + ks;
                return *(const float*)(arow + AT::phys(lc, CPR / 2 + kk / 4) * 16 + (kk % 4) * 4);
            }
        };
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
            if (a.prio) __builtin_amdgcn_s_setprio(1);
            if (active) {
                // fragments of k-steps 0 and 1 first; each fragment register
                // is refilled two k-steps ahead as soon as the MFMAs that
                // read it have issued, so a refill has CT + 2 CT MFMAs to
                // arrive (the schedule is pinned below)
                frag fa[2], fb[2];
                fa[0] = frag_re(cur, 0);
                fb[0] = frag_im(cur, 0);
                if (NK > 1) {
                    fa[1] = frag_re(cur, 1);
                    fb[1] = frag_im(cur, 1);
                }
#pragma unroll
                for (int ks = 0; ks < NK; ks++) {
                    const int s = ks & 1;
#pragma unroll
                    for (int ct = 0; ct < CT; ct++) {
                        f32x16 z = {0.f};
                        X[ct] = O::mma(fa[s], Bre[ct][ks], ks == 0 ? z : X[ct]);
                    }
#pragma unroll
                    for (int ct = 0; ct < CT; ct++) {
                        f32x16 z = {0.f};
                        Y[ct] = O::mma(fb[s], Bim[ct][ks], ks == 0 ? z : Y[ct]);
                    }
                    if (ks + 2 < NK) {
                        fa[s] = frag_re(cur, ks + 2);
                        fb[s] = frag_im(cur, ks + 2);
                    }
                }
                if constexpr (O::KSTEP == 16) {
                    // scheduling groups (0x100: LDS reads, 0x008: MFMAs)
                    __builtin_amdgcn_sched_group_barrier(0x100, NK > 1 ? 4 : 2, 0);
#pragma unroll
                    for (int ks = 0; ks < NK; ks++) {
                        __builtin_amdgcn_sched_group_barrier(0x008, 2 * CT, 0);
                        if (ks + 2 < NK) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
                    }
                }
            }
            if (a.prio) __builtin_amdgcn_s_setprio(0);
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

using PairFn = void (*)(PairArgs);
using AstkFn = void (*)(Geom, const float2*, const float2*, int4*, const int*, int);

// Instantiations, one set per (operand type, grid symmetry): ct = column
// tiles per wave, NK = k-steps of the padded ray (Rp / k-step), emit = second
// pass.  Null for a combination that is not instantiated.
#define CLM_DECLARE_PICKERS(OPN, SYMN)                                \
    PairFn pick_pair_##OPN##_##SYMN(int ct, int NK, bool emit);       \
    AstkFn pick_astk_##OPN##_##SYMN(int NK);
CLM_DECLARE_PICKERS(f32, 0)
CLM_DECLARE_PICKERS(f32, 1)
CLM_DECLARE_PICKERS(bf16, 0)
CLM_DECLARE_PICKERS(bf16, 1)
CLM_DECLARE_PICKERS(f16, 0)
CLM_DECLARE_PICKERS(f16, 1)
#undef CLM_DECLARE_PICKERS

}  // namespace clm
