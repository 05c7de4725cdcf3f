// clm_kernels.hip -- gfx950 kernels of the common-lines engine (the pair
// kernel and A-stack templates are in clm_pair.h).
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
#include "clm_pair.h"

namespace clm {

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
                            unsigned long long* __restrict__ n_small,
                            const double* __restrict__ lsh) { // [K][L] per-line shift (px) or null
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
    const double a = lsh ? lsh[(size_t)k * L + l] : 0.0;
    for (int r = 0; r < R; r++) {
        const int rr = n_r - 1 - r;
        double re = ((double)zc[2 * rr] * H[r]) / nr;
        double im = ((double)zc[2 * rr + 1] * H[r]) / nr;
        if (lsh) {
            // line moved by a px: times exp(-2*pi*i*rk*a/(2R+1)), rk = r - R,
            // the phase the shift search applies for shift a
            const double th = ((-2.0 * M_PI * (double)(r - R)) * a) / (double)(2 * R + 1);
            const double c = cos(th), sn = sin(th);
            const double xr = re * c - im * sn, xi = re * sn + im * c;
            re = xr;
            im = xi;
        }
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

// ------------------------------------------------------------ re-scoring
struct RescoreArgs {
    const StackDev* stacks;   // the whole table (SlotRef.stack indexes it)
    int stack0, nstacks;      // mode 0: stacks [stack0, stack0 + nstacks)
    int mode;                 // 0: every pair of those stacks (pass-1 records),
                              // 1: list with pass-2 candidates, 2: list with pass-1 records
    long long n_items;        // mode 0: total pairs of all stacks
    const int* n_items_dev;   // modes 1, 2: number of list entries (device)
    int item_cap;             // modes 1, 2: entries beyond this are not processed
    const SlotRef* ulist;     // modes 1, 2: u -> (stack, slot)
    int stack_shift;          // modes 1, 2: added to the listed stack index
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

// Term r of C(l1, l2p) at shift j in FP64, the oracle's arithmetic: the row
// a = P1(r).*ph_j(r) (conj(P1) for the C2 half, flip), then a.re*b.re +
// a.im*b.im.
__device__ inline double term64(double2 p, double2 q, double2 b, bool flip) {
    double xr, xi;
    if (!flip) { xr = p.x * q.x - p.y * q.y; xi = p.x * q.y + p.y * q.x; }
    else       { xr = p.x * q.x + p.y * q.y; xi = p.x * q.y - p.y * q.x; }
    const double t = xr * b.x, u = xi * b.y;
    return t + u;
}

__device__ inline double readlane64(double x, int l) {
    const unsigned long long u = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(u & 0xffffffffu), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
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

// C at tie key `key` of pair (k1, k2) in FP64, bit-identical to the oracle:
// lane r forms term r (r = lane, lane + 64, ...) and the terms are added in
// order r = 0..R-1, acc = acc + term, as the oracle adds them.  Called by the
// whole wave with a wave-uniform key; the result is the same in every lane.
__device__ inline double eval64(const StackDev& sd, int k1, int k2, long long key, int lane) {
    const Geom& g = sd.g;
    const int l1 = (int)(key % g.L);
    const int l2p = (int)((key / g.L) % (2 * g.L));
    const int j = (int)(key / ((long long)g.L * 2 * g.L));
    const double2* P1 = sd.P64 + ((size_t)k1 * g.L + l1) * g.R;
    const double2* P2 = sd.P64 + ((size_t)k2 * g.L + (l2p % g.L)) * g.R;
    const double2* ph = sd.ph64 + (size_t)j * g.R;
    const bool flip = l2p >= g.L;
    double acc = 0.0;
    for (int c0 = 0; c0 < g.R; c0 += 64) {
        const int r = c0 + lane;
        const double tr = r < g.R ? term64(P1[r], ph[r], P2[r], flip) : 0.0;
        const int n = min(64, g.R - c0);
        for (int q = 0; q < n; q++) acc = acc + readlane64(tr, q);
    }
    return 2.0 * acc;
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
// per step, the row's lanes reading adjacent 16-byte pieces), values kept in
// LDS; then FP64 re-score of every candidate within tau of the FP32 best,
// the whole wave on one candidate (one lane per term, the terms added in
// the oracle's order).
__global__ __launch_bounds__(256) void rescore_kernel(RescoreArgs a) {
    __shared__ float s_val[4][RS_MAXV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int W = blockDim.x >> 6;
    float* sval = s_val[wave];
    const int q = lane & 3, rl = lane >> 2;  // k-quarter, row within the step
    const bool listed = a.mode != 0, emit = a.mode == 1;
    const long long n_items = listed ? (long long)min(*a.n_items_dev, a.item_cap) : a.n_items;
    // statistics, summed in lane 0 over the wave's items and added once at
    // the end: an atomic per pair on one address serialises across the chip
    unsigned long long c_cell = 0, c_emit = 0, c_64 = 0;
    for (long long item = (long long)blockIdx.x * W + wave; item < n_items; item += (long long)gridDim.x * W) {
        int stk, slot;
        if (!listed) {
            stk = a.stack0 + (a.nstacks > 1 ? find_stack(a.stacks + a.stack0, a.nstacks, item) : 0);
            slot = (int)(item - a.stacks[stk].pair_base);
        } else {
            stk = a.ulist[item].stack + a.stack_shift;
            slot = a.ulist[item].slot;
        }
        // the stack's geometry and pointers in registers: results are stored
        // through sd, so references into it would be re-read on every row
        const StackDev& sd = a.stacks[__builtin_amdgcn_readfirstlane(stk)];
        const Geom g = sd.g;
        const float2* const P32 = sd.P32;
        const float2* const ph32 = sd.ph32;
        const int kq = g.Rp / 4;  // multiple of 4
        const PairRec& rec = sd.recs[slot];
        const int k1 = rec.k1, k2 = rec.k2;
        const EmitRec* er = emit ? a.erecs + item : nullptr;
        const int cmode = emit ? 1 : 0;
        int nrows;
        if (!emit) {
            if (rec.flags & F_EMPTY) {
                if (lane == 0) write_result(sd, slot, 0.0, 0);
                continue;
            }
            if (rec.flags & (F_UNCERTAIN | F_OVERFLOW)) {
                if (lane == 0) { const int u = atomicAdd(a.unc_count, 1); a.unc_list[u] = SlotRef{stk, slot}; }
                continue;
            }
            nrows = 16 * rec.ncells;
            c_cell += (unsigned long long)rec.ncells;
        } else {
            if (er->overflow) {
                if (lane == 0) { const int u = atomicAdd(a.fb_count, 1); a.fb_list[u] = SlotRef{stk, slot}; }
                continue;
            }
            nrows = er->n;
            c_emit += (unsigned long long)nrows;
        }
        // ---- FP32 re-score
        float bv = -INFINITY;
        long long bk = 0x7fffffffffffffffLL;
        for (int base = 0; base < nrows; base += 16) {
            const int r = base + rl;
            int row = -1, col = 0;
            if (r < nrows) cand_rc(g, cmode, rec, er, r, row, col);
            float X = 0.f, Y = 0.f;
            int e = 0, l1 = 0;
            if (row >= 0) {
                e = row / g.L;
                l1 = row - e * g.L;
                // the 4 lanes of a row read 4 adjacent 16-byte pieces per
                // load (lane q: pieces q, q + 4, ...), so one load touches
                // 16 half cache lines instead of 64 lines
                const float4* p = (const float4*)(P32 + ((size_t)k1 * g.L + l1) * g.Rp) + q;
                const float4* bb = (const float4*)(P32 + ((size_t)k2 * g.L + col) * g.Rp) + q;
                const float4* ph = (const float4*)(ph32 + (size_t)e * g.Rp) + q;
#pragma unroll 4
                for (int i = 0; i < kq / 2; i++) {
                    const float4 pv = p[4 * i], qv = ph[4 * i], b4 = bb[4 * i];
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
            // ---- FP64 re-score of every candidate within tau of the FP32
            // best, one candidate at a time across the wave (the best by
            // value, then tie key, does not depend on the visiting order)
            const float lim = bv - sd.tau;
            double b64 = -INFINITY;
            long long k64 = 0x7fffffffffffffffLL;
            int n64 = 0;
            const int nv = min(2 * nrows, RS_MAXV);
            for (int base = 0; base < nv; base += 64) {
                const int iv = base + lane;
                unsigned long long mask = __ballot(iv < nv && sval[iv] >= lim);
                while (mask) {
                    const int ic = base + __builtin_ctzll(mask);
                    mask &= mask - 1;
                    int row, col;
                    cand_rc(g, cmode, rec, er, ic >> 1, row, col);
                    const int e = row / g.L, l1 = row - e * g.L;
                    long long key;
                    if (g.sym) key = (ic & 1) ? tie_key(g.S - 1 - e, g.L + col, l1, g.L) : tie_key(e, col, l1, g.L);
                    else key = e < g.S ? tie_key(e, col, l1, g.L) : tie_key(e - g.S, g.L + col, l1, g.L);
                    const double v = eval64(sd, k1, k2, key, lane);
                    n64++;
                    if (!isnan(v) && better64(v, key, b64, k64)) { b64 = v; k64 = key; }
                }
            }
            c_64 += (unsigned long long)n64;
            fv = b64;
            fk = k64;
        }
        if (lane == 0) write_result(sd, slot, fv, fk);
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) {
        if (a.ncell && c_cell) atomicAdd(a.ncell, c_cell);
        if (a.nemit && c_emit) atomicAdd(a.nemit, c_emit);
        if (a.n64 && c_64) atomicAdd(a.n64, c_64);
    }
}

// Pass-2 (or refinement) tiles from an uncertain list, on the device: one
// pair per tile (entry u -> tile u and explicit pair u); the tiles name stack
// r.stack + stack_shift (the FP32 twin of r.stack for the refinement).
// Entries past `cap` have no pass-2 buffers and go straight to the FP64
// exhaustive list.
__global__ void build_pass2_kernel(const StackDev* stacks, const int* n_unc, const SlotRef* ulist, int cap,
                                   int4* tiles, int4* plist, int* ntiles, int* fb_count, SlotRef* fb_list,
                                   int stack_shift) {
    const int n = *n_unc;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *ntiles = min(n, cap);
    for (int u = i; u < n; u += gridDim.x * blockDim.x) {
        const SlotRef r = ulist[u];
        if (u < cap) {
            const PairRec& rec = stacks[r.stack].recs[r.slot];
            tiles[u] = make_int4(rec.k1, u, 1 | ((r.stack + stack_shift) << 4), -1);
            plist[u] = make_int4(rec.k2, r.slot, u, 0);
        } else {
            const int f = atomicAdd(fb_count, 1);
            fb_list[f] = r;
        }
    }
}

// ------------------------------------------------------ FP64 exhaustive
// Every candidate of a pair in FP64 with the oracle's arithmetic: for shift
// j and map m the rows a(r) = P1(r).*ph_j(r) (conj(P1) for C2) are formed
// exactly as the oracle forms them, and each entry accumulates
// acc = acc + (a.re*b.re + a.im*b.im) over r = 0..R-1 in order, so every value
// is bit-identical to the oracle's.  Computed as 64 x 64 output tiles (rows
// l1, columns l2 of one map) with 16 outputs per thread and the operands of
// 32 consecutive r staged in LDS.  The best (value, tie key) wins; the key
// q = (j*2L + l2p)*L + l1 is the reference's first-max order.
struct Exact64Args {
    const StackDev* stacks;   // the whole table (SlotRef.stack indexes it)
    int stack0, nstacks;      // without a list: stacks [stack0, stack0 + nstacks)
    const SlotRef* list;      // pairs to compute, or null: every pair of those stacks
    const int* n_dev;         // list length (device)
    long long n_all;          // pairs when list is null
    long long item0;          // this launch handles items [item0, item0 + item_cap)
    long long item_cap;
    int sj;                   // shift ranges per item (gridDim.y); 1: a workgroup does the whole item
    double* sv;               // sj > 1: best value and key per (item - item0, range)
    long long* sk;
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

__device__ inline void exact_item(const Exact64Args& a, long long it, int& stk, int& slot) {
    if (a.list) { stk = a.list[it].stack; slot = a.list[it].slot; }
    else {
        stk = a.stack0 + (a.nstacks > 1 ? find_stack(a.stacks + a.stack0, a.nstacks, it) : 0);
        slot = (int)(it - a.stacks[stk].pair_base);
    }
}

__global__ __launch_bounds__(256) void exact64_kernel(Exact64Args a) {
    __shared__ double2 sA[32][64];  // a(r) of 64 rows l1, 32 consecutive r
    __shared__ double2 sB[32][64];  // P2(r) of 64 columns l2
    const long long n_tot = a.list ? (long long)*a.n_dev : a.n_all;
    const long long n = n_tot < a.item0 + a.item_cap ? n_tot : a.item0 + a.item_cap;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (long long it = a.item0 + blockIdx.x; it < n; it += gridDim.x) {
        int stk, slot;
        exact_item(a, it, stk, slot);
        const StackDev& sd = a.stacks[stk];
        const int S = sd.g.S, L = sd.g.L, R = sd.g.R;
        int k1, k2;
        slot_to_pair(slot, sd.g.K, k1, k2);
        const int j0 = (int)((long long)blockIdx.y * S / a.sj), j1 = (int)((long long)(blockIdx.y + 1) * S / a.sj);
        const int NT = (L + 63) / 64;
        const double2* P1 = sd.P64 + (size_t)k1 * L * R;
        const double2* P2 = sd.P64 + (size_t)k2 * L * R;
        double bv = -INFINITY;
        long long bk = 0x7fffffffffffffffLL;
        for (int j = j0; j < j1; j++) {
            const double2* ph = sd.ph64 + (size_t)j * R;
            for (int m = 0; m < 2; m++)
                for (int rt = 0; rt < NT; rt++)
                    for (int ct = 0; ct < NT; ct++) {
                        double acc[4][4];
#pragma unroll
                        for (int i = 0; i < 4; i++)
#pragma unroll
                            for (int k = 0; k < 4; k++) acc[i][k] = 0.0;
                        for (int r0 = 0; r0 < R; r0 += 32) {
                            const int rn = R - r0 < 32 ? R - r0 : 32;
                            __syncthreads();
                            for (int idx = tid; idx < 32 * 64; idx += 256) {
                                const int rr = idx >> 6, col = idx & 63;
                                double2 av = make_double2(0.0, 0.0), bb = make_double2(0.0, 0.0);
                                const int l1 = rt * 64 + col, l2 = ct * 64 + col;
                                if (rr < rn && l1 < L) {
                                    const double2 pv = P1[(size_t)l1 * R + r0 + rr], qv = ph[r0 + rr];
                                    const double pr = pv.x, pi = pv.y, qr = qv.x, qi = qv.y;
                                    if (m == 0) av = make_double2(pr * qr - pi * qi, pr * qi + pi * qr);
                                    else        av = make_double2(pr * qr + pi * qi, pr * qi - pi * qr);
                                }
                                if (rr < rn && l2 < L) bb = P2[(size_t)l2 * R + r0 + rr];
                                sA[rr][col] = av;
                                sB[rr][col] = bb;
                            }
                            __syncthreads();
                            for (int rr = 0; rr < rn; rr++) {
                                double2 A[4], B[4];
#pragma unroll
                                for (int i = 0; i < 4; i++) {
                                    A[i] = sA[rr][ty + 16 * i];
                                    B[i] = sB[rr][tx + 16 * i];
                                }
#pragma unroll
                                for (int i = 0; i < 4; i++)
#pragma unroll
                                    for (int k = 0; k < 4; k++) {
                                        const double t = A[i].x * B[k].x, u = A[i].y * B[k].y;
                                        acc[i][k] = acc[i][k] + (t + u);
                                    }
                            }
                        }
#pragma unroll
                        for (int i = 0; i < 4; i++)
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                const int l1 = rt * 64 + ty + 16 * i, l2 = ct * 64 + tx + 16 * k;
                                if (l1 < L && l2 < L) {
                                    const double v = 2.0 * acc[i][k];
                                    const long long key = tie_key(j, l2 + m * L, l1, L);
                                    if (!isnan(v) && better64(v, key, bv, bk)) { bv = v; bk = key; }
                                }
                            }
                    }
        }
        // workgroup best (LDS of the tiles reused)
        __syncthreads();
        double* s_v = (double*)&sA[0][0];
        long long* s_k = (long long*)&sB[0][0];
        s_v[tid] = bv;
        s_k[tid] = bk;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o && better64(s_v[tid + o], s_k[tid + o], s_v[tid], s_k[tid])) {
                s_v[tid] = s_v[tid + o];
                s_k[tid] = s_k[tid + o];
            }
            __syncthreads();
        }
        if (tid == 0) {
            if (a.sj == 1) write_result(sd, slot, s_v[0], s_k[0]);
            else {
                const long long e = (it - a.item0) * a.sj + blockIdx.y;
                a.sv[e] = s_v[0];
                a.sk[e] = s_k[0];
            }
        }
        __syncthreads();
    }
}

// Best over an item's shift ranges (sj > 1), one thread per item.
__global__ void exact64_reduce_kernel(Exact64Args a) {
    const long long n_tot = a.list ? (long long)*a.n_dev : a.n_all;
    const long long it = a.item0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n_tot || it >= a.item0 + a.item_cap) return;
    double bv = -INFINITY;
    long long bk = 0x7fffffffffffffffLL;
    for (int y = 0; y < a.sj; y++) {
        const long long e = (it - a.item0) * a.sj + y;
        if (better64(a.sv[e], a.sk[e], bv, bk)) { bv = a.sv[e]; bk = a.sk[e]; }
    }
    int stk, slot;
    exact_item(a, it, stk, slot);
    write_result(a.stacks[stk], slot, bv, bk);
}

// FP32 operand planes [K][2][Lp][Rp] of a stack from its FP32 half lines
// (the values prep_kernel writes for FP32 operands; zero padding).
__global__ void bop32_kernel(Geom g, const float2* __restrict__ P32, float* __restrict__ Bop) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)g.K * 2 * g.Lp * g.Rp;
    if (i >= n) return;
    const int r = (int)(i % g.Rp);
    const int l = (int)((i / g.Rp) % g.Lp);
    const int plane = (int)((i / ((long long)g.Rp * g.Lp)) % 2);
    const long long k = i / ((long long)g.Rp * g.Lp * 2);
    float v = 0.f;
    if (l < g.L) {
        const float2 z = P32[((size_t)k * g.L + l) * g.Rp + r];
        v = plane ? z.y : z.x;
    }
    Bop[i] = v;
}

__global__ void fill_i32(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------ windowed search
// Per-line 1-D shift of image k implied by a proposed 2-D shift (x, y):
// a(l) = sin(l*pi/L)*x + cos(l*pi/L)*y, the coefficients of the shift
// equations (cryo_clmatrix_cpu.m:722-738), rounded to a multiple of `step`
// so that window shifts land on the reference grid.
__global__ void lineshift_kernel(const float* __restrict__ prop, int K, int L, double step,
                                 double* __restrict__ lsh) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)K * L) return;
    const int k = (int)(i / L), l = (int)(i % L);
    const double th = (double)l * M_PI / (double)L;
    const double a = sin(th) * (double)prop[2 * k] + cos(th) * (double)prop[2 * k + 1];
    lsh[i] = rint(a / step) * step;
}

// Window result -> reference grid.  Line l1 of image k1 was moved by a1 and
// line l2 of k2 by a2, so window shift s at a C1 column is grid shift
// s + a1 - a2 and at a C2 column s - a1 - a2.  Pairs whose window peak is
// below `gate`, or whose grid shift is off the grid, are flagged for the
// exhaustive search; the others are written in the reference's form.
__global__ void window_merge_kernel(int K, int L, const double* __restrict__ lsh,
                                    const int32_t* __restrict__ wcl1, const int32_t* __restrict__ wcl2,
                                    const double* __restrict__ wcorr, const double* __restrict__ wshift,
                                    double gate, double max_shift, double step, int S_full,
                                    const double* __restrict__ shiftv_full,
                                    int32_t* cl1, int32_t* cl2, int32_t* jidx, double* corr, double* shift1d,
                                    uint8_t* __restrict__ flag, unsigned long long* __restrict__ counts) {
    const long long P = (long long)K * (K - 1) / 2;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = e < P;
    bool ok = false, gated = false;
    if (valid) {
        int k1, k2;
        slot_to_pair(e, K, k1, k2);
        const int c1 = wcl1[e], c2 = wcl2[e];
        const double v = wcorr[e];
        int j = -1;
        gated = c1 > 0 && v >= gate;
        if (gated) {
            const double a1 = lsh[(size_t)k1 * L + (c1 - 1)];
            const int l2 = (c2 - 1) % L;
            const double a2 = lsh[(size_t)k2 * L + l2];
            const double sfull = c2 <= L ? wshift[e] + a1 - a2 : wshift[e] - a1 - a2;
            j = (int)rint((sfull + max_shift) / step);
            ok = j >= 0 && j < S_full;
        }
        if (ok) {
            if (cl1) cl1[e] = c1;
            if (cl2) cl2[e] = c2;
            if (jidx) jidx[e] = j;
            if (corr) corr[e] = v;
            if (shift1d) shift1d[e] = shiftv_full[j];
            flag[e] = 0;
        } else {
            flag[e] = 1;
        }
    }
    // counts 0: accepted, 1: below the gate, 2: off the grid; one atomic per
    // wave and kind (an atomic per pair on one address serialises)
    const unsigned long long m0 = __ballot(valid && ok);
    const unsigned long long m1 = __ballot(valid && !gated);
    const unsigned long long m2 = __ballot(valid && gated && !ok);
    if ((threadIdx.x & 63) == 0) {
        if (m0) atomicAdd(counts + 0, (unsigned long long)__popcll(m0));
        if (m1) atomicAdd(counts + 1, (unsigned long long)__popcll(m1));
        if (m2) atomicAdd(counts + 2, (unsigned long long)__popcll(m2));
    }
}

// Common lines of known rotations and the best 1-D shift along each: the
// per-pair loop of cryo_estimate_shifts.m:142-223 (pair order k1 < k2; the
// reference draws its pairs in random order, which the solution does not
// depend on).  One wave per pair; lane s evaluates the shifts s, s + 64, ...
// with the FP64 terms of eval64 added in order over the ray, then the wave
// finds the first maximum of C1 and of C2 (NaN ignored, as MATLAB's max).
__device__ inline bool first_max(double v, int s, double bv, int bs) {
    return !isnan(v) && (isnan(bv) || v > bv || (v == bv && s < bs));
}

// Read-only FP64 data read through the constant address space: loads from a
// wave-uniform address become scalar loads.
typedef const double __attribute__((address_space(4)))* ConstD;

// One wave takes 64 consecutive pairs at a time: each lane writes the
// outputs of its pair if the selection leaves it out (coalesced), then the
// wave computes the selected ones one after another (a ballot lists them).
// With PH_LDS the shift phases are staged in LDS once per workgroup as
// [r][s], so that the lanes (one per shift) read consecutive entries at
// each ray index; otherwise they are read from the [s][r] table in global
// memory.  Eight waves per workgroup share one staged table.

template <bool PH_LDS>
__global__ __launch_bounds__(512) void rotlines_kernel(const StackDev* __restrict__ sdp,
                                                       const double* __restrict__ rot,
                                                       const uint8_t* __restrict__ sel, int n_theta,
                                                       int32_t* __restrict__ cl1, int32_t* __restrict__ cl2,
                                                       double* __restrict__ shift1d) {
    extern __shared__ double2 s_ph[];  // [R][S] when PH_LDS
    const StackDev& sd = *sdp;
    const Geom g = sd.g;
    const long long K = g.K, P = K * (K - 1) / 2;
    const int lane = threadIdx.x & 63;
    if (PH_LDS) {
        for (int i = threadIdx.x; i < g.S * g.R; i += blockDim.x) {
            const int s = i / g.R, r = i - s * g.R;
            s_ph[r * g.S + s] = sd.ph64[i];
        }
        __syncthreads();
    }
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    // the wave's index, made visibly uniform (scalar) for the compiler
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (long long base = ((long long)blockIdx.x * (blockDim.x >> 6) + wid) * 64; base < P; base += nw * 64) {
        const long long el = base + lane;
        const bool act = el < P && (!sel || sel[el]);
        if (el < P && !act) {
            cl1[el] = -1;
            cl2[el] = -1;
            shift1d[el] = 0.0;
        }
        unsigned long long todo = __ballot(act);
        while (todo) {
            const long long e = base + __builtin_ctzll(todo);
            todo &= todo - 1;
            // (k1, k2) of pair e
            long long k1 = (long long)floor(((2.0 * K - 1.0) - sqrt((2.0 * K - 1.0) * (2.0 * K - 1.0) - 8.0 * (double)e)) / 2.0);
            if (k1 < 0) k1 = 0;
            if (k1 > K - 2) k1 = K - 2;
            auto first = [&](long long x) { return x * K - x * (x + 1) / 2; };
            while (k1 > 0 && first(k1) > e) k1--;
            while (k1 + 1 <= K - 2 && first(k1 + 1) <= e) k1++;
            const long long k2 = e - first(k1) + k1 + 1;
            // commonline_R(R_k1.', R_k2.', n_theta): Ut = R_k2.' * R_k1, entry
            // (a, b) = sum over m of R_k2(m, a) R_k1(m, b), m in order
            const double* Ri = rot + 9 * k1;
            const double* Rj = rot + 9 * k2;
            auto ut = [&](int a, int b) {
                return (Rj[3 * a] * Ri[3 * b] + Rj[1 + 3 * a] * Ri[1 + 3 * b]) + Rj[2 + 3 * a] * Ri[2 + 3 * b];
            };
            const double PI = M_PI;  // 4*atan(1.0) in commonline_R.m:27
            const double aij = atan2(ut(2, 0), -ut(2, 1)) + PI;
            const double aji = atan2(-ut(0, 2), ut(1, 2)) + PI;
            // round half away from zero, then mod n_theta; angles are in [0, 2 pi]
            // (a rotation with NaN entries gives line 0)
            double xij = round(aij / (2 * PI) * n_theta), xji = round(aji / (2 * PI) * n_theta);
            if (!(xij >= 0.0 && xij <= n_theta)) xij = 0.0;
            if (!(xji >= 0.0 && xji <= n_theta)) xji = 0.0;
            int cij = (int)xij % n_theta;
            int cji = (int)xji % n_theta;
            const int L2 = n_theta / 2;
            if (cij >= L2) {  // cryo_estimate_shifts.m:155-163
                cij -= L2;
                cji -= L2;
            }
            if (cji < 0) cji += n_theta;
            cij += 1;
            cji += 1;
            // the two rays are the same for every lane: read through the
            // constant address space, so that they come in by scalar loads
            // (one per wave) instead of a vector load of one address per lane
            // (values computed in vector registers from uniform inputs; read
            // from the first lane so that the addresses are scalar)
            const int lj = __builtin_amdgcn_readfirstlane(cji <= L2 ? cji - 1 : cji - 1 - L2);
            const int li = __builtin_amdgcn_readfirstlane(cij - 1);
            const int q1 = __builtin_amdgcn_readfirstlane((int)k1), q2 = __builtin_amdgcn_readfirstlane((int)k2);
            const ConstD ri = (ConstD)(sd.P64 + ((size_t)q1 * g.L + li) * g.R);
            const ConstD rj = (ConstD)(sd.P64 + ((size_t)q2 * g.L + lj) * g.R);
            double b1 = NAN, b2 = NAN;
            int s1 = 0, s2 = 0;
            // the shift loop is uniform (lanes past the last shift compute
            // the last one again and are ignored), so the ray loop runs
            // outside divergent control and its ray addresses stay scalar
            for (int s0 = 0; s0 < g.S; s0 += 64) {
                const int s = s0 + lane;
                const int sc = s < g.S ? s : g.S - 1;
                double a1 = 0.0, a2 = 0.0;
#pragma unroll 4
                for (int r = 0; r < g.R; r++) {
                    const double2 pv = make_double2(ri[2 * r], ri[2 * r + 1]);
                    const double2 bv = make_double2(rj[2 * r], rj[2 * r + 1]);
                    const double2 qv = PH_LDS ? s_ph[r * g.S + sc] : sd.ph64[(size_t)sc * g.R + r];
                    a1 = a1 + term64(pv, qv, bv, false);
                    a2 = a2 + term64(pv, qv, bv, true);
                }
                const double c1 = 2.0 * a1, c2 = 2.0 * a2;
                if (s < g.S && first_max(c1, s, b1, s1)) { b1 = c1; s1 = s; }
                if (s < g.S && first_max(c2, s, b2, s2)) { b2 = c2; s2 = s; }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double v1 = __shfl_xor(b1, o, 64), v2 = __shfl_xor(b2, o, 64);
                const int t1 = __shfl_xor(s1, o, 64), t2 = __shfl_xor(s2, o, 64);
                if (first_max(v1, t1, b1, s1)) { b1 = v1; s1 = t1; }
                if (first_max(v2, t2, b2, s2)) { b2 = v2; s2 = t2; }
            }
            if (lane == 0) {
                // an all-NaN max is NaN at index 1 (s = 0) in MATLAB, as here
                cl1[e] = cij;
                cl2[e] = cji;
                shift1d[e] = sd.shiftv[b1 > b2 ? s1 : s2];
            }
        }
    }
}

// Partner-subset search: the pairs left out keep the outputs of a pair with
// no common line (cryo_clmatrix_gpu_1.m:346-350 searches only the drawn
// partners; clstack and corrstack stay zero elsewhere).  The selected pairs'
// outputs are written by the search itself.
__global__ void subset_init_kernel(long long P, const uint8_t* __restrict__ sel, int32_t* __restrict__ cl1,
                                   int32_t* __restrict__ cl2, int32_t* __restrict__ jidx, double* __restrict__ corr,
                                   double* __restrict__ shift1d) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < P; e += stride) {
        if (sel[e]) continue;
        cl1[e] = 0;
        cl2[e] = 0;
        jidx[e] = -1;
        corr[e] = 0.0;
        shift1d[e] = 0.0;
    }
}

// Flagged pairs -> explicit pass-1 tiles, in pair order.  Three steps: per
// row counts, one scan, per row ordered compaction.
__global__ void row_count_kernel(int K, const uint8_t* __restrict__ flag, int ppw, int* __restrict__ cnt,
                                 int* __restrict__ tcnt) {
    __shared__ int s_c;
    const int k1 = blockIdx.x;
    if (k1 >= K - 1) return;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    const long long f = (long long)k1 * K - (long long)k1 * (k1 + 1) / 2;
    const int n = K - 1 - k1;
    int c = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) c += flag[f + i] ? 1 : 0;
    atomicAdd(&s_c, c);
    __syncthreads();
    if (threadIdx.x == 0) {
        cnt[k1] = s_c;
        tcnt[k1] = (s_c + ppw - 1) / ppw;
    }
}

// Exclusive scans of cnt and tcnt over n rows (one workgroup of 1024).
// Totals go to tot[0] (pairs) and tot[1] (tiles).
__global__ __launch_bounds__(1024) void row_scan_kernel(int n, const int* __restrict__ cnt,
                                                        const int* __restrict__ tcnt, int* __restrict__ off,
                                                        int* __restrict__ toff, int* __restrict__ tot) {
    __shared__ int sa[1024], sb[1024];
    const int t = threadIdx.x, T = blockDim.x;
    const int per = (n + T - 1) / T;
    const int lo = min(n, t * per), hi = min(n, lo + per);
    int a = 0, b = 0;
    for (int i = lo; i < hi; i++) { a += cnt[i]; b += tcnt[i]; }
    sa[t] = a;
    sb[t] = b;
    __syncthreads();
    for (int o = 1; o < T; o <<= 1) {
        const int va = t >= o ? sa[t - o] : 0, vb = t >= o ? sb[t - o] : 0;
        __syncthreads();
        sa[t] += va;
        sb[t] += vb;
        __syncthreads();
    }
    int ea = sa[t] - a, eb = sb[t] - b;
    for (int i = lo; i < hi; i++) {
        off[i] = ea;
        toff[i] = eb;
        ea += cnt[i];
        eb += tcnt[i];
    }
    if (t == T - 1) {
        tot[0] = sa[t];
        tot[1] = sb[t];
    }
}

__global__ __launch_bounds__(256) void row_emit_kernel(int K, int stack, const uint8_t* __restrict__ flag, int ppw,
                                                       const int* __restrict__ off, const int* __restrict__ toff,
                                                       SlotRef* __restrict__ list, int4* __restrict__ plist,
                                                       int4* __restrict__ tiles) {
    __shared__ int s_w[4];
    const int k1 = blockIdx.x;
    if (k1 >= K - 1) return;
    const long long f = (long long)k1 * K - (long long)k1 * (k1 + 1) / 2;
    const int n = K - 1 - k1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int base = off[k1];
    const int tb = toff[k1], ob = off[k1];
    for (int i0 = 0; i0 < n; i0 += blockDim.x) {
        const int i = i0 + threadIdx.x;
        const bool on = i < n && flag[f + i];
        const unsigned long long m = __ballot(on);
        if (lane == 0) s_w[wave] = __popcll(m);
        __syncthreads();
        int pre = 0;
        for (int w = 0; w < wave; w++) pre += s_w[w];
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (on) {
            const int pos = base + pre + __popcll(m & ((1ull << lane) - 1ull));
            const int slot = (int)(f + i);
            list[pos] = SlotRef{stack, slot};
            plist[pos] = make_int4(k1 + 1 + i, slot, pos, 0);
        }
        base += tot;
        __syncthreads();
    }
    // tiles of this row: PPW consecutive listed pairs each
    const int cnt = base - ob;
    for (int q = threadIdx.x * ppw; q < cnt; q += blockDim.x * ppw)
        tiles[tb + q / ppw] = make_int4(k1, ob + q, min(ppw, cnt - q) | (stack << 4), -1);
}

}  // namespace clm
