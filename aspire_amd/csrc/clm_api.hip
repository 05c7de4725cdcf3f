// clm_api.hip -- C ABI of libclm (declared in include/clm.h): context,
// geometry, orchestration of the device passes, host-side FP64 assembly.
#include "clm.h"
#include "clm_kernels.hip"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace clm;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

}  // namespace

struct clm_ctx {
    int device = 0;
    int n_cu = 0;
    int lds_max = 65536;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf P64, P32, Bop, nrm, recs, tiles, plist, ctr, ulist, erecs, fblist;
    DevBuf ph32, ph64, shiftv, H, pf, out_cl1, out_cl2, out_j, out_corr, out_shift, uslots;
    hipEvent_t ev[6] = {};
    std::vector<int4> h_tiles;
    int tiles_K = -1, tiles_PPW = -1;
};

namespace {

int fail(clm_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail((ctx), CLM_E_HIP, "%s failed: %s (%s:%d)", #expr,              \
                        hipGetErrorString(e_), __FILE__, __LINE__);                    \
    } while (0)

int ensure(clm_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.n >= bytes) return CLM_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.n = 0;
        if (e != hipSuccess) return fail(c, CLM_E_HIP, "hipFree: %s", hipGetErrorString(e));
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, CLM_E_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    b.n = bytes;
    return CLM_OK;
}

void release(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

// Operand element kinds (prep_kernel's op_kind)
enum { OPK_F32 = 0, OPK_BF16 = 1, OPK_F16 = 2 };

// Kernel variant: column tiles per wave and waves per workgroup.
struct Variant {
    int ct = 0, W = 0, CPP = 0, NCTp = 0;
};

// Supported padded ray lengths (k-steps of 16).
int pad_R(int R) {
    static const int opts[] = {16, 32, 48, 64, 96, 128};
    for (int o : opts)
        if (R <= o) return o;
    return -1;
}

Variant choose_variant(int opk, int NCT, int Rp, int want_ct, int want_W) {
    Variant best;
    double best_waste = 1e9;
    const int cts_mfma16[] = {3, 2, 1};
    const int cts_f32[] = {2, 1};
    const int* cts = opk == OPK_F32 ? cts_f32 : cts_mfma16;
    const int ncts = opk == OPK_F32 ? 2 : 3;
    for (int i = 0; i < ncts; i++) {
        const int ct = cts[i];
        if (want_ct && ct != want_ct) continue;
        if (opk == OPK_F32 && ct == 2 && Rp > 64) continue;
        if (opk != OPK_F32 && ct == 3 && Rp > 64) continue;  // register budget
        const int NCTp = (NCT + ct - 1) / ct * ct;
        const int CPP = NCTp / ct;
        if (CPP > 8) continue;
        int W = want_W ? want_W : (8 / CPP) * CPP;
        if (opk == OPK_F32 && !want_W) W = CPP * std::max(1, 4 / CPP);
        if (W < CPP || W % CPP || W > 8) continue;
        const double waste = (double)NCTp / NCT;
        if (waste < best_waste - 1e-9) {
            best_waste = waste;
            best.ct = ct; best.W = W; best.CPP = CPP; best.NCTp = NCTp;
        }
    }
    return best;
}

using PairFn = void (*)(PairArgs);

template <typename OpT, int CT, int NK, bool SYM>
PairFn pk(bool emit) {
    return emit ? pair_kernel<OpT, CT, NK, SYM, true> : pair_kernel<OpT, CT, NK, SYM, false>;
}
template <typename OpT, int CT, bool SYM>
PairFn pick_nk(int NK, bool emit) {
    switch (NK) {
        case 1: return pk<OpT, CT, 1, SYM>(emit);
        case 2: return pk<OpT, CT, 2, SYM>(emit);
        case 3: return pk<OpT, CT, 3, SYM>(emit);
        case 4: return pk<OpT, CT, 4, SYM>(emit);
        case 6: return pk<OpT, CT, 6, SYM>(emit);
        case 8: return pk<OpT, CT, 8, SYM>(emit);
    }
    return nullptr;
}
template <int CT, bool SYM>
PairFn pick_f32(int NK, bool emit) {
    switch (NK) {
        case 8: return pk<float, CT, 8, SYM>(emit);
        case 16: return pk<float, CT, 16, SYM>(emit);
        case 24: return pk<float, CT, 24, SYM>(emit);
        case 32: return pk<float, CT, 32, SYM>(emit);
        case 48: return pk<float, CT, 48, SYM>(emit);
        case 64: return pk<float, CT, 64, SYM>(emit);
    }
    return nullptr;
}
template <typename OpT, bool SYM>
PairFn pick_ct(int ct, int NK, bool emit) {
    switch (ct) {
        case 1: return pick_nk<OpT, 1, SYM>(NK, emit);
        case 2: return pick_nk<OpT, 2, SYM>(NK, emit);
        case 3: return pick_nk<OpT, 3, SYM>(NK, emit);
    }
    return nullptr;
}

PairFn pick_pair_kernel(int opk, int ct, int Rp, bool sym, bool emit) {
    if (opk == OPK_F32) {
        const int NK = Rp / 2;
        if (ct == 1) return sym ? pick_f32<1, true>(NK, emit) : pick_f32<1, false>(NK, emit);
        if (ct == 2 && NK <= 32) return sym ? pick_f32<2, true>(NK, emit) : pick_f32<2, false>(NK, emit);
        return nullptr;
    }
    const int NK = Rp / 16;
    if (opk == OPK_BF16) return sym ? pick_ct<__bf16, true>(ct, NK, emit) : pick_ct<__bf16, false>(ct, NK, emit);
    return sym ? pick_ct<_Float16, true>(ct, NK, emit) : pick_ct<_Float16, false>(ct, NK, emit);
}

inline long long pair_slot(long long k1, long long k2, long long K) {
    return k1 * K - k1 * (k1 + 1) / 2 + (k2 - k1 - 1);
}

void slot_pair_host(long long s, long long K, int& k1, int& k2) {
    double Kd = (double)K;
    double d = (2 * Kd - 1) * (2 * Kd - 1) - 8.0 * (double)s;
    long long a = (long long)std::floor(((2 * Kd - 1) - std::sqrt(std::max(d, 0.0))) / 2);
    if (a < 0) a = 0;
    auto first = [&](long long x) { return x * K - x * (x + 1) / 2; };
    while (a > 0 && first(a) > s) a--;
    while (first(a + 1) <= s) a++;
    k1 = (int)a;
    k2 = (int)(s - first(a) + a + 1);
}

struct Setup {
    Geom g;
    std::vector<double> shifts;   // s_j, j < S
    std::vector<float2> ph32;     // [NE][Rp]
    std::vector<double2> ph64;    // [S][R]
    std::vector<double> H;        // [2n_r-1]
};

// Everything the reference derives from the arguments before the pair loop.
int setup(clm_ctx* c, const clm_params* p, Setup& s) {
    if (p->n_theta <= 0 || p->n_theta % 2) return fail(c, CLM_E_ARG, "n_theta must be even");
    if (p->n_r < 2) return fail(c, CLM_E_ARG, "n_r must be at least 2 (got %d)", p->n_r);
    if (p->K < 0) return fail(c, CLM_E_ARG, "K must be non-negative");
    if (!(p->shift_step > 0)) return fail(c, CLM_E_ARG, "shift_step must be positive");
    if (!(p->max_shift >= 0)) return fail(c, CLM_E_ARG, "max_shift must be non-negative");
    Geom& g = s.g;
    g.K = p->K;
    g.n_r = p->n_r;
    g.T = p->n_theta;
    g.L = p->n_theta / 2;
    g.R = p->n_r - 1;
    // n_shifts=ceil(2*max_shift/shift_step+1)  cryo_clmatrix_cpu.m:140
    const double ns = std::ceil(2.0 * p->max_shift / p->shift_step + 1.0);
    if (ns > 4096) return fail(c, CLM_E_UNSUP, "too many shifts (%g)", ns);
    g.S = (int)ns;
    s.shifts.resize(g.S);
    for (int j = 0; j < g.S; j++) s.shifts[j] = -p->max_shift + (double)j * p->shift_step;
    bool sym = true;
    for (int j = 0; j < g.S; j++) {
        const double a = s.shifts[j], b = s.shifts[g.S - 1 - j];
        if (std::fabs(a + b) > 1e-12 * std::max(1.0, std::fabs(a))) sym = false;
    }
    g.sym = sym ? 1 : 0;
    g.NE = sym ? g.S : 2 * g.S;
    const int R = g.R;
    g.Rp = pad_R(R);
    // filter H (cryo_clmatrix_cpu.m:279) with rows rk in {-1,0,1} zeroed (:290)
    s.H.resize(2 * g.n_r - 1);
    for (int i = 0; i < 2 * g.n_r - 1; i++) {
        const int rk = i - R;
        const double r = (double)rk, w = (double)R / 4.0;
        s.H[i] = (rk >= -1 && rk <= 1) ? 0.0 : std::sqrt(std::fabs(r)) * std::exp(-(r * r) / (2.0 * (w * w)));
    }
    // phases exp(-2*pi*i*rk2*shift/(2*rmax+1))  :362
    auto theta = [&](int r, double shift) {
        const double rk = (double)(r - R);
        return ((-2.0 * M_PI * rk) * shift) / (double)(2 * R + 1);
    };
    s.ph64.resize((size_t)g.S * R);
    for (int j = 0; j < g.S; j++)
        for (int r = 0; r < R; r++) {
            const double th = theta(r, s.shifts[j]);
            s.ph64[(size_t)j * R + r] = make_double2(std::cos(th), std::sin(th));
        }
    if (g.Rp > 0) {
        s.ph32.assign((size_t)(g.NE + 1) * g.Rp, make_float2(0.f, 0.f));  // row NE stays zero
        for (int e = 0; e < g.NE; e++) {
            const double sh = e < g.S ? s.shifts[e] : -s.shifts[e - g.S];
            for (int r = 0; r < R; r++) {
                const double th = theta(r, sh);
                s.ph32[(size_t)e * g.Rp + r] = make_float2((float)std::cos(th), (float)std::sin(th));
            }
        }
    }
    return CLM_OK;
}

template <typename T>
int upload(clm_ctx* c, DevBuf& b, const std::vector<T>& v) {
    int rc = ensure(c, b, v.size() * sizeof(T));
    if (rc) return rc;
    if (!v.empty()) HIPCHK(c, hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return CLM_OK;
}

int fill(clm_ctx* c, void* p, int n, int v) {
    if (n <= 0) return CLM_OK;
    hipLaunchKernelGGL(fill_i32, dim3((n + 255) / 256), dim3(256), 0, c->stream, (int*)p, n, v);
    HIPCHK(c, hipGetLastError());
    return CLM_OK;
}

float event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

// Device job: everything from pf (device) to the per-pair records (device).
int run_device(clm_ctx* c, const clm_params* p, const void* d_pf, int32_t* d_cl1, int32_t* d_cl2,
               int32_t* d_j, double* d_corr, double* d_shift, clm_stats* st) {
    Setup s;
    int rc = setup(c, p, s);
    if (rc) return rc;
    Geom& g = s.g;
    const long long K = g.K;
    const long long P = K * (K - 1) / 2;
    if (st) {
        std::memset(st, 0, sizeof *st);
        st->pairs = P;
        st->n_eval_shifts = g.NE;
    }
    if (P == 0) return CLM_OK;
    if (P > 0x7fffffffLL) return fail(c, CLM_E_UNSUP, "K too large (%lld pairs)", P);
    const int prec = p->precision;
    if (prec < CLM_FP32 || prec > CLM_FP64_REF) return fail(c, CLM_E_ARG, "unknown precision %d", prec);
    if (p->pf_dtype != CLM_C128 && p->pf_dtype != CLM_C64) return fail(c, CLM_E_ARG, "unknown pf dtype %d", p->pf_dtype);
    if (g.Rp < 0) return fail(c, CLM_E_UNSUP, "n_r=%d exceeds the supported 129", g.n_r);
    const int opk = prec == CLM_BF16_RESCORE ? OPK_BF16 : prec == CLM_FP16_RESCORE ? OPK_F16 : OPK_F32;
    const bool use64 = prec != CLM_FP32;
    const int NCT = (g.L + 31) / 32;
    Variant v = choose_variant(opk, NCT, g.Rp, p->variant / 100, p->variant % 100);
    if (!v.ct) return fail(c, CLM_E_UNSUP, "no kernel variant for L=%d (variant %d)", g.L, p->variant);
    g.NCT = v.NCTp;
    g.Lp = 32 * v.NCTp;
    g.NROWS = g.NE * g.L;
    g.NAT = (g.NROWS + 31) / 32;
    if ((long long)g.NROWS * g.Lp > 0xffffffffLL) return fail(c, CLM_E_UNSUP, "map too large");
    if (st) st->variant = v.ct * 100 + v.W;
    const size_t osz = opk == OPK_F32 ? 4 : 2;

    // ---- buffers
    if ((rc = ensure(c, c->P32, (size_t)K * g.L * g.Rp * sizeof(float2)))) return rc;
    if (use64 && (rc = ensure(c, c->P64, (size_t)K * g.L * g.R * sizeof(double2)))) return rc;
    if ((rc = ensure(c, c->Bop, (size_t)K * 2 * g.Lp * g.Rp * osz))) return rc;
    if ((rc = ensure(c, c->nrm, (size_t)K * sizeof(float)))) return rc;
    if ((rc = ensure(c, c->ctr, 64 * sizeof(unsigned long long)))) return rc;
    if ((rc = upload(c, c->ph32, s.ph32))) return rc;
    if ((rc = upload(c, c->ph64, s.ph64))) return rc;
    if ((rc = upload(c, c->shiftv, s.shifts))) return rc;
    if ((rc = upload(c, c->H, s.H))) return rc;
    unsigned long long* ctr = (unsigned long long*)c->ctr.p;
    // ctr layout (u64): 0 n_small, 1 fp64 candidates, 2 tile counter (as int),
    // 3 uncertain count (int), 4 fallback count (int)
    if ((rc = fill(c, ctr, 2 * 64, 0))) return rc;
    if ((rc = fill(c, c->nrm.p, (int)K, 0))) return rc;

    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    {
        const long long lines = K * g.Lp;
        const int bs = 256;
        const unsigned nb = (unsigned)((lines + bs - 1) / bs);
        if (p->pf_dtype == CLM_C128)
            hipLaunchKernelGGL(prep_kernel<double>, dim3(nb), dim3(bs), 0, c->stream, (const double*)d_pf, g,
                               (const double*)c->H.p, use64 ? (double2*)c->P64.p : nullptr, (float2*)c->P32.p,
                               c->Bop.p, opk, (unsigned*)c->nrm.p, ctr + 0);
        else
            hipLaunchKernelGGL(prep_kernel<float>, dim3(nb), dim3(bs), 0, c->stream, (const float*)d_pf, g,
                               (const double*)c->H.p, use64 ? (double2*)c->P64.p : nullptr, (float2*)c->P32.p,
                               c->Bop.p, opk, (unsigned*)c->nrm.p, ctr + 0);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));

    int n_unc = 0, n_fb = 0;
    if (prec == CLM_FP64_REF) {
        Exact64Args ea{};
        ea.g = g;
        ea.P64 = (const double2*)c->P64.p;
        ea.ph64 = (const double2*)c->ph64.p;
        ea.shiftv = (const double*)c->shiftv.p;
        ea.list = nullptr;
        ea.n = (int)P;
        ea.K = (int)K;
        ea.cl1 = d_cl1; ea.cl2 = d_cl2; ea.jidx = d_j; ea.corr = d_corr; ea.shift1d = d_shift;
        const unsigned nb = (unsigned)std::min<long long>(P, (long long)c->n_cu * 8);
        hipLaunchKernelGGL(exact64_kernel, dim3(nb), dim3(256), 0, c->stream, ea);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    } else {
        // ---- pass 1 tiles: rows of the pair triangle cut into PPW pairs
        const int PPW = v.W / v.CPP;
        if (c->tiles_K != (int)K || c->tiles_PPW != PPW) {
            c->h_tiles.clear();
            for (long long k1 = 0; k1 + 1 < K; k1++)
                for (long long k2 = k1 + 1; k2 < K; k2 += PPW)
                    c->h_tiles.push_back(make_int4((int)k1, (int)k2, (int)std::min<long long>(PPW, K - k2),
                                                   (int)pair_slot(k1, k2, K)));
            if ((rc = upload(c, c->tiles, c->h_tiles))) return rc;
            c->tiles_K = (int)K;
            c->tiles_PPW = PPW;
        }
        const int ntiles = (int)c->h_tiles.size();
        if ((rc = ensure(c, c->recs, (size_t)P * sizeof(PairRec)))) return rc;
        if ((rc = ensure(c, c->ulist, (size_t)P * sizeof(int)))) return rc;

        const float u = opk == OPK_BF16 ? 0x1p-8f : opk == OPK_F16 ? 0x1p-11f : 0x1p-24f;
        const float sub = opk == OPK_F16 ? 0x1p-24f * std::sqrt(2.0f * g.Rp) : 0.0f;
        PairArgs pa{};
        pa.g = g;
        pa.P32 = (const float2*)c->P32.p;
        pa.Bop = c->Bop.p;
        pa.ph32 = (const float2*)c->ph32.p;
        pa.nrm = (const float*)c->nrm.p;
        pa.tiles = (const int4*)c->tiles.p;
        pa.ntiles = ntiles;
        pa.counter = (int*)(ctr + 2);
        pa.plist = nullptr;
        pa.recs = (PairRec*)c->recs.p;
        pa.erecs = nullptr;
        pa.CPP = v.CPP;
        pa.coef = (2.0f * u + u * u + (2.0f * g.Rp + 16.0f) * 0x1p-24f) * 1.01f;
        pa.sub = sub;
        pa.tau_t = use64 ? (float)(0.5 * p->tie_margin) : 0.0f;
        PairFn fn = pick_pair_kernel(opk, v.ct, g.Rp, g.sym != 0, false);
        PairFn fn2 = pick_pair_kernel(opk, v.ct, g.Rp, g.sym != 0, true);
        if (!fn || !fn2) return fail(c, CLM_E_UNSUP, "no kernel for ct=%d Rp=%d", v.ct, g.Rp);
        const int osz_lds = opk == OPK_F32 ? 4 : 2;
        pa.stage = pair_lds_bytes(g.Rp, osz_lds, v.W, g.NE, g.L, 1) <= (size_t)c->lds_max ? 1 : 0;
        const size_t lds = pair_lds_bytes(g.Rp, osz_lds, v.W, g.NE, g.L, pa.stage);
        if (lds > 65536) {
            HIPCHK(c, hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            HIPCHK(c, hipFuncSetAttribute((const void*)fn2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        }
        int occ = 0;
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)fn, 64 * v.W, lds));
        if (occ < 1) occ = 1;
        const unsigned grid = (unsigned)std::min<long long>((long long)c->n_cu * occ, ntiles);
        hipLaunchKernelGGL(fn, dim3(grid), dim3(64 * v.W), lds, c->stream, pa);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream));

        // ---- re-score
        RescoreArgs ra{};
        ra.g = g;
        ra.P32 = (const float2*)c->P32.p;
        ra.P64 = use64 ? (const double2*)c->P64.p : nullptr;
        ra.ph32 = (const float2*)c->ph32.p;
        ra.ph64 = (const double2*)c->ph64.p;
        ra.shiftv = (const double*)c->shiftv.p;
        ra.recs = (const PairRec*)c->recs.p;
        ra.mode = 0;
        ra.n_items = (int)P;
        ra.fp64 = use64 ? 1 : 0;
        ra.tau = (float)p->tie_margin;
        ra.unc_count = (int*)(ctr + 3);
        ra.unc_list = (int*)c->ulist.p;
        ra.fb_count = (int*)(ctr + 4);
        ra.n64 = ctr + 1;
        ra.ncell = ctr + 6;
        ra.nemit = ctr + 7;
        ra.cl1 = d_cl1; ra.cl2 = d_cl2; ra.jidx = d_j; ra.corr = d_corr; ra.shift1d = d_shift;
        const unsigned rgrid = (unsigned)std::min<long long>((P + 3) / 4, (long long)c->n_cu * 16);
        hipLaunchKernelGGL(rescore_kernel, dim3(rgrid), dim3(256), 0, c->stream, ra);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(&n_unc, ctr + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));

        // ---- second pass on uncertain pairs, in batches
        std::vector<int> fb_all;
        if (n_unc > 0) {
            std::vector<int> uslots(n_unc);
            HIPCHK(c, hipMemcpy(uslots.data(), c->ulist.p, sizeof(int) * n_unc, hipMemcpyDeviceToHost));
            std::sort(uslots.begin(), uslots.end());
            const int BATCH = 131072;
            for (int b0 = 0; b0 < n_unc; b0 += BATCH) {
                const int nb = std::min(BATCH, n_unc - b0);
                std::vector<int4> tl, pl;
                std::vector<int> order(nb);
                int cur_k1 = -1, cnt = 0;
                for (int i = 0; i < nb; i++) {
                    int k1, k2;
                    slot_pair_host(uslots[b0 + i], K, k1, k2);
                    if (k1 != cur_k1 || cnt == PPW) {
                        tl.push_back(make_int4(k1, (int)pl.size(), 0, -1));
                        cur_k1 = k1;
                        cnt = 0;
                    }
                    pl.push_back(make_int4(k2, uslots[b0 + i], i, 0));
                    order[i] = uslots[b0 + i];
                    tl.back().z++;
                    cnt++;
                }
                if ((rc = upload(c, c->plist, pl))) return rc;
                if ((rc = upload(c, c->uslots, order))) return rc;
                if ((rc = ensure(c, c->erecs, (size_t)nb * sizeof(EmitRec)))) return rc;
                DevBuf tb;
                if ((rc = upload(c, tb, tl))) return rc;
                if ((rc = fill(c, ctr + 2, 2, 0))) { release(tb); return rc; }
                pa.tiles = (const int4*)tb.p;
                pa.ntiles = (int)tl.size();
                pa.plist = (const int4*)c->plist.p;
                pa.erecs = (EmitRec*)c->erecs.p;
                const unsigned g2 = (unsigned)std::min<long long>((long long)c->n_cu * occ, (long long)tl.size());
                hipLaunchKernelGGL(fn2, dim3(g2), dim3(64 * v.W), lds, c->stream, pa);
                hipError_t le = hipGetLastError();
                if (le != hipSuccess) { release(tb); return fail(c, CLM_E_HIP, "pair pass 2: %s", hipGetErrorString(le)); }
                ra.mode = 1;
                ra.n_items = nb;
                ra.erecs = (const EmitRec*)c->erecs.p;
                ra.ulist = (const int*)c->uslots.p;
                if ((rc = ensure(c, c->fblist, (size_t)n_unc * sizeof(int)))) { release(tb); return rc; }
                ra.fb_list = (int*)c->fblist.p;
                if ((rc = fill(c, ctr + 4, 2, 0))) { release(tb); return rc; }
                const unsigned rg2 = (unsigned)std::min<long long>((nb + 3) / 4, (long long)c->n_cu * 16);
                hipLaunchKernelGGL(rescore_kernel, dim3(rg2), dim3(256), 0, c->stream, ra);
                le = hipGetLastError();
                if (le != hipSuccess) { release(tb); return fail(c, CLM_E_HIP, "rescore pass 2: %s", hipGetErrorString(le)); }
                int nfb = 0;
                HIPCHK(c, hipMemcpyAsync(&nfb, ctr + 4, sizeof(int), hipMemcpyDeviceToHost, c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                release(tb);
                if (nfb > 0) {
                    std::vector<int> fb(nfb);
                    HIPCHK(c, hipMemcpy(fb.data(), c->fblist.p, sizeof(int) * nfb, hipMemcpyDeviceToHost));
                    fb_all.insert(fb_all.end(), fb.begin(), fb.end());
                }
            }
        }
        n_fb = (int)fb_all.size();
        if (n_fb > 0) {
            if (!use64 && (rc = ensure(c, c->P64, (size_t)K * g.L * g.R * sizeof(double2)))) return rc;
            if (!use64) {
                // plain FP32 runs do not keep FP64 rays; rebuild them for the fallback
                const long long lines = K * g.Lp;
                const unsigned nb = (unsigned)((lines + 255) / 256);
                if (p->pf_dtype == CLM_C128)
                    hipLaunchKernelGGL(prep_kernel<double>, dim3(nb), dim3(256), 0, c->stream, (const double*)d_pf, g,
                                       (const double*)c->H.p, (double2*)c->P64.p, (float2*)c->P32.p, c->Bop.p, opk,
                                       (unsigned*)c->nrm.p, ctr + 5);
                else
                    hipLaunchKernelGGL(prep_kernel<float>, dim3(nb), dim3(256), 0, c->stream, (const float*)d_pf, g,
                                       (const double*)c->H.p, (double2*)c->P64.p, (float2*)c->P32.p, c->Bop.p, opk,
                                       (unsigned*)c->nrm.p, ctr + 5);
                HIPCHK(c, hipGetLastError());
            }
            if ((rc = upload(c, c->fblist, fb_all))) return rc;
            Exact64Args ea{};
            ea.g = g;
            ea.P64 = (const double2*)c->P64.p;
            ea.ph64 = (const double2*)c->ph64.p;
            ea.shiftv = (const double*)c->shiftv.p;
            ea.list = (const int*)c->fblist.p;
            ea.n = n_fb;
            ea.K = (int)K;
            ea.cl1 = d_cl1; ea.cl2 = d_cl2; ea.jidx = d_j; ea.corr = d_corr; ea.shift1d = d_shift;
            const unsigned nb = (unsigned)std::min<long long>(n_fb, (long long)c->n_cu * 8);
            hipLaunchKernelGGL(exact64_kernel, dim3(nb), dim3(256), 0, c->stream, ea);
            HIPCHK(c, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st) {
        unsigned long long h[8] = {0};
        HIPCHK(c, hipMemcpy(h, ctr, sizeof h, hipMemcpyDeviceToHost));
        st->small_norm_lines = (int64_t)h[0];
        st->fp64_candidates = (int64_t)h[1];
        st->cells_rescored = (int64_t)h[6];
        st->emitted = (int64_t)h[7];
        st->uncertain_pairs = n_unc;
        st->fallback_pairs = prec == CLM_FP64_REF ? P : n_fb;
        st->ms_prep = event_ms(c->ev[0], c->ev[1]);
        st->ms_pairs = event_ms(c->ev[1], c->ev[2]);
        st->ms_rescore = event_ms(c->ev[2], c->ev[3]);
        st->ms_total = event_ms(c->ev[0], c->ev[3]);
    }
    return CLM_OK;
}

}  // namespace

extern "C" {

int clm_version(void) { return CLM_VERSION; }

void clm_default_params(clm_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->pf_dtype = CLM_C128;
    p->max_shift = 15.0;   // cryo_clmatrix_cpu.m:133
    p->shift_step = 1.0;   // :138
    p->precision = CLM_BF16_RESCORE;
    p->verbose = 1;        // :129
    p->tie_margin = 1e-3;
}

int64_t clm_num_pairs(int32_t K) { return K < 2 ? 0 : (int64_t)K * (K - 1) / 2; }

clm_ctx* clm_create(int device, int* err) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        if (err) *err = CLM_E_NODEV;
        return nullptr;
    }
    if (device < 0 || device >= n) {
        if (err) *err = CLM_E_ARG;
        return nullptr;
    }
    clm_ctx* c = new clm_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        if (err) *err = CLM_E_HIP;
        delete c;
        return nullptr;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->n_cu = prop.multiProcessorCount;
        if (prop.sharedMemPerBlock > 0) c->lds_max = (int)prop.sharedMemPerBlock;
    }
    if (c->n_cu <= 0) c->n_cu = 256;
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    if (err) *err = CLM_OK;
    return c;
}

void clm_destroy(clm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->P64, &c->P32, &c->Bop, &c->nrm, &c->recs, &c->tiles, &c->plist, &c->ctr,
                      &c->ulist, &c->erecs, &c->fblist, &c->ph32, &c->ph64, &c->shiftv, &c->H, &c->pf,
                      &c->out_cl1, &c->out_cl2, &c->out_j, &c->out_corr, &c->out_shift, &c->uslots};
    for (DevBuf* b : bufs) release(*b);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* clm_last_error(const clm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int clm_clmatrix_device(clm_ctx* c, const clm_params* p, const void* d_pf, int32_t* d_cl1, int32_t* d_cl2,
                        int32_t* d_jidx, double* d_corr, double* d_shift1d, clm_stats* st) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    if (!d_pf && p->K > 1) return fail(c, CLM_E_ARG, "null pf");
    HIPCHK(c, hipSetDevice(c->device));
    return run_device(c, p, d_pf, d_cl1, d_cl2, d_jidx, d_corr, d_shift1d, st);
}

int clm_clmatrix(clm_ctx* c, const clm_params* p, const void* pf, int32_t* cl1, int32_t* cl2, int32_t* jidx,
                 double* corr, double* shift1d, clm_stats* st) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    if (p->n_theta % 2) return fail(c, CLM_E_ARG, "n_theta must be even");
    if (p->K < 0 || p->n_r < 0 || p->n_theta < 0) return fail(c, CLM_E_ARG, "negative size");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t esz = p->pf_dtype == CLM_C64 ? 8 : 16;
    const size_t pf_bytes = (size_t)p->n_r * p->n_theta * p->K * esz;
    const long long P = clm_num_pairs(p->K);
    int rc;
    if ((rc = ensure(c, c->pf, pf_bytes))) return rc;
    if ((rc = ensure(c, c->out_cl1, P * 4))) return rc;
    if ((rc = ensure(c, c->out_cl2, P * 4))) return rc;
    if ((rc = ensure(c, c->out_j, P * 4))) return rc;
    if ((rc = ensure(c, c->out_corr, P * 8))) return rc;
    if ((rc = ensure(c, c->out_shift, P * 8))) return rc;
    if (pf_bytes) HIPCHK(c, hipMemcpyAsync(c->pf.p, pf, pf_bytes, hipMemcpyHostToDevice, c->stream));
    rc = run_device(c, p, c->pf.p, (int32_t*)c->out_cl1.p, (int32_t*)c->out_cl2.p, (int32_t*)c->out_j.p,
                    (double*)c->out_corr.p, (double*)c->out_shift.p, st);
    if (rc) return rc;
    if (P > 0) {
        if (cl1) HIPCHK(c, hipMemcpyAsync(cl1, c->out_cl1.p, P * 4, hipMemcpyDeviceToHost, c->stream));
        if (cl2) HIPCHK(c, hipMemcpyAsync(cl2, c->out_cl2.p, P * 4, hipMemcpyDeviceToHost, c->stream));
        if (jidx) HIPCHK(c, hipMemcpyAsync(jidx, c->out_j.p, P * 4, hipMemcpyDeviceToHost, c->stream));
        if (corr) HIPCHK(c, hipMemcpyAsync(corr, c->out_corr.p, P * 8, hipMemcpyDeviceToHost, c->stream));
        if (shift1d) HIPCHK(c, hipMemcpyAsync(shift1d, c->out_shift.p, P * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return CLM_OK;
}

int clm_assemble(int32_t K, const int32_t* cl1, const int32_t* cl2, const double* corr, double* clstack,
                 double* corrstack, double* eqmap) {
    if (K < 0) return CLM_E_ARG;
    const size_t KK = (size_t)K * K;
    if (clstack) std::memset(clstack, 0, KK * sizeof(double));
    if (corrstack) std::memset(corrstack, 0, KK * sizeof(double));
    if (eqmap) std::memset(eqmap, 0, KK * sizeof(double));
    long long e = 0;
    for (long long k1 = 0; k1 < K; k1++)
        for (long long k2 = k1 + 1; k2 < K; k2++, e++) {
            const size_t a = (size_t)(k1 + k2 * K), b = (size_t)(k2 + k1 * K);  // (k1,k2), (k2,k1)
            if (clstack) {
                clstack[a] = cl1 ? cl1[e] : 0;
                clstack[b] = cl2 ? cl2[e] : 0;
            }
            if (corrstack && corr) corrstack[a] = corr[e] != 0.0 ? 1.0 - corr[e] : 0.0;  // :764
            if (eqmap) eqmap[a] = (double)(e + 1);
        }
    return CLM_OK;
}

int clm_shift_equations_csc(int32_t K, int32_t n_theta, const int32_t* cl1, const int32_t* cl2,
                            const double* shift1d, int64_t* jc, int64_t* ir, double* pr, int64_t* nnz) {
    if (K < 0 || n_theta <= 0 || n_theta % 2 || !cl1 || !cl2 || !shift1d) return CLM_E_ARG;
    const int L = n_theta / 2;
    const double dtheta = M_PI / (double)L;  // :245
    // sin/cos of (cl-1)*dtheta and of (cl-1)*dtheta - pi for every possible cl
    std::vector<double> sa(2 * L + 1), ca(2 * L + 1), sb(2 * L + 1), cb(2 * L + 1);
    std::vector<char> lowb(2 * L + 1);
    for (int cl = 0; cl <= 2 * L; cl++) {
        const double a = (double)(cl - 1) * dtheta;
        sa[cl] = std::sin(a);
        ca[cl] = std::cos(a);
        double b = a;
        lowb[cl] = b < M_PI;
        if (!lowb[cl]) b = b - M_PI;
        sb[cl] = std::sin(b);
        cb[cl] = std::cos(b);
    }
    // coefficient of pair e on (image k1: x,y) and (image k2: x,y)   :731-738
    auto coefs = [&](long long e, double out[4]) {
        const int a = cl1[e], b = cl2[e];
        if (a < 0 || a > 2 * L || b < 0 || b > 2 * L) { out[0] = out[1] = out[2] = out[3] = NAN; return; }
        if (lowb[b]) { out[0] = sa[a]; out[1] = ca[a]; }
        else { out[0] = -sa[a]; out[1] = -ca[a]; }
        out[2] = -sb[b];
        out[3] = -cb[b];
    };
    const long long P = clm_num_pairs(K);
    auto slot = [&](long long k1, long long k2) { return k1 * K - k1 * (k1 + 1) / 2 + (k2 - k1 - 1); };
    const bool fill_pass = ir != nullptr && pr != nullptr;
    long long n = 0;
    if (jc) jc[0] = 0;
    for (long long k = 0; k < K; k++) {
        for (int xy = 0; xy < 2; xy++) {
            for (long long j = 0; j < k; j++) {  // rows where k is the second image
                const long long e = slot(j, k);
                double c4[4];
                coefs(e, c4);
                const double v = c4[2 + xy];
                if (v != 0.0) {
                    if (fill_pass) { ir[n] = e; pr[n] = v; }
                    n++;
                }
            }
            for (long long m = k + 1; m < K; m++) {  // rows where k is the first image
                const long long e = slot(k, m);
                double c4[4];
                coefs(e, c4);
                const double v = c4[xy];
                if (v != 0.0) {
                    if (fill_pass) { ir[n] = e; pr[n] = v; }
                    n++;
                }
            }
            if (jc) jc[2 * k + xy + 1] = n;
        }
    }
    for (long long e = 0; e < P; e++) {  // last column: right-hand side
        if (shift1d[e] != 0.0) {
            if (fill_pass) { ir[n] = e; pr[n] = shift1d[e]; }
            n++;
        }
    }
    if (jc) jc[2 * (long long)K + 1] = n;
    if (nnz) *nnz = n;
    return CLM_OK;
}

}  // extern "C"
