// clm_api.hip -- C ABI of libclm (declared in include/clm.h): context,
// geometry, orchestration of the device passes, host-side FP64 assembly.
#include "clm.h"
#include "clm_ctx.h"
#include "clm_kernels.hip"
#include "clm_clmap.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace clm;
using namespace clmh;

namespace {

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
        if (opk == OPK_F32 && ct == 2 && Rp > 32) continue;  // register budget
        if (opk != OPK_F32 && ct == 3 && Rp > 64) continue;  // register budget
        const int NCTp = (NCT + ct - 1) / ct * ct;
        const int CPP = NCTp / ct;
        if (CPP > 8) continue;
        int W = want_W ? want_W : (8 / CPP) * CPP;
        if (opk == OPK_F32 && !want_W) W = CPP * ((4 + CPP - 1) / CPP);  // at least 4 waves
        if (W < CPP || W % CPP || W > 8) continue;
        const double waste = (double)NCTp / NCT;
        if (waste < best_waste - 1e-9) {
            best_waste = waste;
            best.ct = ct; best.W = W; best.CPP = CPP; best.NCTp = NCTp;
        }
    }
    return best;
}

PairFn pick_pair_kernel(int opk, int ct, int Rp, bool sym, bool emit) {
    if (opk == OPK_F32) return sym ? pick_pair_f32_1(ct, Rp / 2, emit) : pick_pair_f32_0(ct, Rp / 2, emit);
    if (opk == OPK_BF16) return sym ? pick_pair_bf16_1(ct, Rp / 16, emit) : pick_pair_bf16_0(ct, Rp / 16, emit);
    return sym ? pick_pair_f16_1(ct, Rp / 16, emit) : pick_pair_f16_0(ct, Rp / 16, emit);
}

AstkFn pick_astk(int opk, int Rp, bool sym) {
    if (opk == OPK_F32) return sym ? pick_astk_f32_1(Rp / 2) : pick_astk_f32_0(Rp / 2);
    if (opk == OPK_BF16) return sym ? pick_astk_bf16_1(Rp / 16) : pick_astk_bf16_0(Rp / 16);
    return sym ? pick_astk_f16_1(Rp / 16) : pick_astk_f16_0(Rp / 16);
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

float event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

// Largest number of pairs given a second (exact-threshold) pass per group
// and call; uncertain pairs beyond it go to the FP64 exhaustive kernel.
constexpr int UCAP = 32768;

// 64-bit counters per group after the 16 globals (see enqueue_job)
constexpr int GSTRIDE = 8;
// pass-1 work queues, one per XCD (workgroup b runs on XCD b % 8)
constexpr int NQ = 8;

// Split k1-major pass-1 tiles into NQ contiguous runs of about equal pair
// counts and append the NQ+1 run bounds (ints, padded to whole int4s) after
// the tiles: the pair kernel reads them at tiles + ntiles.
void append_queue_bounds(std::vector<int4>& h) {
    const int ntiles = (int)h.size();
    long long tot = 0;
    for (int i = 0; i < ntiles; i++) tot += h[i].z & 15;
    int b[NQ + 1];
    b[0] = 0;
    long long acc = 0;
    int q = 1;
    for (int i = 0; i < ntiles && q < NQ; i++) {
        acc += h[i].z & 15;
        while (q < NQ && acc * NQ >= tot * q) b[q++] = i + 1;
    }
    while (q < NQ) b[q++] = ntiles;
    b[NQ] = ntiles;
    static_assert(NQ == 8, "bounds packing");
    h.push_back(make_int4(b[0], b[1], b[2], b[3]));
    h.push_back(make_int4(b[4], b[5], b[6], b[7]));
    h.push_back(make_int4(b[8], 0, 0, 0));
}

// FP64 exhaustive search of items (pairs): the first EX_CAP items split their
// shift grid over up to EX_SJ workgroups each (few items still fill the
// GPU), the rest take one workgroup per item.
constexpr long long EX_CAP = 4096;
constexpr int EX_SJ = 64;

// One stack of a call (one cryo_clmatrix problem).
struct StackJob {
    clm_params p;
    clm_stack_io io;
    Setup s;
    int opk = 0, exact = 0, use64 = 0;
    Variant v;
    int lag = 0;
    long long P = 0;
    int group = -1;
    const double* lsh = nullptr;     // per-line pre-shift [K][L] (windowed search) or null
    int force64 = 0;                 // FP64 near-tie re-score even for FP32 operands (refinement)
    int twin_src = -1;               // FP32 twin: table index of the stack it refines
    int filt = 0;                    // map_filter_radius (> 0: FP64 search of filtered maps)
    // selected pairs only (device-built): list of (stack, slot), count, pass-1 tiles
    const SlotRef* sel_list = nullptr;
    const int* sel_count = nullptr;
    const int4* sel_tiles = nullptr;
    const int4* sel_plist = nullptr;
    const int* sel_ntiles = nullptr;
};

// Stacks that share one pair-kernel instantiation (or the FP64 kernel).
struct Group {
    int opk = 0, exact = 0, ct = 0, Rp = 0, sym = 0, W = 0, CPP = 0, lag = 0;
    int sel = 0;                // one stack of selected pairs (device-built tiles and list)
    int filt = 0;               // (FP64 groups) map_filter_radius of its stacks
    int norefine = 0;           // no FP32 twin: uncertain pairs take the second operand-type pass
    int twin = -1;              // index of the FP32 twin group refining this group's uncertain pairs
    int twin_of = -1;           // (twin groups) the group it refines
    std::vector<int> members;   // job indices
    int first = 0;              // first table index
    long long P = 0;            // pairs of all members
    long long ubase = 0;        // offset of its uncertain list
    PairFn fn = nullptr, fn2 = nullptr;
    size_t lds = 0;
    int occ = 1;
    const TileCache* tc = nullptr;
};

struct Plan {
    std::vector<StackJob> jobs;
    std::vector<Group> groups;
    std::vector<int> order;      // table index -> job index
    std::vector<StackDev> table;
    int sb_base = 0;             // first per-stack buffer set of the context used
    long long tw_max = 0;        // largest group refined by a twin (pairs)
    long long P = 0;
    int n_unc_cap = 0;           // entries of the shared pass-2 buffers
    // FP32 twins built only when their group has uncertain pairs (runs with
    // a host wait between the phases; a captured job builds them up front)
    bool twin_lazy = false;
    bool twins_built = false;
};

int enqueue_fill(clm_ctx* c, void* p, long long n, int v) {
    if (n <= 0) return CLM_OK;
    hipLaunchKernelGGL(fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (int*)p, (int)n, v);
    HIPCHK(c, hipGetLastError());
    return CLM_OK;
}

// Validate one stack and derive its geometry and kernel variant.
// Shape key of the autotuner's table: what the kernel variants' speed depends on.
unsigned long long tune_key(int opk, const Geom& g) {
    return (unsigned long long)opk | ((unsigned long long)g.L << 2) | ((unsigned long long)g.Rp << 16) |
           ((unsigned long long)g.NE << 24) | ((unsigned long long)g.sym << 40);
}

int plan_stack(clm_ctx* c, const clm_params* p, StackJob& j) {
    int rc = setup(c, p, j.s);
    if (rc) return rc;
    Geom& g = j.s.g;
    const long long K = g.K;
    j.P = K * (K - 1) / 2;
    if (j.P > 0x7fffffffLL) return fail(c, CLM_E_UNSUP, "K too large (%lld pairs)", j.P);
    const int prec = p->precision;
    if (prec < CLM_FP32 || prec > CLM_FP64_REF) return fail(c, CLM_E_ARG, "unknown precision %d", prec);
    if (p->pf_dtype != CLM_C128 && p->pf_dtype != CLM_C64) return fail(c, CLM_E_ARG, "unknown pf dtype %d", p->pf_dtype);
    if (g.Rp < 0) return fail(c, CLM_E_UNSUP, "n_r=%d exceeds the supported 129", g.n_r);
    {
        // The margin sets both the pass-1 keep window (tie_margin/2 beyond the
        // certified bound) and the FP64 re-score window below the FP32 best.
        // Below a few times the FP32 re-score's own error (rays are unit norm,
        // so (2Rp+16)·2^-24 in correlation units) the FP64 stage could miss
        // the true maximum; NaN or negative would re-score nothing.
        const double floor_tm = 4.0 * (2.0 * g.Rp + 16.0) * 0x1p-24;
        if (!(p->tie_margin >= floor_tm) || !std::isfinite(p->tie_margin))
            return fail(c, CLM_E_ARG, "tie_margin %g is below the FP32 re-score error floor %.3g (or not finite)",
                        p->tie_margin, floor_tm);
    }
    j.opk = prec == CLM_BF16_RESCORE ? OPK_BF16 : prec == CLM_FP16_RESCORE ? OPK_F16 : OPK_F32;
    j.exact = prec == CLM_FP64_REF;
    j.use64 = prec != CLM_FP32;
    const int NCT = (g.L + 31) / 32;
    int want_ct = p->variant / 100, want_W = p->variant % 100, want_lag = 1;
    if (p->variant == 0) {
        const unsigned long long key = tune_key(j.opk, g);
        for (const auto& e : c->tuned)
            if (e.first == key) { want_ct = e.second.x; want_W = e.second.y; want_lag = e.second.z; }
    }
    j.v = choose_variant(j.opk, NCT, g.Rp, want_ct, want_W);
    if (!j.v.ct) return fail(c, CLM_E_UNSUP, "no kernel variant for L=%d (variant %d)", g.L, p->variant);
    g.NCT = j.v.NCTp;
    g.Lp = 32 * j.v.NCTp;
    g.NROWS = g.NE * g.L;
    g.NAT = (g.NROWS + 31) / 32;
    if ((long long)g.NROWS * g.Lp > 0xffffffffLL) return fail(c, CLM_E_UNSUP, "map too large");
    j.lag = (j.v.W % 2 == 0 && want_lag && !(p->reserved & 1)) ? 1 : 0;
    return CLM_OK;
}

// Allocate and upload everything the device job reads; build the table.
// Nothing here is captured: the job itself (enqueue_job) only launches kernels.
int prepare(clm_ctx* c, Plan& pl, int n, const clm_params* ps, const clm_stack_io* ios, bool build_tiles = true,
            const StackJob* extra = nullptr, int sb_base = 0, bool twins = true) {
    int rc;
    pl.jobs.resize(n);
    pl.sb_base = sb_base;
    for (int i = 0; i < n; i++) {
        StackJob& j = pl.jobs[i];
        if (extra) j = extra[i];
        j.p = ps[i];
        j.io = ios[i];
        if ((rc = plan_stack(c, &ps[i], j))) {
            if (n > 1) c->err = "stack " + std::to_string(i) + ": " + c->err;
            return rc;
        }
        if (j.filt > 0) {  // filtered maps: the FP64 search of every entry
            // cryo_average_clmap's fft2(W, 2L, 2L) would truncate a wider weight square
            if (2 * j.filt + 1 > 2 * j.s.g.L)
                return fail(c, CLM_E_UNSUP, "map_filter_radius %d: the (2d+1)-square exceeds the 2L = %d map", j.filt,
                            2 * j.s.g.L);
            j.exact = 1;
            j.use64 = 1;
        }
        if (j.P > 0 && !j.io.pf) return fail(c, CLM_E_ARG, "null pf (stack %d)", i);
    }
    // groups: same kernel instantiation (pairs-free stacks need none)
    for (int i = 0; i < n; i++) {
        StackJob& j = pl.jobs[i];
        if (j.P == 0) continue;
        const Geom& g = j.s.g;
        const int lag = j.lag;
        const int norefine = (j.p.reserved & 2) ? 1 : 0;
        int gi = -1;
        for (size_t q = 0; q < pl.groups.size() && !j.sel_list; q++) {
            const Group& G = pl.groups[q];
            if (!G.sel && G.exact == j.exact && G.filt == j.filt && G.norefine == norefine &&
                (j.exact || (G.opk == j.opk && G.ct == j.v.ct && G.Rp == g.Rp && G.sym == g.sym && G.W == j.v.W &&
                             G.CPP == j.v.CPP && G.lag == lag))) {
                gi = (int)q;
                break;
            }
        }
        if (gi < 0) {
            Group G;
            G.opk = j.opk; G.exact = j.exact; G.ct = j.v.ct; G.Rp = g.Rp; G.sym = g.sym;
            G.W = j.v.W; G.CPP = j.v.CPP; G.lag = lag;
            G.sel = j.sel_list != nullptr;
            G.norefine = norefine;
            G.filt = j.filt;
            pl.groups.push_back(G);
            gi = (int)pl.groups.size() - 1;
        }
        pl.groups[gi].members.push_back(i);
        j.group = gi;
    }
    // FP32 twins: for every BF16/FP16 group, FP32 copies of its stacks (same
    // rays, records and outputs; FP32 operand planes and A-stacks) that re-run
    // the group's uncertain pairs with a 16x tighter error bound.  A broad
    // correlation peak (clean, high-SNR data) leaves many candidates inside
    // the BF16 bound, and this keeps such pairs off the FP64 exhaustive path.
    const size_t n_orig_groups = pl.groups.size();
    for (size_t gi = 0; gi < n_orig_groups && twins; gi++) {
        if (pl.groups[gi].exact || pl.groups[gi].opk == OPK_F32 || pl.groups[gi].norefine) continue;
        const StackJob& j0 = pl.jobs[pl.groups[gi].members[0]];
        const Variant v32 = choose_variant(OPK_F32, (j0.s.g.L + 31) / 32, j0.s.g.Rp, 0, 0);
        if (!v32.ct) continue;
        Group T;
        T.opk = OPK_F32; T.ct = v32.ct; T.Rp = pl.groups[gi].Rp; T.sym = pl.groups[gi].sym;
        T.W = v32.W; T.CPP = v32.CPP; T.lag = (v32.W % 2 == 0) ? 1 : 0;
        T.sel = 1;
        T.twin_of = (int)gi;
        for (int i : pl.groups[gi].members) {
            StackJob t = pl.jobs[i];
            t.opk = OPK_F32;
            t.v = v32;
            t.s.g.NCT = v32.NCTp;
            t.s.g.Lp = 32 * v32.NCTp;
            t.force64 = t.use64;
            t.twin_src = i;  // job index for now; table index below
            t.group = (int)pl.groups.size();
            pl.jobs.push_back(t);
            T.members.push_back((int)pl.jobs.size() - 1);
        }
        pl.groups[gi].twin = (int)pl.groups.size();
        pl.groups.push_back(T);
    }
    // table in group order (twins last); pair_base relative to the group
    pl.order.clear();
    pl.P = 0;
    long long ubase = 0;
    std::vector<int> table_of(pl.jobs.size(), -1);
    for (Group& G : pl.groups) {
        G.first = (int)pl.order.size();
        G.P = 0;
        for (int i : G.members) {
            table_of[i] = (int)pl.order.size();
            pl.order.push_back(i);
            G.P += pl.jobs[i].P;
        }
        G.ubase = ubase;
        if (!G.exact) ubase += G.P;
        if (G.twin_of < 0) pl.P += G.P;
    }
    pl.tw_max = 0;
    for (StackJob& j : pl.jobs)
        if (j.twin_src >= 0) j.twin_src = table_of[j.twin_src];
    for (const Group& G : pl.groups)
        if (G.twin_of >= 0) pl.tw_max = std::max(pl.tw_max, G.P);
    const int nt = (int)pl.order.size();
    if (c->sb.size() < (size_t)(sb_base + nt)) c->sb.resize(sb_base + nt);
    pl.table.assign(nt, StackDev{});
    for (int ti = 0; ti < nt; ti++) {
        StackJob& j = pl.jobs[pl.order[ti]];
        StackBufs& B = c->sb[pl.sb_base + ti];
        const Geom& g = j.s.g;
        const long long K = g.K;
        const size_t osz = j.opk == OPK_F32 ? 4 : 2;
        if (j.twin_src >= 0) {
            // FP32 twin: own operand planes and A-stacks (allocated here, or
            // by build_twins when they are first needed), everything else shared
            if (!pl.twin_lazy) {
                if ((rc = ensure(c, B.Bop, (size_t)K * 2 * g.Lp * g.Rp * 4))) return rc;
                if ((rc = ensure(c, B.astk, (size_t)K * g.NAT * (32 * 2 * (size_t)g.Rp * 4)))) return rc;
            }
            StackDev& d = pl.table[ti];
            d = pl.table[j.twin_src];
            d.g = g;
            d.astk = pl.twin_lazy ? nullptr : B.astk.p;
            d.Bop = pl.twin_lazy ? nullptr : B.Bop.p;
            d.coef = (2.0f * 0x1p-24f + 0x1p-48f + (2.0f * g.Rp + 16.0f) * 0x1p-24f) * 1.01f;
            d.sub = 0.0f;
            d.tau_t = j.force64 ? (float)(0.5 * j.p.tie_margin) : 0.0f;
            d.fp64 = j.force64;
            d.CPP = j.v.CPP;
            continue;
        }
        if ((rc = ensure(c, B.P64, (size_t)K * g.L * g.R * sizeof(double2)))) return rc;
        if ((rc = ensure(c, B.P32, (size_t)K * g.L * g.Rp * sizeof(float2)))) return rc;
        if ((rc = ensure(c, B.Bop, (size_t)K * 2 * g.Lp * g.Rp * osz))) return rc;
        if ((rc = ensure(c, B.nrm, (size_t)K * sizeof(float)))) return rc;
        if (!j.exact) {
            const size_t tile_bytes = 32 * 2 * (size_t)g.Rp * osz;
            if ((rc = ensure(c, B.astk, (size_t)K * g.NAT * tile_bytes))) return rc;
            if ((rc = ensure(c, B.recs, (size_t)j.P * sizeof(PairRec)))) return rc;
        }
        if ((rc = upload(c, B.ph32, j.s.ph32))) return rc;
        if ((rc = upload(c, B.ph64, j.s.ph64))) return rc;
        if ((rc = upload(c, B.shiftv, j.s.shifts))) return rc;
        if ((rc = upload(c, B.H, j.s.H))) return rc;
        if (j.filt > 0) {
            // W0 = 1.25.^(21:-1:1); W(i,j) = W0(1 + |i-d| + |j-d|)  (cryo_average_clmap.m:24-26)
            const int w = 2 * j.filt + 1;
            std::vector<double> wv((size_t)w * w);
            for (int a = 0; a < w; a++)
                for (int b = 0; b < w; b++)
                    wv[(size_t)a * w + b] = std::pow(1.25, (double)(21 - (std::abs(a - j.filt) + std::abs(b - j.filt))));
            if ((rc = upload(c, B.W, wv))) return rc;
        }
        StackDev& d = pl.table[ti];
        d.g = g;
        d.astk = B.astk.p;
        d.Bop = B.Bop.p;
        d.nrm = (const float*)B.nrm.p;
        d.recs = (PairRec*)B.recs.p;
        d.P32 = (const float2*)B.P32.p;
        d.P64 = (const double2*)B.P64.p;
        d.ph32 = (const float2*)B.ph32.p;
        d.ph64 = (const double2*)B.ph64.p;
        d.shiftv = (const double*)B.shiftv.p;
        d.cl1 = j.io.cl1; d.cl2 = j.io.cl2; d.jidx = j.io.jidx; d.corr = j.io.corr; d.shift1d = j.io.shift1d;
        const float u = j.opk == OPK_BF16 ? 0x1p-8f : j.opk == OPK_F16 ? 0x1p-11f : 0x1p-24f;
        d.coef = (2.0f * u + u * u + (2.0f * g.Rp + 16.0f) * 0x1p-24f) * 1.01f;
        d.sub = j.opk == OPK_F16 ? 0x1p-24f * std::sqrt(2.0f * g.Rp) : 0.0f;
        d.tau_t = (j.use64 || j.force64) ? (float)(0.5 * j.p.tie_margin) : 0.0f;
        d.tau = (float)j.p.tie_margin;
        d.fp64 = j.use64 || j.force64;
        d.CPP = j.v.CPP;
        d.npairs = j.P;
        d.filt = j.filt;
        d.wts = j.filt > 0 ? (const double*)B.W.p : nullptr;
    }
    for (Group& G : pl.groups) {
        long long base = 0;
        for (int ti = G.first; ti < G.first + (int)G.members.size(); ti++) {
            pl.table[ti].pair_base = base;
            base += pl.table[ti].npairs;
        }
    }
    if ((rc = upload(c, c->table, pl.table))) return rc;
    // shared buffers: counters (16 globals + GSTRIDE per group), lists, pass-2 tiles
    if ((rc = ensure(c, c->ctr, (16 + GSTRIDE * pl.groups.size()) * sizeof(unsigned long long)))) return rc;
    if ((rc = ensure(c, c->ulist, (size_t)std::max<long long>(ubase, 1) * sizeof(SlotRef)))) return rc;
    if ((rc = ensure(c, c->fblist, (size_t)std::max<long long>(pl.P, 1) * sizeof(SlotRef)))) return rc;
    long long gmax = 0;
    for (const Group& G : pl.groups)
        if (!G.exact) gmax = std::max(gmax, G.P);
    pl.n_unc_cap = (int)std::min<long long>(gmax, UCAP);
    if ((rc = ensure(c, c->erecs, (size_t)std::max(pl.n_unc_cap, 1) * sizeof(EmitRec)))) return rc;
    if ((rc = ensure(c, c->plist, (size_t)std::max(pl.n_unc_cap, 1) * sizeof(int4)))) return rc;
    if ((rc = ensure(c, c->tiles2, (size_t)std::max(pl.n_unc_cap, 1) * sizeof(int4)))) return rc;
    if ((rc = ensure(c, c->tw_tiles, (size_t)std::max<long long>(pl.tw_max, 1) * sizeof(int4)))) return rc;
    if ((rc = ensure(c, c->ex_v, (size_t)EX_CAP * EX_SJ * sizeof(double)))) return rc;
    if ((rc = ensure(c, c->ex_k, (size_t)EX_CAP * EX_SJ * sizeof(long long)))) return rc;
    if ((rc = ensure(c, c->tw_plist, (size_t)std::max<long long>(pl.tw_max, 1) * sizeof(int4)))) return rc;
    // kernels and pass-1 tiles per group
    for (Group& G : pl.groups) {
        if (G.exact) continue;
        G.fn = pick_pair_kernel(G.opk, G.ct, G.Rp, G.sym != 0, false);
        G.fn2 = pick_pair_kernel(G.opk, G.ct, G.Rp, G.sym != 0, true);
        if (!G.fn || !G.fn2) return fail(c, CLM_E_UNSUP, "no kernel for ct=%d Rp=%d", G.ct, G.Rp);
        if (G.W < 4 || G.W > 8) return fail(c, CLM_E_UNSUP, "workgroup of %d waves", G.W);
        G.lds = pair_lds_bytes(G.W);
        G.occ = 0;
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&G.occ, (const void*)G.fn, 64 * G.W, G.lds));
        if (G.occ < 1) G.occ = 1;
        if (!build_tiles || G.sel) continue;
        const int PPW = G.W / G.CPP;
        std::vector<long long> key;
        for (int ti = G.first; ti < G.first + (int)G.members.size(); ti++) {
            key.push_back(ti);
            key.push_back(pl.table[ti].g.K);
            key.push_back(PPW);
        }
        bool hit = false;
        for (const TileCache& tc : c->tcache) hit = hit || tc.key == key;
        if (!hit) {
            std::vector<int4> h;
            for (int ti = G.first; ti < G.first + (int)G.members.size(); ti++) {
                const long long K = pl.table[ti].g.K;
                for (long long k1 = 0; k1 + 1 < K; k1++)
                    for (long long k2 = k1 + 1; k2 < K; k2 += PPW)
                        h.push_back(make_int4((int)k1, (int)k2, (int)std::min<long long>(PPW, K - k2) | (ti << 4),
                                              (int)pair_slot(k1, k2, K)));
            }
            if (c->tcache.size() >= std::max<size_t>(4, pl.groups.size())) {
                if (!c->arena) release(c->tcache.back().tiles);
                c->tcache.pop_back();
            }
            c->tcache.insert(c->tcache.begin(), TileCache{});
            TileCache& tc = c->tcache.front();
            tc.key = key;
            tc.ntiles = (int)h.size();
            append_queue_bounds(h);
            if ((rc = upload(c, tc.tiles, h))) return rc;
            HIPCHK(c, hipStreamSynchronize(c->stream));  // h is released on return
        }
    }
    // tile caches can move while inserting: resolve pointers afterwards
    for (Group& G : pl.groups) {
        if (G.exact || !build_tiles || G.sel) continue;
        const int PPW = G.W / G.CPP;
        std::vector<long long> key;
        for (int ti = G.first; ti < G.first + (int)G.members.size(); ti++) {
            key.push_back(ti);
            key.push_back(pl.table[ti].g.K);
            key.push_back(PPW);
        }
        for (const TileCache& tc : c->tcache)
            if (tc.key == key) G.tc = &tc;
        if (!G.tc) return fail(c, CLM_E_UNSUP, "tile cache too small for %zu groups", pl.groups.size());
    }
    return CLM_OK;
}

// Per-stack preprocessing (lines, filter, normalization, operand planes) and
// A-stacks.
// An FP32 twin's operand planes and A-stacks, from its source stack's P32
// (which the source's prep has written): A-stacks of all images, or of the
// nimg images listed in imgs (device) in that order.
int enqueue_twin_prep(clm_ctx* c, const Plan& pl, size_t ti, const int* imgs = nullptr, int nimg = -1) {
    const StackJob& j = pl.jobs[pl.order[ti]];
    const StackBufs& B = c->sb[pl.sb_base + ti];
    const Geom& g = j.s.g;
    const StackDev& src = pl.table[j.twin_src];
    if (nimg < 0) nimg = g.K;
    const long long nb32 = (long long)g.K * 2 * g.Lp * g.Rp;
    hipLaunchKernelGGL(bop32_kernel, dim3((unsigned)((nb32 + 255) / 256)), dim3(256), 0, c->stream, g, src.P32,
                       (float*)B.Bop.p);
    HIPCHK(c, hipGetLastError());
    const long long nch = (long long)nimg * g.NAT * (long long)(32 * 2 * (size_t)g.Rp * 4 / 16);
    AstkFn af = pick_astk(OPK_F32, g.Rp, g.sym != 0);
    if (!af) return fail(c, CLM_E_UNSUP, "no FP32 A-stack kernel for Rp=%d", g.Rp);
    if (nch > 0) {
        hipLaunchKernelGGL(af, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, c->stream, g, src.P32, src.ph32,
                           (int4*)B.astk.p, imgs, nimg);
        HIPCHK(c, hipGetLastError());
    }
    return CLM_OK;
}

// Lazy twins, once some group turned out to need them: the images k1 of the
// uncertain pairs (the twins re-run only those pairs, whose A side is k1)
// get FP32 A-stacks, the device table maps every image to its A-stack
// (-1: none), and the twins' operand planes are built for all images.
// ctr: the counters as read back after the re-score.
int build_twins(clm_ctx* c, Plan& pl, const std::vector<unsigned long long>& ctr) {
    int rc;
    // uncertain pairs of every group refined by a twin, by table index
    std::vector<std::vector<int>> k1s(pl.order.size());
    for (size_t gi = 0; gi < pl.groups.size(); gi++) {
        const Group& G = pl.groups[gi];
        if (G.twin < 0) continue;
        const int n = (int)(unsigned)(ctr[16 + GSTRIDE * gi + 1] & 0xffffffffu);
        if (n <= 0) continue;
        std::vector<SlotRef> ul((size_t)std::min<long long>(n, G.P));
        HIPCHK(c, hipMemcpy(ul.data(), (const SlotRef*)c->ulist.p + G.ubase, ul.size() * sizeof(SlotRef),
                            hipMemcpyDeviceToHost));
        for (const SlotRef& r : ul) {
            const long long K = pl.table[r.stack].g.K;
            // slot -> k1 (pairs k1 < k2 in row-major order)
            long long k1 = (long long)std::floor(((2.0 * K - 1.0) -
                                                  std::sqrt((2.0 * K - 1.0) * (2.0 * K - 1.0) - 8.0 * (double)r.slot)) / 2.0);
            k1 = std::max(0LL, std::min(k1, K - 2));
            while (k1 > 0 && pair_slot(k1, k1 + 1, K) > r.slot) k1--;
            while (k1 + 1 < K - 1 && pair_slot(k1 + 1, k1 + 2, K) <= r.slot) k1++;
            k1s[r.stack].push_back((int)k1);
        }
    }
    std::vector<std::vector<int>> hmaps(pl.order.size());
    for (size_t ti = 0; ti < pl.order.size(); ti++) {
        const StackJob& j = pl.jobs[pl.order[ti]];
        if (j.twin_src < 0) continue;
        StackBufs& B = c->sb[pl.sb_base + ti];
        const Geom& g = j.s.g;
        const long long K = g.K;
        // image -> A-stack index, then the images in index order
        std::vector<int>& h = hmaps[ti];
        h.assign((size_t)K, -1);
        std::vector<int> imgs;
        for (int k : k1s[j.twin_src])
            if (h[k] < 0) {
                h[k] = (int)imgs.size();
                imgs.push_back(k);
            }
        h.insert(h.end(), imgs.begin(), imgs.end());
        const int nimg = (int)imgs.size();
        if ((rc = ensure(c, B.Bop, (size_t)K * 2 * g.Lp * g.Rp * 4))) return rc;
        if ((rc = ensure(c, B.astk, (size_t)std::max(nimg, 1) * g.NAT * (32 * 2 * (size_t)g.Rp * 4)))) return rc;
        if ((rc = upload(c, B.amap, h))) return rc;
        pl.table[ti].astk = B.astk.p;
        pl.table[ti].Bop = B.Bop.p;
        pl.table[ti].amap = (const int*)B.amap.p;
    }
    if ((rc = upload(c, c->table, pl.table))) return rc;
    pl.twins_built = true;
    for (size_t ti = 0; ti < pl.order.size(); ti++) {
        if (pl.jobs[pl.order[ti]].twin_src < 0) continue;
        const long long K = pl.jobs[pl.order[ti]].s.g.K;
        const int nimg = (int)(hmaps[ti].size() - (size_t)K);
        if ((rc = enqueue_twin_prep(c, pl, ti, (const int*)c->sb[pl.sb_base + ti].amap.p + K, nimg))) return rc;
    }
    // the host maps are released on return: the uploads must be done
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int enqueue_prep(clm_ctx* c, const Plan& pl) {
    int rc;
    unsigned long long* ctr = (unsigned long long*)c->ctr.p;
    for (size_t ti = 0; ti < pl.order.size(); ti++) {
        const StackJob& j = pl.jobs[pl.order[ti]];
        const StackBufs& B = c->sb[pl.sb_base + ti];
        const Geom& g = j.s.g;
        if (j.twin_src >= 0) {
            if (!pl.twin_lazy && (rc = enqueue_twin_prep(c, pl, ti))) return rc;
            continue;
        }
        if ((rc = enqueue_fill(c, B.nrm.p, g.K, 0))) return rc;
        const long long lines = (long long)g.K * g.Lp;
        const unsigned nb = (unsigned)((lines + 255) / 256);
        if (j.p.pf_dtype == CLM_C128)
            hipLaunchKernelGGL(prep_kernel<double>, dim3(nb), dim3(256), 0, c->stream, (const double*)j.io.pf, g,
                               (const double*)B.H.p, (double2*)B.P64.p, (float2*)B.P32.p, B.Bop.p, j.opk,
                               (unsigned*)B.nrm.p, ctr + 0, j.lsh);
        else
            hipLaunchKernelGGL(prep_kernel<float>, dim3(nb), dim3(256), 0, c->stream, (const float*)j.io.pf, g,
                               (const double*)B.H.p, (double2*)B.P64.p, (float2*)B.P32.p, B.Bop.p, j.opk,
                               (unsigned*)B.nrm.p, ctr + 0, j.lsh);
        HIPCHK(c, hipGetLastError());
        if (!j.exact) {
            const size_t osz = j.opk == OPK_F32 ? 4 : 2;
            const long long nch = (long long)g.K * g.NAT * (long long)(32 * 2 * (size_t)g.Rp * osz / 16);
            AstkFn af = pick_astk(j.opk, g.Rp, g.sym != 0);
            if (!af) return fail(c, CLM_E_UNSUP, "no A-stack kernel for Rp=%d", g.Rp);
            hipLaunchKernelGGL(af, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, c->stream, g,
                               (const float2*)B.P32.p, (const float2*)B.ph32.p, (int4*)B.astk.p, nullptr, g.K);
            HIPCHK(c, hipGetLastError());
        }
    }
    return CLM_OK;
}

// The device job: kernel launches only (no copies, no host waits), so that it
// can be captured into a HIP graph.  With `timed`, events split it into
// preprocessing, pair contraction and re-scoring.
enum { PHASE_A = 1, PHASE_B = 2, PHASE_ALL = 3 };


int enqueue_exact(clm_ctx* c, Exact64Args ea, long long n_upper, int S) {
    if (n_upper <= 0) return CLM_OK;
    const int sj = std::max(1, std::min(S, EX_SJ));
    const long long cap = std::min<long long>(n_upper, EX_CAP);
    ea.item0 = 0;
    ea.item_cap = cap;
    ea.sj = sj;
    ea.sv = (double*)c->ex_v.p;
    ea.sk = (long long*)c->ex_k.p;
    hipLaunchKernelGGL(exact64_kernel, dim3((unsigned)std::min<long long>(cap, 256), (unsigned)sj), dim3(256), 0,
                       c->stream, ea);
    HIPCHK(c, hipGetLastError());
    if (sj > 1) {
        hipLaunchKernelGGL(exact64_reduce_kernel, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, c->stream, ea);
        HIPCHK(c, hipGetLastError());
    }
    if (n_upper > EX_CAP) {
        ea.item0 = EX_CAP;
        ea.item_cap = n_upper - EX_CAP;
        ea.sj = 1;
        hipLaunchKernelGGL(exact64_kernel,
                           dim3((unsigned)std::min<long long>(n_upper - EX_CAP, (long long)c->n_cu * 8), 1), dim3(256),
                           0, c->stream, ea);
        HIPCHK(c, hipGetLastError());
    }
    return CLM_OK;
}

int enqueue_job(clm_ctx* c, const Plan& pl, bool timed, int phases = PHASE_ALL) {
    int rc;
    unsigned long long* ctr = (unsigned long long*)c->ctr.p;
    const StackDev* dtab = (const StackDev*)c->table.p;
    if (phases & PHASE_A) {
    // ctr (u64): 0 small-norm lines, 1 FP64 candidates, 4 fallback count (int),
    // 6 cells re-scored, 7 second-pass candidates; per group at 16+4g:
    // 0 pass-1 tile counter (one queue), 1 uncertain count, 2 pass-2 tile counter,
    // 3 pass-2 tiles, 4-7 pass-1 tile counters of the NQ XCD queues (int each)
    if ((rc = enqueue_fill(c, ctr, 2 * (16 + GSTRIDE * (long long)pl.groups.size()), 0))) return rc;
    if (timed) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    if ((rc = enqueue_prep(c, pl))) return rc;
    if (timed) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    // pass 1, one persistent launch per group (twins run in phase B)
    for (size_t gi = 0; gi < pl.groups.size(); gi++) {
        const Group& G = pl.groups[gi];
        if (G.twin_of >= 0) continue;
        if (G.exact) {
            Exact64Args ea{};
            ea.stacks = dtab;
            ea.stack0 = G.first;
            ea.nstacks = (int)G.members.size();
            ea.list = nullptr;
            ea.n_dev = nullptr;
            ea.n_all = G.P;
            if (G.sel) {
                ea.list = pl.jobs[G.members[0]].sel_list;
                ea.n_dev = pl.jobs[G.members[0]].sel_count;
            }
            if (G.filt > 0) {
                ClmapArgs ca{dtab, G.first, (int)G.members.size(), G.P};
                const long long nb = std::min<long long>(G.P, (long long)c->n_cu * 2);
                hipLaunchKernelGGL(clmap_kernel, dim3((unsigned)nb), dim3(256), 0, c->stream, ca);
                HIPCHK(c, hipGetLastError());
                continue;
            }
            int S = 1;
            for (int i : G.members) S = std::max(S, pl.jobs[i].s.g.S);
            if ((rc = enqueue_exact(c, ea, G.P, S))) return rc;
            continue;
        }
        unsigned long long* cg = ctr + 16 + GSTRIDE * gi;
        PairArgs pa{};
        pa.stacks = dtab;
        pa.counter = (int*)(cg + 0);
        pa.nq = 1;
        pa.erecs = nullptr;
        pa.CPP = G.CPP;
        pa.lag = G.lag;
        pa.prio = (pl.jobs[G.members[0]].p.reserved & 4) ? 0 : 1;
        long long max_tiles;
        if (G.sel) {
            const StackJob& j = pl.jobs[G.members[0]];
            pa.tiles = j.sel_tiles;
            pa.ntiles = 0;
            pa.ntiles_dev = j.sel_ntiles;
            pa.plist = j.sel_plist;
            max_tiles = G.P;  // at least one pair per tile
        } else {
            pa.tiles = (const int4*)G.tc->tiles.p;
            pa.ntiles = G.tc->ntiles;
            pa.ntiles_dev = nullptr;
            pa.plist = nullptr;
            max_tiles = G.tc->ntiles;
        }
        const unsigned grid = (unsigned)std::min<long long>((long long)c->n_cu * G.occ, max_tiles);
        if (!G.sel && grid >= (unsigned)NQ) {
            pa.nq = NQ;
            pa.qbeg = (const int*)(pa.tiles + pa.ntiles);
            pa.counter = (int*)(cg + 4);
        }
        hipLaunchKernelGGL(G.fn, dim3(grid), dim3(64 * G.W), G.lds, c->stream, pa);
        HIPCHK(c, hipGetLastError());
    }
    if (timed) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    // re-score; uncertain pairs are listed for the refinement or second pass
    for (size_t gi = 0; gi < pl.groups.size(); gi++) {
        const Group& G = pl.groups[gi];
        if (G.exact || G.twin_of >= 0) continue;
        unsigned long long* cg = ctr + 16 + GSTRIDE * gi;
        SlotRef* ul = (SlotRef*)c->ulist.p + G.ubase;
        RescoreArgs ra{};
        ra.stacks = dtab;
        ra.stack0 = G.first;
        ra.nstacks = (int)G.members.size();
        ra.mode = 0;
        ra.n_items = G.P;
        if (G.sel) {
            const StackJob& j = pl.jobs[G.members[0]];
            ra.mode = 2;
            ra.n_items_dev = j.sel_count;
            ra.item_cap = (int)G.P;
            ra.ulist = j.sel_list;
        }
        ra.unc_count = (int*)(cg + 1);
        ra.unc_list = ul;
        ra.fb_count = (int*)(ctr + 4);
        ra.fb_list = (SlotRef*)c->fblist.p;
        ra.n64 = ctr + 1;
        ra.ncell = ctr + 6;
        ra.nemit = ctr + 7;
        const unsigned rgrid = (unsigned)std::min<long long>((G.P + 3) / 4, (long long)c->n_cu * 16);
        hipLaunchKernelGGL(rescore_kernel, dim3(rgrid), dim3(256), 0, c->stream, ra);
        HIPCHK(c, hipGetLastError());
    }
    }  // phase A
    if (!(phases & PHASE_B)) return CLM_OK;
    for (size_t gi0 = 0; gi0 < pl.groups.size(); gi0++) {
        if (pl.groups[gi0].exact || pl.groups[gi0].twin_of >= 0) continue;
        size_t gi = gi0;  // the group whose uncertain list takes the second pass
        if (pl.groups[gi0].twin >= 0) {
            // refinement: the FP32 twin re-runs the uncertain pairs (pass 1 over
            // one-pair tiles, re-score of its records); its own uncertain pairs
            // then take the FP32 second pass below
            const Group& G0 = pl.groups[gi0];
            const Group& T = pl.groups[G0.twin];
            unsigned long long* c0 = ctr + 16 + GSTRIDE * gi0;
            unsigned long long* cT = ctr + 16 + GSTRIDE * (size_t)G0.twin;
            const SlotRef* ul0 = (const SlotRef*)c->ulist.p + G0.ubase;
            const int shift = T.first - G0.first;
            hipLaunchKernelGGL(build_pass2_kernel, dim3((unsigned)std::min<long long>((G0.P + 255) / 256, 1024)),
                               dim3(256), 0, c->stream, dtab, (const int*)(c0 + 1), ul0, (int)G0.P,
                               (int4*)c->tw_tiles.p, (int4*)c->tw_plist.p, (int*)(cT + 3), (int*)(ctr + 4),
                               (SlotRef*)c->fblist.p, shift);
            HIPCHK(c, hipGetLastError());
            PairArgs pa{};
            pa.stacks = dtab;
            pa.tiles = (const int4*)c->tw_tiles.p;
            pa.ntiles = 0;
            pa.ntiles_dev = (const int*)(cT + 3);
            pa.counter = (int*)(cT + 0);
            pa.plist = (const int4*)c->tw_plist.p;
            pa.erecs = nullptr;
            pa.CPP = T.CPP;
            pa.lag = T.lag;
            const unsigned g1 = (unsigned)std::min<long long>((long long)c->n_cu * T.occ, G0.P);
            hipLaunchKernelGGL(T.fn, dim3(g1), dim3(64 * T.W), T.lds, c->stream, pa);
            HIPCHK(c, hipGetLastError());
            RescoreArgs rt{};
            rt.stacks = dtab;
            rt.mode = 2;
            rt.n_items_dev = (const int*)(c0 + 1);
            rt.item_cap = (int)G0.P;
            rt.ulist = ul0;
            rt.stack_shift = shift;
            rt.unc_count = (int*)(cT + 1);
            rt.unc_list = (SlotRef*)c->ulist.p + T.ubase;
            rt.fb_count = (int*)(ctr + 4);
            rt.fb_list = (SlotRef*)c->fblist.p;
            rt.n64 = ctr + 1;
            rt.ncell = ctr + 6;
            rt.nemit = ctr + 7;
            const unsigned rg = (unsigned)std::min<long long>((G0.P + 3) / 4, (long long)c->n_cu * 16);
            hipLaunchKernelGGL(rescore_kernel, dim3(rg), dim3(256), 0, c->stream, rt);
            HIPCHK(c, hipGetLastError());
            gi = (size_t)G0.twin;
        }
        const Group& G = pl.groups[gi];
        unsigned long long* cg = ctr + 16 + GSTRIDE * gi;
        SlotRef* ul = (SlotRef*)c->ulist.p + G.ubase;
        RescoreArgs ra{};
        ra.stacks = dtab;
        ra.stack0 = G.first;
        ra.nstacks = (int)G.members.size();
        ra.fb_count = (int*)(ctr + 4);
        ra.fb_list = (SlotRef*)c->fblist.p;
        ra.n64 = ctr + 1;
        ra.ncell = ctr + 6;
        ra.nemit = ctr + 7;
        const int cap = (int)std::min<long long>(G.P, pl.n_unc_cap);
        hipLaunchKernelGGL(build_pass2_kernel, dim3((unsigned)std::min<long long>((G.P + 255) / 256, 1024)),
                           dim3(256), 0, c->stream, dtab, (const int*)(cg + 1), (const SlotRef*)ul, cap,
                           (int4*)c->tiles2.p, (int4*)c->plist.p, (int*)(cg + 3), (int*)(ctr + 4),
                           (SlotRef*)c->fblist.p, 0);
        HIPCHK(c, hipGetLastError());
        PairArgs pa{};
        pa.stacks = dtab;
        pa.tiles = (const int4*)c->tiles2.p;
        pa.ntiles = 0;
        pa.ntiles_dev = (const int*)(cg + 3);
        pa.counter = (int*)(cg + 2);
        pa.plist = (const int4*)c->plist.p;
        pa.erecs = (EmitRec*)c->erecs.p;
        pa.CPP = G.CPP;
        pa.lag = G.lag;
        pa.prio = (pl.jobs[G.members[0]].p.reserved & 4) ? 0 : 1;
        const unsigned g2 = (unsigned)std::min<long long>((long long)c->n_cu * G.occ, cap);
        hipLaunchKernelGGL(G.fn2, dim3(g2), dim3(64 * G.W), G.lds, c->stream, pa);
        HIPCHK(c, hipGetLastError());
        ra.mode = 1;
        ra.n_items = 0;
        ra.n_items_dev = (const int*)(cg + 1);
        ra.item_cap = cap;
        ra.ulist = ul;
        ra.erecs = (const EmitRec*)c->erecs.p;
        const unsigned rg2 = (unsigned)std::min<long long>((cap + 3) / 4, (long long)c->n_cu * 16);
        hipLaunchKernelGGL(rescore_kernel, dim3(rg2), dim3(256), 0, c->stream, ra);
        HIPCHK(c, hipGetLastError());
    }
    // pairs no pass could settle: FP64 exhaustive search
    if (pl.P > 0) {
        Exact64Args ea{};
        ea.stacks = dtab;
        ea.list = (const SlotRef*)c->fblist.p;
        ea.n_dev = (const int*)(ctr + 4);
        int S = 1;
        for (const StackJob& j : pl.jobs) S = std::max(S, j.s.g.S);
        if ((rc = enqueue_exact(c, ea, pl.P, S))) return rc;
    }
    if (timed) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    return CLM_OK;
}

// Add a finished plan's counters to st (which the caller cleared).
int add_stats(clm_ctx* c, const Plan& pl, clm_stats* st) {
    if (!st || pl.P == 0) return CLM_OK;
    std::vector<unsigned long long> h(16 + GSTRIDE * pl.groups.size());
    HIPCHK(c, hipMemcpy(h.data(), c->ctr.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    auto as_int = [](unsigned long long x) { return (int64_t)(int)(unsigned)(x & 0xffffffffu); };
    st->small_norm_lines += (int64_t)h[0];
    st->fp64_candidates += (int64_t)h[1];
    st->cells_rescored += (int64_t)h[6];
    st->emitted += (int64_t)h[7];
    st->fallback_pairs += as_int(h[4]);
    for (size_t gi = 0; gi < pl.groups.size(); gi++) {
        const Group& G = pl.groups[gi];
        if (G.exact && !G.sel) st->fallback_pairs += G.P;
        else if (!G.exact && G.twin_of < 0) {
            st->uncertain_pairs += as_int(h[16 + GSTRIDE * gi + 1]);
            if (G.twin >= 0) st->refined_pairs += as_int(h[16 + GSTRIDE * gi + 1]);
        }
    }
    return CLM_OK;
}

void begin_stats(const Plan& pl, clm_stats* st) {
    if (!st) return;
    std::memset(st, 0, sizeof *st);
    for (const StackJob& j : pl.jobs)
        if (j.twin_src < 0) st->pairs += j.P;
    if (!pl.jobs.empty()) {
        st->n_eval_shifts = pl.jobs[0].s.g.NE;
        st->variant = pl.jobs[0].v.ct * 100 + pl.jobs[0].v.W;
    }
}

void timed_stats(clm_ctx* c, clm_stats* st) {
    if (!st) return;
    st->ms_prep = event_ms(c->ev[0], c->ev[1]);
    st->ms_pairs = event_ms(c->ev[1], c->ev[2]);
    st->ms_rescore = event_ms(c->ev[2], c->ev[3]);
    st->ms_total = event_ms(c->ev[0], c->ev[3]);
}

// Run a plan to completion (no graph).  With lazy twins the host waits
// after the re-score: the twins are built only if a group refined by them
// has uncertain pairs (else their refinement launches find nothing to do).
int run_plan(clm_ctx* c, Plan& pl, bool timed, clm_stats* st) {
    int rc;
    if (!pl.twin_lazy) {
        if ((rc = enqueue_job(c, pl, timed))) return rc;
    } else {
        if ((rc = enqueue_job(c, pl, timed, PHASE_A))) return rc;
        bool any = false;
        for (const Group& G : pl.groups) any = any || G.twin >= 0;
        if (any) {
            std::vector<unsigned long long> h(16 + GSTRIDE * pl.groups.size());
            HIPCHK(c, hipMemcpyAsync(h.data(), c->ctr.p, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            bool need = false;
            for (size_t gi = 0; gi < pl.groups.size(); gi++)
                if (pl.groups[gi].twin >= 0 && (int)(unsigned)(h[16 + GSTRIDE * gi + 1] & 0xffffffffu) > 0) need = true;
            if (need && (rc = build_twins(c, pl, h))) return rc;
        }
        if ((rc = enqueue_job(c, pl, timed, PHASE_B))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return add_stats(c, pl, st);
}

void drop_graph(clm_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->gsig.clear();
}

// Every argument the captured job depends on, buffer addresses included.
std::vector<unsigned long long> signature(const clm_ctx* c, const Plan& pl) {
    std::vector<unsigned long long> s;
    auto add = [&](const void* p, size_t n) {
        const size_t w = (n + 7) / 8;
        const size_t o = s.size();
        s.resize(o + w, 0ull);
        std::memcpy(s.data() + o, p, n);
    };
    for (const StackJob& j : pl.jobs) {
        add(&j.p, sizeof j.p);
        add(&j.io, sizeof j.io);
    }
    add(pl.table.data(), pl.table.size() * sizeof(StackDev));
    for (size_t ti = 0; ti < pl.order.size(); ti++) {
        const StackBufs& B = c->sb[pl.sb_base + ti];
        const void* sp[] = {B.P64.p, B.P32.p, B.Bop.p, B.nrm.p, B.astk.p, B.recs.p, B.ph32.p, B.ph64.p, B.shiftv.p, B.H.p,
                            B.W.p};
        add(sp, sizeof sp);
    }
    const void* bufs[] = {c->table.p, c->ctr.p, c->ulist.p, c->erecs.p, c->plist.p, c->tiles2.p, c->fblist.p,
                          c->tw_tiles.p, c->tw_plist.p, c->ex_v.p, c->ex_k.p};
    add(bufs, sizeof bufs);
    for (const Group& G : pl.groups) {
        const void* t = G.tc ? G.tc->tiles.p : nullptr;
        add(&t, sizeof t);
    }
    return s;
}

// Device job of n stacks (results in device memory).  With `graph`, the job
// is captured once and the graph replayed while the signature repeats.
int run_job_core(clm_ctx* c, int n, const clm_params* ps, const clm_stack_io* ios, bool graph, clm_stats* st,
                 const StackJob* extra) {
    Plan pl;
    // lazy twins need a host wait between the phases, and grow memory after
    // kernels have run: not in a captured job, nor in an allocation-free call
    // (which checks every buffer against its scratch array before any kernel)
    pl.twin_lazy = !graph && !c->arena;
    int rc = prepare(c, pl, n, ps, ios, true, extra);
    if (rc) return rc;
    begin_stats(pl, st);
    if (pl.P == 0) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return CLM_OK;
    }
    if (!graph) {
        if ((rc = run_plan(c, pl, true, st))) return rc;
        timed_stats(c, st);
        return CLM_OK;
    }
    std::vector<unsigned long long> sig = signature(c, pl);
    if (!c->gexec || sig != c->gsig) {
        drop_graph(c);
        HIPCHK(c, hipStreamSynchronize(c->stream));  // uploads done before capture
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_job(c, pl, false);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return fail(c, CLM_E_HIP, "graph capture: %s", hipGetErrorString(e));
        c->graph = gr;
        HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->gsig = sig;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    rc = add_stats(c, pl, st);
    if (st) st->ms_total = event_ms(c->ev[0], c->ev[3]);
    return rc;
}

void log_line(clm_ctx* c, int level, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->log_fn(c->log_user, level, buf);
}

// What verbose bit 1 reports (the reference prints per-pair progress,
// cryo_clmatrix_cpu.m:744-749): the call's shape and, once the device job is
// done, its stages and counts.  The job is one device launch sequence with no
// host wait, so the report follows its completion.  Rays of near-zero norm
// are reported whatever the verbose bits, as cryo_raynormalize.m:20-22 warns.
void log_job(clm_ctx* c, int n, const clm_params* ps, const clm_stats* st) {
    if (!c->log_fn || !st) return;
    bool progress = false;
    for (int i = 0; i < n; i++) progress = progress || (ps[i].verbose & 1);
    static const char* pname[] = {"FP32", "BF16 + re-score", "FP16 + re-score", "FP64"};
    if (progress) {
        for (int i = 0; i < n; i++) {
            const clm_params& p = ps[i];
            const int S = (int)std::ceil(2.0 * p.max_shift / p.shift_step + 1.0);
            log_line(c, 0, "%s%sK=%d n_r=%d n_theta=%d max_shift=%g shift_step=%g (%d shifts), %s",
                     n > 1 ? "stack " : "cryo_clmatrix: ", n > 1 ? (std::to_string(i) + ": ").c_str() : "", p.K,
                     p.n_r, p.n_theta, p.max_shift, p.shift_step, S,
                     p.precision >= 0 && p.precision <= 3 ? pname[p.precision] : "?");
        }
        log_line(c, 0, "%lld pairs in %.2f ms (%.4g pairs/s): preprocessing %.2f ms, contraction %.2f ms, "
                       "re-score %.2f ms; kernel variant %d, %d evaluated shifts",
                 (long long)st->pairs, st->ms_total, st->ms_total > 0 ? st->pairs / (st->ms_total * 1e-3) : 0.0,
                 st->ms_prep, st->ms_pairs, st->ms_rescore, st->variant, st->n_eval_shifts);
        log_line(c, 0, "%lld cells re-scored in FP32, %lld candidates in FP64; %lld uncertain pairs (%lld refined "
                       "with FP32 operands), %lld finished by the FP64 exhaustive search",
                 (long long)st->cells_rescored, (long long)st->fp64_candidates, (long long)st->uncertain_pairs,
                 (long long)st->refined_pairs, (long long)st->fallback_pairs);
    }
    if (st->small_norm_lines)
        log_line(c, 1, "Ray norm is close to zero for %lld lines (norm < 1e-13)", (long long)st->small_norm_lines);
}

int run_job(clm_ctx* c, int n, const clm_params* ps, const clm_stack_io* ios, bool graph, clm_stats* st,
            const StackJob* extra = nullptr) {
    clm_stats local;
    if (!st && c->log_fn) st = &local;
    const int rc = run_job_core(c, n, ps, ios, graph, st, extra);
    if (rc == CLM_OK) log_job(c, n, ps, st);
    return rc;
}

// Learned form: exhaustive search cut to a window around each pair's shift
// as predicted from per-image 2-D shift proposals, with the exhaustive search
// as the fallback for pairs whose window peak is below `gate` or lands off
// the reference grid.
int run_windowed(clm_ctx* c, const clm_params* p, const void* d_pf, const float* d_prop, double window,
                 double gate, const clm_stack_io& out, uint8_t* d_flag, clm_window_stats* ws) {
    int rc;
    if (ws) std::memset(ws, 0, sizeof *ws);
    Setup sf;
    if ((rc = setup(c, p, sf))) return rc;
    const double step = p->shift_step;
    const double q = p->max_shift / step, wq = window / step;
    if (std::fabs(q - std::rint(q)) > 1e-9)
        return fail(c, CLM_E_UNSUP, "windowed search needs max_shift (%g) to be a multiple of shift_step (%g)",
                    p->max_shift, step);
    if (!(window >= 0) || std::fabs(wq - std::rint(wq)) > 1e-9)
        return fail(c, CLM_E_ARG, "window (%g) must be a non-negative multiple of shift_step (%g)", window, step);
    if (p->precision == CLM_FP64_REF) return fail(c, CLM_E_UNSUP, "windowed search runs the MFMA precisions");
    const long long K = p->K, P = K * (K - 1) / 2;
    const int L = sf.g.L;
    if (ws) {
        ws->pairs = P;
        ws->full_shifts = sf.g.S;
        ws->window_shifts = 2 * (int)std::rint(wq) + 1;
    }
    if (P == 0) return CLM_OK;
    if (!d_pf || !d_prop) return fail(c, CLM_E_ARG, "null pf or proposals");
    // full-grid variant first: its pairs per workgroup shape the fallback tiles
    StackJob probe;
    if ((rc = plan_stack(c, p, probe))) return rc;
    const int ppw = probe.v.W / probe.v.CPP;
    if ((rc = ensure(c, c->w_cl1, P * 4)) || (rc = ensure(c, c->w_cl2, P * 4)) || (rc = ensure(c, c->w_j, P * 4)) ||
        (rc = ensure(c, c->w_corr, P * 8)) || (rc = ensure(c, c->w_shift, P * 8)) ||
        (rc = ensure(c, c->w_lsh, (size_t)K * L * 8)) || (rc = ensure(c, c->w_flag, (size_t)P)) ||
        (rc = ensure(c, c->w_rows, (size_t)(4 * K + 4) * 4)) ||
        (rc = ensure(c, c->w_list, (size_t)P * sizeof(SlotRef))) ||
        (rc = ensure(c, c->w_plist, (size_t)P * sizeof(int4))) ||
        (rc = ensure(c, c->w_tiles, (size_t)P * sizeof(int4))) ||
        (rc = ensure(c, c->w_counts, 8 * sizeof(unsigned long long))))
        return rc;
    if ((rc = upload(c, c->w_shiftv, sf.shifts))) return rc;
    int* rows = (int*)c->w_rows.p;
    int *cnt = rows, *tcnt = rows + K, *off = rows + 2 * K, *toff = rows + 3 * K, *tot = rows + 4 * K;
    unsigned long long* counts = (unsigned long long*)c->w_counts.p;

    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    if ((rc = enqueue_fill(c, counts, 16, 0)) || (rc = enqueue_fill(c, tot, 2, 0))) return rc;
    hipLaunchKernelGGL(lineshift_kernel, dim3((unsigned)((K * L + 255) / 256)), dim3(256), 0, c->stream, d_prop,
                       (int)K, L, step, (double*)c->w_lsh.p);
    HIPCHK(c, hipGetLastError());
    // windowed search: the same engine on rays moved per line, 2*window/step+1 shifts
    clm_params pw = *p;
    pw.max_shift = window;
    StackJob xw;
    xw.lsh = (const double*)c->w_lsh.p;
    const clm_stack_io iow{d_pf, (int32_t*)c->w_cl1.p, (int32_t*)c->w_cl2.p, (int32_t*)c->w_j.p,
                           (double*)c->w_corr.p, (double*)c->w_shift.p};
    {
        Plan plw;
        plw.twin_lazy = !c->arena;
        if ((rc = prepare(c, plw, 1, &pw, &iow, true, &xw, 0))) return rc;
        if ((rc = run_plan(c, plw, false, nullptr))) return rc;
        hipLaunchKernelGGL(window_merge_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, c->stream, (int)K,
                           L, (const double*)c->w_lsh.p, (const int32_t*)c->w_cl1.p, (const int32_t*)c->w_cl2.p,
                           (const double*)c->w_corr.p, (const double*)c->w_shift.p, gate, p->max_shift, step,
                           sf.g.S, (const double*)c->w_shiftv.p, out.cl1, out.cl2, out.jidx, out.corr, out.shift1d,
                           (uint8_t*)c->w_flag.p, counts);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(row_count_kernel, dim3((unsigned)(K - 1)), dim3(256), 0, c->stream, (int)K,
                           (const uint8_t*)c->w_flag.p, ppw, cnt, tcnt);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(row_scan_kernel, dim3(1), dim3(1024), 0, c->stream, (int)(K - 1), (const int*)cnt,
                           (const int*)tcnt, off, toff, tot);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(row_emit_kernel, dim3((unsigned)(K - 1)), dim3(256), 0, c->stream, (int)K, 0,
                           (const uint8_t*)c->w_flag.p, ppw, (const int*)off, (const int*)toff,
                           (SlotRef*)c->w_list.p, (int4*)c->w_plist.p, (int4*)c->w_tiles.p);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));  // the window plan's shared buffers are reused below
    }
    // exhaustive search of the flagged pairs, results into the same outputs
    StackJob xf;
    xf.sel_list = (const SlotRef*)c->w_list.p;
    xf.sel_count = tot;
    xf.sel_tiles = (const int4*)c->w_tiles.p;
    xf.sel_plist = (const int4*)c->w_plist.p;
    xf.sel_ntiles = tot + 1;
    {
        Plan plf;
        plf.twin_lazy = !c->arena;
        if ((rc = prepare(c, plf, 1, p, &out, true, &xf, 2))) return rc;
        if (plf.jobs[0].v.W / plf.jobs[0].v.CPP != ppw) return fail(c, CLM_E_HIP, "fallback tile shape changed");
        if ((rc = run_plan(c, plf, false, nullptr))) return rc;
    }
    if (d_flag) HIPCHK(c, hipMemcpyAsync(d_flag, c->w_flag.p, (size_t)P, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (ws) {
        unsigned long long h[4] = {0};
        HIPCHK(c, hipMemcpy(h, counts, sizeof h, hipMemcpyDeviceToHost));
        ws->accepted = (int64_t)h[0];
        ws->below_gate = (int64_t)h[1];
        ws->off_grid = (int64_t)h[2];
        ws->ms_window = event_ms(c->ev[0], c->ev[1]);
        ws->ms_fallback = event_ms(c->ev[1], c->ev[2]);
        ws->ms_total = event_ms(c->ev[0], c->ev[2]);
    }
    return CLM_OK;
}

// Partner-subset search: the pairs flagged in d_sel (one byte per pair, in
// equation order), each with the full search of the other entry points; the
// rest keep the outputs of a pair with no common line.  The flags become
// explicit pass-1 tiles on the device, as the windowed search's fallback
// pairs do.
int run_subset(clm_ctx* c, const clm_params* p, const uint8_t* d_sel, const clm_stack_io& out, clm_stats* st) {
    int rc;
    if (st) std::memset(st, 0, sizeof *st);
    StackJob probe;
    if ((rc = plan_stack(c, p, probe))) return rc;
    const long long K = p->K, P = probe.P;
    if (P == 0) return CLM_OK;
    if (!out.pf || !d_sel || !out.cl1 || !out.cl2 || !out.jidx || !out.corr || !out.shift1d)
        return fail(c, CLM_E_ARG, "null pf, selection or output");
    const int ppw = probe.v.W / probe.v.CPP;
    if ((rc = ensure(c, c->w_rows, (size_t)(4 * K + 4) * 4)) ||
        (rc = ensure(c, c->w_list, (size_t)P * sizeof(SlotRef))) ||
        (rc = ensure(c, c->w_plist, (size_t)P * sizeof(int4))) ||
        (rc = ensure(c, c->w_tiles, (size_t)P * sizeof(int4))))
        return rc;
    int* rows = (int*)c->w_rows.p;
    int *cnt = rows, *tcnt = rows + K, *off = rows + 2 * K, *toff = rows + 3 * K, *tot = rows + 4 * K;
    const unsigned ig = (unsigned)std::min<long long>((P + 255) / 256, (long long)c->n_cu * 16);
    hipLaunchKernelGGL(subset_init_kernel, dim3(ig), dim3(256), 0, c->stream, P, d_sel, out.cl1, out.cl2, out.jidx,
                       out.corr, out.shift1d);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(row_count_kernel, dim3((unsigned)(K - 1)), dim3(256), 0, c->stream, (int)K, d_sel, ppw, cnt,
                       tcnt);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(row_scan_kernel, dim3(1), dim3(1024), 0, c->stream, (int)(K - 1), (const int*)cnt,
                       (const int*)tcnt, off, toff, tot);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(row_emit_kernel, dim3((unsigned)(K - 1)), dim3(256), 0, c->stream, (int)K, 0, d_sel, ppw,
                       (const int*)off, (const int*)toff, (SlotRef*)c->w_list.p, (int4*)c->w_plist.p,
                       (int4*)c->w_tiles.p);
    HIPCHK(c, hipGetLastError());
    int nsel = 0;
    HIPCHK(c, hipMemcpyAsync(&nsel, tot, sizeof nsel, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    StackJob xs;
    xs.sel_list = (const SlotRef*)c->w_list.p;
    xs.sel_count = tot;
    xs.sel_tiles = (const int4*)c->w_tiles.p;
    xs.sel_plist = (const int4*)c->w_plist.p;
    xs.sel_ntiles = tot + 1;
    Plan pl;
    pl.twin_lazy = !c->arena;
    if ((rc = prepare(c, pl, 1, p, &out, true, &xs, 0))) return rc;
    if (!pl.jobs[0].exact && pl.jobs[0].v.W / pl.jobs[0].v.CPP != ppw)
        return fail(c, CLM_E_HIP, "subset tile shape changed");
    begin_stats(pl, st);
    if (st) st->pairs = nsel;
    if (nsel == 0) return CLM_OK;
    if ((rc = run_plan(c, pl, true, st))) return rc;
    timed_stats(c, st);
    return CLM_OK;
}

// Candidate variants of a shape: every (CT, W) the kernels support, W a
// multiple of the waves per pair between 4 and 8.
std::vector<std::pair<int, int>> candidate_variants(int opk, int NCT, int Rp) {
    std::vector<std::pair<int, int>> out;
    for (int ct = 1; ct <= 3; ct++) {
        if (opk == OPK_F32 && (ct == 3 || (ct == 2 && Rp > 32))) continue;
        if (opk != OPK_F32 && ct == 3 && Rp > 64) continue;
        const int CPP = (NCT + ct - 1) / ct;
        if (CPP > 8) continue;
        for (int W = CPP; W <= 8; W += CPP)
            if (W >= 4) out.push_back({ct, W});
    }
    return out;
}

int autotune(clm_ctx* c, const clm_params* p, const void* d_pf, int sample_pairs, clm_tune_entry* out,
             int max_out, int* n_out) {
    int rc;
    if (n_out) *n_out = 0;
    StackJob base;
    if ((rc = plan_stack(c, p, base))) return rc;
    if (base.exact || base.P == 0) return CLM_OK;  // nothing to choose
    if (!d_pf) return fail(c, CLM_E_ARG, "null pf");
    const Geom& g0 = base.s.g;
    const long long Np = std::min<long long>(base.P, sample_pairs > 0 ? sample_pairs : 200000);
    auto cands = candidate_variants(base.opk, (g0.L + 31) / 32, g0.Rp);
    double best = 1e30;
    int4 best_v = make_int4(0, 0, 0, 0);
    int n = 0;
    for (const auto& cw : cands) {
        for (int lag = (cw.second % 2 == 0) ? 1 : 0; lag >= 0; lag--) {
            clm_params q = *p;
            q.variant = cw.first * 100 + cw.second;
            q.reserved = (p->reserved & ~1) | (lag ? 0 : 1);
            const clm_stack_io io{d_pf, nullptr, nullptr, nullptr, nullptr, nullptr};
            Plan pl;
            if ((rc = prepare(c, pl, 1, &q, &io, false, nullptr, 0, false))) return rc;
            const Group& G = pl.groups[0];
            const long long K = g0.K;
            const int PPW = G.W / G.CPP;
            std::vector<int4> h;
            long long taken = 0;
            for (long long k1 = 0; k1 + 1 < K && taken < Np; k1++)
                for (long long k2 = k1 + 1; k2 < K && taken < Np; k2 += PPW) {
                    const int cnt = (int)std::min<long long>({(long long)PPW, K - k2, Np - taken});
                    h.push_back(make_int4((int)k1, (int)k2, cnt, (int)pair_slot(k1, k2, K)));
                    taken += cnt;
                }
            const int ntiles = (int)h.size();
            append_queue_bounds(h);
            if ((rc = upload(c, c->tune_tiles, h))) return rc;
            HIPCHK(c, hipStreamSynchronize(c->stream));  // h is released on return
            if ((rc = enqueue_fill(c, c->ctr.p, 2 * (16 + GSTRIDE), 0))) return rc;
            if ((rc = enqueue_prep(c, pl))) return rc;
            PairArgs pa{};
            pa.stacks = (const StackDev*)c->table.p;
            pa.tiles = (const int4*)c->tune_tiles.p;
            pa.ntiles = ntiles;
            pa.counter = (int*)((unsigned long long*)c->ctr.p + 16);
            pa.nq = 1;
            pa.CPP = G.CPP;
            pa.lag = G.lag;
            pa.prio = (p->reserved & 4) ? 0 : 1;
            const unsigned grid = (unsigned)std::min<long long>((long long)c->n_cu * G.occ, (long long)ntiles);
            if (grid >= (unsigned)NQ) {  // the same XCD queues as the measured job
                pa.nq = NQ;
                pa.qbeg = (const int*)(pa.tiles + ntiles);
                pa.counter = (int*)((unsigned long long*)c->ctr.p + 16 + 4);
            }
            float ms = 0.f;
            for (int rep = 0; rep < 2; rep++) {  // first launch warms caches and code
                if ((rc = enqueue_fill(c, pa.counter, pa.nq, 0))) return rc;
                HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
                hipLaunchKernelGGL(G.fn, dim3(grid), dim3(64 * G.W), G.lds, c->stream, pa);
                HIPCHK(c, hipGetLastError());
                HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                ms = event_ms(c->ev[4], c->ev[5]);
            }
            const double per = (double)ms / (double)taken;
            if (out && n < max_out) out[n] = clm_tune_entry{q.variant, G.lag, (double)ms, (int64_t)taken};
            n++;
            if (per < best) {
                best = per;
                best_v = make_int4(cw.first, cw.second, G.lag, 0);
            }
        }
    }
    if (n_out) *n_out = n;
    if (best_v.x) {
        const unsigned long long key = tune_key(base.opk, base.s.g);
        bool found = false;
        for (auto& e : c->tuned)
            if (e.first == key) { e.second = best_v; found = true; }
        if (!found) c->tuned.push_back({key, best_v});
    }
    return CLM_OK;
}

}  // namespace

// Device memory of the allocation-free entry point
// (clm_clmatrix_device_noalloc): one static array, of which every GPU the
// library runs on holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel or a copy: it uses the per-device
// address from hipGetSymbolAddress, taken once per device at clm_create.
constexpr size_t NOALLOC_SCRATCH_BYTES = CLM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};
hipError_t g_scratch_err[MAX_DEVICES] = {};

// This device's address of `scratch`; the device must be current.
void* scratch_address(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device] && g_scratch_err[device] == hipSuccess) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e == hipSuccess && ptr) g_scratch_ptr[device] = ptr;
        else g_scratch_err[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
    }
    return g_scratch_ptr[device];
}

// For the duration of an allocation-free call, the context's per-call device
// buffers are swapped for empty ones that ensure() carves from the scratch
// array; the context's own (hipMalloc'd) buffers come back unchanged after.
struct NoAllocScope {
    clm_ctx* c;
    std::vector<StackBufs> sb;
    std::vector<TileCache> tcache;
    DevBuf saved[11];
    DevBuf* shared[11];
    explicit NoAllocScope(clm_ctx* ctx)
        : c(ctx), shared{&ctx->table, &ctx->ctr, &ctx->ulist, &ctx->erecs, &ctx->plist, &ctx->tiles2,
                         &ctx->fblist, &ctx->tw_tiles, &ctx->tw_plist, &ctx->ex_v, &ctx->ex_k} {
        sb.swap(c->sb);
        tcache.swap(c->tcache);
        for (int i = 0; i < 11; i++) {
            saved[i] = *shared[i];
            *shared[i] = DevBuf{};
        }
        c->arena = (char*)c->scratch;
        c->arena_cap = c->scratch_cap;
        c->arena_off = 0;
    }
    ~NoAllocScope() {
        c->arena = nullptr;
        c->arena_cap = c->arena_off = 0;
        for (int i = 0; i < 11; i++) *shared[i] = saved[i];
        sb.swap(c->sb);
        tcache.swap(c->tcache);
    }
};
}  // namespace

namespace {
// f(begin, end) over [0, n) split into contiguous ranges, one per host
// thread (at most 32, at least `grain` items each); runs inline when a
// thread cannot be started.  Nothing here throws across the C ABI.
template <typename F>
void host_parallel(long long n, long long grain, F f) {
    if (n <= 0) return;
    const long long hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::max<long long>(1, std::min<long long>({hw, 32LL, n / std::max(grain, 1LL)}));
    std::vector<std::thread> th;
    int started = 0;
    if (nt > 1) {
        try {
            for (int t = 1; t < nt; t++) {
                th.emplace_back([&f, n, nt, t] { f(n * t / nt, n * (t + 1) / nt); });
                started = t;
            }
        } catch (...) {
        }
    }
    // this thread: range 0, then every range whose thread did not start
    f(0, n / nt);
    for (int t = started + 1; t < nt; t++) f(n * t / nt, n * (t + 1) / nt);
    for (auto& x : th) x.join();
}

// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014): the per-row generator of
// clm_partner_subset, so that each row's draw depends on (seed, row) only.
inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Equation number of every pair among the selected ones (0-based; the
// unselected pairs' entries are not used).  Null without a selection.
std::vector<int32_t> selected_rank(long long P, const uint8_t* sel) {
    std::vector<int32_t> r;
    if (!sel) return r;
    r.resize((size_t)P);
    int32_t n = 0;
    for (long long e = 0; e < P; e++) {
        r[(size_t)e] = n;
        n += sel[e] != 0;
    }
    return r;
}

int assemble_impl(int32_t K, const uint8_t* sel, const int32_t* cl1, const int32_t* cl2, const double* corr,
                  double* clstack, double* corrstack, double* eqmap) {
    if (K < 0) return CLM_E_ARG;
    const std::vector<int32_t> rank = selected_rank(clm_num_pairs(K), eqmap ? sel : nullptr);
    // one column of the K x K column-major outputs per step, columns split
    // over host threads: entry (r, c) is pair (r, c) above the diagonal
    // (cl1, 1 - corr, its equation number), pair (c, r) below it (cl2, 0, 0)
    host_parallel(K, 65536 / std::max(K, 1) + 1, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; c++) {
            const size_t col = (size_t)c * K;
            for (long long r = 0; r < K; r++) {
                double v1 = 0.0, v2 = 0.0, v3 = 0.0;
                if (r < c) {
                    const long long e = pair_slot(r, c, K);
                    v1 = cl1 ? cl1[e] : 0;
                    v2 = corr && corr[e] != 0.0 ? 1.0 - corr[e] : 0.0;  // :764
                    if (!sel) v3 = (double)(e + 1);
                    else if (sel[e]) v3 = (double)rank[(size_t)e] + 1.0;
                } else if (r > c) {
                    v1 = cl2 ? cl2[pair_slot(c, r, K)] : 0;
                }
                if (clstack) clstack[col + r] = v1;
                if (corrstack) corrstack[col + r] = v2;
                if (eqmap) eqmap[col + r] = v3;
            }
        }
    });
    return CLM_OK;
}

int shift_equations_impl(int32_t K, int32_t n_theta, const uint8_t* sel, int32_t flags, const int32_t* cl1,
                         const int32_t* cl2, const double* shift1d, int64_t* jc, int64_t* ir, double* pr,
                         int64_t* nnz);
}  // namespace

extern "C" {

int clm_version(void) { return CLM_VERSION; }

void clm_default_params(clm_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->pf_dtype = CLM_C128;
    p->max_shift = 15.0;   // cryo_clmatrix_cpu.m:133
    p->shift_step = 1.0;   // :138
    // FP16: the same MFMA rate as BF16 with an 8x tighter rounding bound,
    // so broad high-SNR peaks stay certain without the FP32 refinement
    p->precision = CLM_FP16_RESCORE;
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
    // the allocation-free entry point's scratch array on this device (an
    // unavailable array is reported by that entry point, not here)
    c->scratch = scratch_address(device);
    c->scratch_cap = c->scratch ? NOALLOC_SCRATCH_BYTES : 0;
    if (err) *err = CLM_OK;
    return c;
}

void clm_destroy(clm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (StackBufs& B : c->sb) {
        DevBuf* sbufs[] = {&B.P64, &B.P32, &B.Bop, &B.nrm, &B.astk, &B.recs, &B.ph32, &B.ph64, &B.shiftv, &B.H, &B.W,
                           &B.amap};
        for (DevBuf* b : sbufs) release(*b);
    }
    for (TileCache& tc : c->tcache) release(tc.tiles);
    release(c->tune_tiles);
    release(c->tw_tiles);
    release(c->ex_v);
    release(c->ex_k);
    release(c->tw_plist);
    DevBuf* snb[] = {&c->sn_h1, &c->sn_h2, &c->sn_z, &c->sn_y, &c->sn_dy, &c->sn_dz, &c->sn_dz2, &c->sn_dz1};
    for (DevBuf* b : snb) release(*b);
    DevBuf* wb[] = {&c->w_cl1, &c->w_cl2, &c->w_j, &c->w_corr, &c->w_shift, &c->w_lsh, &c->w_flag, &c->w_rows,
                    &c->w_list, &c->w_plist, &c->w_tiles, &c->w_counts, &c->w_shiftv};
    for (DevBuf* b : wb) release(*b);
    pft_release(c);
    DevBuf* pb[] = {&c->pft_ax, &c->pft_ay, &c->pft_xt, &c->pft_img, &c->pft_out,
                    &c->pft_grid, &c->pft_base, &c->pft_wt, &c->pft_dec,
                    &c->ls_pack, &c->ls_u, &c->ls_part, &c->ls_vec, &c->ls_tab, &c->ls_state, &c->ls_tiles,
                    &c->es_rot, &c->es_x,
                    &c->vt_m, &c->vt_mt, &c->vt_pk, &c->vt_cs,
                    &c->s3_rs, &c->s3_vs, &c->s3_ss, &c->s3_tiles, &c->s3_thr, &c->s3_hist};
    for (DevBuf* b : pb) release(*b);
    DevBuf* bufs[] = {&c->table, &c->ctr, &c->ulist, &c->erecs, &c->plist, &c->tiles2, &c->fblist, &c->pf,
                      &c->out_cl1, &c->out_cl2, &c->out_j, &c->out_corr, &c->out_shift};
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
    const clm_stack_io io{d_pf, d_cl1, d_cl2, d_jidx, d_corr, d_shift1d};
    return run_job(c, 1, p, &io, false, st);
}

int clm_clmatrix_device_noalloc(clm_ctx* c, const clm_params* p, const void* d_pf, int32_t* d_cl1, int32_t* d_cl2,
                                int32_t* d_jidx, double* d_corr, double* d_shift1d, clm_stats* st) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    if (!d_pf && p->K > 1) return fail(c, CLM_E_ARG, "null pf");
    if (!c->scratch)
        return fail(c, CLM_E_HIP, "the scratch array of device %d is unavailable (hipGetSymbolAddress: %s)",
                    c->device,
                    hipGetErrorString(c->device >= 0 && c->device < MAX_DEVICES ? g_scratch_err[c->device]
                                                                                : hipErrorInvalidDevice));
    HIPCHK(c, hipSetDevice(c->device));
    const clm_stack_io io{d_pf, d_cl1, d_cl2, d_jidx, d_corr, d_shift1d};
    NoAllocScope scope(c);
    return run_job(c, 1, p, &io, false, st);
}

int clm_clmatrix_windowed(clm_ctx* c, const clm_params* p, const void* d_pf, const float* d_prop, double window,
                          double gate, int32_t* d_cl1, int32_t* d_cl2, int32_t* d_jidx, double* d_corr,
                          double* d_shift1d, uint8_t* d_flag, clm_window_stats* ws) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    HIPCHK(c, hipSetDevice(c->device));
    const clm_stack_io out{d_pf, d_cl1, d_cl2, d_jidx, d_corr, d_shift1d};
    return run_windowed(c, p, d_pf, d_prop, window, gate, out, d_flag, ws);
}

int clm_clmatrix_subset_device(clm_ctx* c, const clm_params* p, const void* d_pf, const uint8_t* d_sel,
                               int32_t* d_cl1, int32_t* d_cl2, int32_t* d_jidx, double* d_corr, double* d_shift1d,
                               clm_stats* st) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    HIPCHK(c, hipSetDevice(c->device));
    const clm_stack_io out{d_pf, d_cl1, d_cl2, d_jidx, d_corr, d_shift1d};
    clm_stats local;
    if (!st && c->log_fn) st = &local;
    const int rc = run_subset(c, p, d_sel, out, st);
    if (rc == CLM_OK && st && st->pairs > 0) log_job(c, 1, p, st);
    return rc;
}

int clm_common_lines_from_rotations(clm_ctx* c, const clm_params* p, const void* d_pf, const double* d_rot,
                                    const uint8_t* d_sel, int32_t* d_cl1, int32_t* d_cl2, double* d_shift1d) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    const long long P = clm_num_pairs(p->K);
    if (P > 0 && (!d_pf || !d_rot || !d_cl1 || !d_cl2 || !d_shift1d)) return fail(c, CLM_E_ARG, "null pointer");
    HIPCHK(c, hipSetDevice(c->device));
    // the rays: the FP64 preprocessing of the search (lines, filter, band,
    // normalization; cryo_estimate_shifts.m:63,121,180-188 is that of
    // cryo_clmatrix), without its A-stacks or records
    clm_params q = *p;
    q.precision = CLM_FP64_REF;
    q.variant = 0;
    const clm_stack_io io{d_pf, d_cl1, d_cl2, nullptr, nullptr, d_shift1d};
    Plan pl;
    int rc;
    if ((rc = prepare(c, pl, 1, &q, &io, false, nullptr, 0, false))) return rc;
    if (P == 0) return CLM_OK;
    if ((rc = enqueue_fill(c, c->ctr.p, 32, 0))) return rc;
    if ((rc = enqueue_prep(c, pl))) return rc;
    // a wave per 64 consecutive pairs, eight waves per workgroup; the shift
    // phases in LDS when they fit in 64 KB (41 shifts x 64 ray points: 42 KB)
    const Geom& g = pl.jobs[0].s.g;
    const size_t ph_bytes = (size_t)g.S * g.R * sizeof(double2);
    const unsigned grid = (unsigned)std::min<long long>((P + 511) / 512, (long long)c->n_cu * 8);
    if (ph_bytes <= 65536)
        hipLaunchKernelGGL(rotlines_kernel<true>, dim3(grid), dim3(512), ph_bytes, c->stream,
                           (const StackDev*)c->table.p, d_rot, d_sel, (int)p->n_theta, d_cl1, d_cl2, d_shift1d);
    else
        hipLaunchKernelGGL(rotlines_kernel<false>, dim3(grid), dim3(512), 0, c->stream, (const StackDev*)c->table.p,
                           d_rot, d_sel, (int)p->n_theta, d_cl1, d_cl2, d_shift1d);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_cryo_estimate_shifts(clm_ctx* c, const clm_params* p, const void* pf, const double* rot, const uint8_t* sel,
                             double tol, int32_t maxit, int32_t* cl1, int32_t* cl2, double* shift1d, double* x,
                             clm_lsqr_stats* st) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    if (p->n_theta <= 0 || p->n_theta % 2) return fail(c, CLM_E_ARG, "n_theta must be even");
    if (p->K < 2 || p->n_r < 2) return fail(c, CLM_E_ARG, "at least two images of two radial points are needed");
    if (!pf || !rot || !cl1 || !cl2 || !shift1d || !x) return fail(c, CLM_E_ARG, "null pf, rotations or output");
    HIPCHK(c, hipSetDevice(c->device));
    const long long P = clm_num_pairs(p->K);
    const size_t esz = p->pf_dtype == CLM_C64 ? 8 : 16;
    const size_t pf_bytes = (size_t)p->n_r * p->n_theta * p->K * esz;
    int rc;
    if ((rc = ensure(c, c->pf, pf_bytes)) || (rc = ensure(c, c->es_rot, (size_t)p->K * 9 * sizeof(double))) ||
        (rc = ensure(c, c->es_x, (size_t)p->K * 2 * sizeof(double))) || (rc = ensure(c, c->out_cl1, P * 4)) ||
        (rc = ensure(c, c->out_cl2, P * 4)) || (rc = ensure(c, c->out_shift, P * 8)) ||
        (sel && (rc = ensure(c, c->w_flag, (size_t)P))))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->pf.p, pf, pf_bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->es_rot.p, rot, (size_t)p->K * 9 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (sel) HIPCHK(c, hipMemcpyAsync(c->w_flag.p, sel, (size_t)P, hipMemcpyHostToDevice, c->stream));
    if ((rc = clm_common_lines_from_rotations(c, p, c->pf.p, (const double*)c->es_rot.p,
                                              sel ? (const uint8_t*)c->w_flag.p : nullptr, (int32_t*)c->out_cl1.p,
                                              (int32_t*)c->out_cl2.p, (double*)c->out_shift.p)))
        return rc;
    if ((rc = clm_estimate_shifts_ex(c, p->K, p->n_theta, (const int32_t*)c->out_cl1.p,
                                     (const int32_t*)c->out_cl2.p, (const double*)c->out_shift.p, tol, maxit, 0.0,
                                     CLM_FLIP_BY_INDEX, (double*)c->es_x.p, st)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(cl1, c->out_cl1.p, P * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cl2, c->out_cl2.p, P * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(shift1d, c->out_shift.p, P * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(x, c->es_x.p, (size_t)p->K * 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CLM_OK;
}

int clm_autotune(clm_ctx* c, const clm_params* p, const void* d_pf, int32_t sample_pairs, clm_tune_entry* entries,
                 int32_t max_entries, int32_t* n_entries) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    if (max_entries < 0 || (max_entries > 0 && !entries)) return fail(c, CLM_E_ARG, "invalid entry buffer");
    HIPCHK(c, hipSetDevice(c->device));
    return autotune(c, p, d_pf, sample_pairs, entries, max_entries, n_entries);
}

int clm_clmatrix_grouped(clm_ctx* c, int32_t n, const clm_params* p, const clm_stack_io* io, int32_t flags,
                         clm_stats* st) {
    if (!c) return CLM_E_ARG;
    if (n < 0 || (n > 0 && (!p || !io))) return fail(c, CLM_E_ARG, "invalid stack list");
    if (flags & ~CLM_GROUP_GRAPH) return fail(c, CLM_E_ARG, "unknown flags 0x%x", flags);
    HIPCHK(c, hipSetDevice(c->device));
    return run_job(c, n, p, io, (flags & CLM_GROUP_GRAPH) != 0, st);
}

// Host-buffer form of one stack: copy pf in, run, copy the outputs back.
static int host_job(clm_ctx* c, const clm_params* p, const void* pf, int32_t* cl1, int32_t* cl2, int32_t* jidx,
                    double* corr, double* shift1d, clm_stats* st, const StackJob* extra) {
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
    const clm_stack_io io{c->pf.p, (int32_t*)c->out_cl1.p, (int32_t*)c->out_cl2.p, (int32_t*)c->out_j.p,
                          (double*)c->out_corr.p, (double*)c->out_shift.p};
    rc = run_job(c, 1, p, &io, false, st, extra);
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

int clm_clmatrix(clm_ctx* c, const clm_params* p, const void* pf, int32_t* cl1, int32_t* cl2, int32_t* jidx,
                 double* corr, double* shift1d, clm_stats* st) {
    return host_job(c, p, pf, cl1, cl2, jidx, corr, shift1d, st, nullptr);
}

int clm_clmatrix_subset(clm_ctx* c, const clm_params* p, const void* pf, const uint8_t* sel, int32_t* cl1,
                        int32_t* cl2, int32_t* jidx, double* corr, double* shift1d, clm_stats* st) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    if (p->n_theta % 2) return fail(c, CLM_E_ARG, "n_theta must be even");
    if (p->K < 0 || p->n_r < 0 || p->n_theta < 0) return fail(c, CLM_E_ARG, "negative size");
    const long long P = clm_num_pairs(p->K);
    if (P > 0 && (!pf || !sel || !cl1 || !cl2 || !jidx || !corr || !shift1d))
        return fail(c, CLM_E_ARG, "null pf, selection or output");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t esz = p->pf_dtype == CLM_C64 ? 8 : 16;
    const size_t pf_bytes = (size_t)p->n_r * p->n_theta * p->K * esz;
    int rc;
    if ((rc = ensure(c, c->pf, pf_bytes)) || (rc = ensure(c, c->w_flag, (size_t)P)) ||
        (rc = ensure(c, c->out_cl1, P * 4)) || (rc = ensure(c, c->out_cl2, P * 4)) ||
        (rc = ensure(c, c->out_j, P * 4)) || (rc = ensure(c, c->out_corr, P * 8)) ||
        (rc = ensure(c, c->out_shift, P * 8)))
        return rc;
    if (pf_bytes) HIPCHK(c, hipMemcpyAsync(c->pf.p, pf, pf_bytes, hipMemcpyHostToDevice, c->stream));
    if (P > 0) HIPCHK(c, hipMemcpyAsync(c->w_flag.p, sel, (size_t)P, hipMemcpyHostToDevice, c->stream));
    rc = clm_clmatrix_subset_device(c, p, c->pf.p, (const uint8_t*)c->w_flag.p, (int32_t*)c->out_cl1.p,
                                    (int32_t*)c->out_cl2.p, (int32_t*)c->out_j.p, (double*)c->out_corr.p,
                                    (double*)c->out_shift.p, st);
    if (rc) return rc;
    if (P > 0) {
        HIPCHK(c, hipMemcpyAsync(cl1, c->out_cl1.p, P * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(cl2, c->out_cl2.p, P * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(jidx, c->out_j.p, P * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(corr, c->out_corr.p, P * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(shift1d, c->out_shift.p, P * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return CLM_OK;
}

static int check_radius(clm_ctx* c, int32_t r) {
    // cryo_average_clmap's weights W0 = 1.25.^(21:-1:1) cover 1 + 2d <= 21
    if (r < 1 || r > 10) return fail(c, CLM_E_ARG, "map_filter_radius %d outside [1, 10]", r);
    return CLM_OK;
}

int clm_clmatrix_filtered(clm_ctx* c, const clm_params* p, int32_t map_filter_radius, const void* pf, int32_t* cl1,
                          int32_t* cl2, int32_t* jidx, double* corr, double* shift1d, clm_stats* st) {
    if (!c) return CLM_E_ARG;
    int rc = check_radius(c, map_filter_radius);
    if (rc) return rc;
    StackJob ex;
    ex.filt = map_filter_radius;
    return host_job(c, p, pf, cl1, cl2, jidx, corr, shift1d, st, &ex);
}

int clm_clmatrix_filtered_device(clm_ctx* c, const clm_params* p, int32_t map_filter_radius, const void* d_pf,
                                 int32_t* d_cl1, int32_t* d_cl2, int32_t* d_jidx, double* d_corr, double* d_shift1d,
                                 clm_stats* st) {
    if (!c) return CLM_E_ARG;
    if (!p) return fail(c, CLM_E_ARG, "null params");
    int rc = check_radius(c, map_filter_radius);
    if (rc) return rc;
    if (!d_pf && p->K > 1) return fail(c, CLM_E_ARG, "null pf");
    HIPCHK(c, hipSetDevice(c->device));
    StackJob ex;
    ex.filt = map_filter_radius;
    const clm_stack_io io{d_pf, d_cl1, d_cl2, d_jidx, d_corr, d_shift1d};
    return run_job(c, 1, p, &io, false, st, &ex);
}

void clm_set_log(clm_ctx* c, clm_log_fn fn, void* user) {
    if (!c) return;
    c->log_fn = fn;
    c->log_user = user;
}

int clm_clstack_mask(int32_t K, int32_t n_theta, const double* clstack, const double* ref_clmatrix, double* mask,
                     int64_t* matched) {
    if (K < 0 || n_theta <= 0 || n_theta % 2 || (K > 0 && (!clstack || !ref_clmatrix || !mask))) return CLM_E_ARG;
    // cryo_clmatrix_cpu.m:567-617 with n_theta := n_theta/2 (:119): lines are
    // matched when both lie within 5 degrees of the reference's (chord
    // distance on the unit circle), or both within 5 degrees of its antipodes
    const int L = n_theta / 2;
    const double max_angle = 5.0 / 180.0 * M_PI;
    const double tol = 2.0 * std::sin(max_angle / 2.0) + 1.0e-10;
    const double dth = 2.0 * M_PI / (2.0 * L);
    auto chord = [&](double a, double b, bool flip) {
        const double ta = dth * (a - 1.0) + (flip ? M_PI : 0.0), tb = dth * (b - 1.0);
        return std::hypot(std::cos(ta) - std::cos(tb), std::sin(ta) - std::sin(tb));
    };
    int64_t n = 0;
    const size_t KK = (size_t)K * K;
    for (size_t i = 0; i < KK; i++) mask[i] = 0.0;
    for (int k1 = 0; k1 < K; k1++)
        for (int k2 = k1 + 1; k2 < K; k2++) {
            const size_t e12 = (size_t)k2 * K + k1, e21 = (size_t)k1 * K + k2;  // (k1,k2), (k2,k1) column-major
            const double l1 = clstack[e12], l2 = clstack[e21], t1 = ref_clmatrix[e12], t2 = ref_clmatrix[e21];
            const bool same = chord(l1, t1, false) <= tol && chord(l2, t2, false) <= tol;
            const bool flip = chord(l1, t1, true) <= tol && chord(l2, t2, true) <= tol;
            if (same || flip) {
                mask[e12] = 1.0;
                n++;
            }
        }
    if (matched) *matched = n;
    return CLM_OK;
}

int clm_assemble(int32_t K, const int32_t* cl1, const int32_t* cl2, const double* corr, double* clstack,
                 double* corrstack, double* eqmap) {
    return assemble_impl(K, nullptr, cl1, cl2, corr, clstack, corrstack, eqmap);
}

int clm_n_equations(int32_t K, double memoryfactor, int64_t* n) {
    // cryo_estimate_shifts.m:58-94, in the reference's double arithmetic
    if (K < 0 || !n) return CLM_E_ARG;
    if (!(memoryfactor >= 0.0) || (memoryfactor > 1.0 && memoryfactor < 100.0)) return CLM_E_ARG;
    const double k = (double)K;
    const double total = std::ceil(k * (k - 1.0) / 2.0);
    const double memtotal = total * 2.0 * k * 8.0;
    double neq;
    if (memoryfactor <= 1.0) {
        neq = std::ceil(k * (k - 1.0) * memoryfactor / 2.0);
    } else {
        const double f = (memoryfactor * 1e6) / memtotal;
        neq = f < 1.0 ? std::ceil(k * (k - 1.0) * f / 2.0) : total;
    }
    *n = (int64_t)std::max(neq, k);
    return CLM_OK;
}

int clm_equation_pairs(int32_t K, int64_t n, uint64_t seed, uint8_t* sel, int64_t* count) {
    // n of the K(K-1)/2 pairs, every n-subset equally likely (the pairs
    // cryo_estimate_shifts.m:130-137 draws with randperm; MATLAB's stream is
    // not reproduced): Floyd's sampling (J. Bentley and R. Floyd, CACM
    // 30(9), 1987), one draw per pair taken, the selection itself as the
    // membership set.  For j = P-n .. P-1: t uniform in [0, j]; take t, or j
    // if t is taken already.
    if (K < 0 || n < 0) return CLM_E_ARG;
    const long long P = clm_num_pairs(K);
    if (P > 0 && !sel) return CLM_E_ARG;
    const long long want = std::min<long long>(n, P);
    if (P > 0) std::memset(sel, 0, (size_t)P);
    uint64_t st = seed ^ 0x9e3779b97f4a7c15ull;
    for (long long j = P - want; j < P; j++) {
        // uniform in [0, j] by a 64 x 64 -> 128-bit product (relative bias below (j+1)/2^64)
        const long long t = (long long)(((unsigned __int128)splitmix64(st) * (unsigned __int128)(j + 1)) >> 64);
        if (sel[t]) sel[j] = 1;
        else sel[t] = 1;
    }
    if (count) *count = want;
    return CLM_OK;
}

int clm_assemble_subset(int32_t K, const uint8_t* sel, const int32_t* cl1, const int32_t* cl2, const double* corr,
                        double* clstack, double* corrstack, double* eqmap) {
    if (K > 1 && !sel) return CLM_E_ARG;
    return assemble_impl(K, sel, cl1, cl2, corr, clstack, corrstack, eqmap);
}

int clm_partner_subset(int32_t K, int32_t NK, uint64_t seed, uint8_t* sel, int64_t* count) {
    if (K < 0 || NK < -1) return CLM_E_ARG;
    const long long P = clm_num_pairs(K);
    if (P > 0 && !sel) return CLM_E_ARG;
    // cryo_clmatrix_gpu_1.m:150-151,346-350: image k1 (0-based) takes
    // n2 = min(K-1-k1, NK) of its K-1-k1 later images, drawn uniformly
    // without replacement and kept in increasing order; NK = -1 means all.
    // Selection sampling (Knuth, TAOCP vol. 2, 3.4.2, Algorithm S) visits the
    // candidates in order and takes each with probability (needed / left),
    // which yields every n2-subset with equal probability, already sorted.
    std::vector<long long> taken(K > 1 ? K - 1 : 0, 0);
    host_parallel(K > 1 ? K - 1 : 0, 64, [&](long long r0, long long r1) {
        for (long long k1 = r0; k1 < r1; k1++) {
            const long long m = K - 1 - k1;
            const long long n2 = NK < 0 ? m : std::min<long long>(m, NK);
            uint8_t* row = sel + pair_slot(k1, k1 + 1, K);
            if (n2 == m) {
                std::memset(row, 1, (size_t)m);
                taken[k1] = m;
                continue;
            }
            uint64_t st = seed ^ (0xd1b54a32d192ed03ull * (uint64_t)(k1 + 1));
            long long got = 0;
            for (long long i = 0; i < m; i++) {
                const double u = (double)(splitmix64(st) >> 11) * 0x1p-53;  // [0, 1)
                const bool take = (double)(m - i) * u < (double)(n2 - got);
                row[i] = take ? 1 : 0;
                got += take;
            }
            taken[k1] = got;
        }
    });
    if (count) {
        long long n = 0;
        for (long long t : taken) n += t;
        *count = n;
    }
    return CLM_OK;
}

int clm_shift_equations_csc(int32_t K, int32_t n_theta, const int32_t* cl1, const int32_t* cl2,
                            const double* shift1d, int64_t* jc, int64_t* ir, double* pr, int64_t* nnz) {
    return shift_equations_impl(K, n_theta, nullptr, 0, cl1, cl2, shift1d, jc, ir, pr, nnz);
}

int clm_shift_equations_csc_ex(int32_t K, int32_t n_theta, const uint8_t* sel, int32_t flags, const int32_t* cl1,
                               const int32_t* cl2, const double* shift1d, int64_t* jc, int64_t* ir, double* pr,
                               int64_t* nnz) {
    if (flags & ~(CLM_FLIP_BY_INDEX | CLM_COMPACT_RECORDS)) return CLM_E_ARG;
    if ((flags & CLM_COMPACT_RECORDS) && !sel && K > 1) return CLM_E_ARG;
    return shift_equations_impl(K, n_theta, sel, flags, cl1, cl2, shift1d, jc, ir, pr, nnz);
}

int clm_shift_equations_csc_subset(int32_t K, int32_t n_theta, const uint8_t* sel, const int32_t* cl1,
                                   const int32_t* cl2, const double* shift1d, int64_t* jc, int64_t* ir, double* pr,
                                   int64_t* nnz) {
    if (K > 1 && !sel) return CLM_E_ARG;
    return shift_equations_impl(K, n_theta, sel, 0, cl1, cl2, shift1d, jc, ir, pr, nnz);
}

}  // extern "C"

namespace {
// shift_equations in CSC form; with a selection, only the selected pairs'
// rows, numbered in pair order (cryo_clmatrix_gpu_1.m:809-827,853-858)
int shift_equations_impl(int32_t K, int32_t n_theta, const uint8_t* sel, int32_t flags, const int32_t* cl1,
                         const int32_t* cl2, const double* shift1d, int64_t* jc, int64_t* ir, double* pr,
                         int64_t* nnz) {
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
        // the second line keeps its orientation: angle test (cryo_clmatrix)
        // or index test (cryo_estimate_shifts, CLM_FLIP_BY_INDEX)
        lowb[cl] = (flags & CLM_FLIP_BY_INDEX) ? cl <= L : b < M_PI;
        if (!lowb[cl]) b = b - M_PI;
        sb[cl] = std::sin(b);
        cb[cl] = std::cos(b);
    }
    const long long P = clm_num_pairs(K);
    const std::vector<int32_t> rank = selected_rank(P, sel);
    auto row_of = [&](long long e) -> long long { return sel ? (long long)rank[(size_t)e] : e; };
    // the records of pair e: at e, or at its row among the selected pairs
    // when only those are given (CLM_COMPACT_RECORDS)
    const bool compact = (flags & CLM_COMPACT_RECORDS) != 0;
    auto rec_of = [&](long long e) -> long long { return compact ? row_of(e) : e; };
    // coefficient of pair e on (image k1: x,y) and (image k2: x,y)   :731-738
    auto coefs = [&](long long e, double out[4]) {
        const long long q = rec_of(e);
        const int a = cl1[q], b = cl2[q];
        if (a < 0 || a > 2 * L || b < 0 || b > 2 * L) { out[0] = out[1] = out[2] = out[3] = NAN; return; }
        if (lowb[b]) { out[0] = sa[a]; out[1] = ca[a]; }
        else { out[0] = -sa[a]; out[1] = -ca[a]; }
        out[2] = -sb[b];
        out[3] = -cb[b];
    };
    const long long ncol = 2 * (long long)K;
    // The nonzeros in pair order: pair e = (k1, k2) has its coefficients in
    // columns 2k1 + xy (c4[xy]) and 2k2 + xy (c4[2 + xy]) and, when its 1-D
    // shift is nonzero, an entry in column 2K (the right-hand side).  Walking
    // the pairs in order appends every column's rows in increasing order, as
    // the compressed-column form needs.  Chunks of consecutive pairs are
    // counted per column, then filled from offsets that follow the chunk
    // order, so the result does not depend on the number of host threads.
    const long long nch = P == 0 ? 0 : std::min<long long>(std::max<long long>(P / 262144, 1), 256);
    const long long W = ncol + 1;
    std::vector<long long> cnt((size_t)(nch * W), 0);
    auto first_of = [&](long long k) { return k * K - k * (k + 1) / 2; };
    auto walk = [&](long long c, auto&& emit) {
        const long long e0 = P * c / nch, e1 = P * (c + 1) / nch;
        if (e0 >= e1) return;
        const double t = 2.0 * K - 1.0;
        long long k1 = (long long)std::floor((t - std::sqrt(std::max(t * t - 8.0 * (double)e0, 0.0))) / 2.0);
        k1 = std::min<long long>(std::max<long long>(k1, 0), K - 2);
        while (k1 > 0 && first_of(k1) > e0) k1--;
        while (k1 + 1 <= K - 2 && first_of(k1 + 1) <= e0) k1++;
        long long k2 = e0 - first_of(k1) + k1 + 1;
        double c4[4];
        for (long long e = e0; e < e1; e++) {
            if (!sel || sel[e]) {
                const long long row = row_of(e);
                coefs(e, c4);
                for (int xy = 0; xy < 2; xy++)
                    if (c4[xy] != 0.0) emit(2 * k1 + xy, row, c4[xy]);
                for (int xy = 0; xy < 2; xy++)
                    if (c4[2 + xy] != 0.0) emit(2 * k2 + xy, row, c4[2 + xy]);
                const double b = shift1d[rec_of(e)];
                if (b != 0.0) emit(ncol, row, b);
            }
            if (++k2 == K) {
                k1++;
                k2 = k1 + 1;
            }
        }
    };
    host_parallel(nch, 1, [&](long long c0, long long c1) {
        for (long long c = c0; c < c1; c++) {
            long long* cc = cnt.data() + c * W;
            walk(c, [&](long long col, long long, double) { cc[col]++; });
        }
    });
    // column starts, and each chunk's offset within every column
    std::vector<long long> start((size_t)W + 1, 0);
    for (long long col = 0; col < W; col++) {
        long long n = start[col];
        for (long long c = 0; c < nch; c++) {
            const long long k = cnt[c * W + col];
            cnt[c * W + col] = n;
            n += k;
        }
        start[col + 1] = n;
    }
    if (jc)
        for (long long q = 0; q <= ncol + 1; q++) jc[q] = start[q];
    if (nnz) *nnz = start[W];
    if (ir && pr) {
        host_parallel(nch, 1, [&](long long c0, long long c1) {
            for (long long c = c0; c < c1; c++) {
                long long* off = cnt.data() + c * W;
                walk(c, [&](long long col, long long row, double v) {
                    const long long i = off[col]++;
                    ir[i] = row;
                    pr[i] = v;
                });
            }
        });
    }
    return CLM_OK;
}
}  // namespace
