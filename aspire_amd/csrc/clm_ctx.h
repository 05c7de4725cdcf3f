// clm_ctx.h -- host-side context of libclm shared by its translation units:
// the context (one HIP stream, grow-only device buffers, caches) and the
// error/allocation helpers every C-ABI entry point uses.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "clm.h"

namespace clmh {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
};

struct StackBufs {
    DevBuf P64, P32, Bop, nrm, astk, recs, ph32, ph64, shiftv, H, W;
    DevBuf amap;  // a twin's image -> A-stack index map (K ints), then its image list
};

// Pass-1 tiles of one group, kept while the group's shapes repeat.
struct TileCache {
    std::vector<long long> key;  // (table index, K, PPW) per member
    DevBuf tiles;
    int ntiles = 0;
};

}  // namespace clmh

struct clm_ctx {
    int device = 0;
    int n_cu = 0;
    int lds_max = 65536;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<clmh::StackBufs> sb;          // per stack of the current call
    std::vector<clmh::TileCache> tcache;      // most recent first
    clmh::DevBuf table, ctr, ulist, erecs, plist, tiles2, fblist;
    clmh::DevBuf pf, out_cl1, out_cl2, out_j, out_corr, out_shift;  // host-buffer form
    hipEvent_t ev[6] = {};
    // windowed search: window records, per-line shifts, flags, selected-pair lists
    clmh::DevBuf w_cl1, w_cl2, w_j, w_corr, w_shift, w_lsh, w_flag, w_rows, w_list, w_plist, w_tiles, w_counts,
        w_shiftv;
    // refinement of uncertain pairs by the FP32 twins: one-pair tiles and pairs
    clmh::DevBuf tw_tiles, tw_plist;
    // FP64 exhaustive search: per (item, shift range) best values and keys
    clmh::DevBuf ex_v, ex_k;
    // shift network: activations and gradients of the last batch
    clmh::DevBuf sn_h1, sn_h2, sn_z, sn_y, sn_dy, sn_dz, sn_dz2, sn_dz1;
    // polar Fourier transform: per-axis tables, staged images, host-form buffers;
    // the NUFFT method's oversampled grids, windows, kernel weights and
    // deconvolution factors, and its hipFFT plans (opaque handles) with their
    // grid side and batch
    clmh::DevBuf pft_ax, pft_ay, pft_xt, pft_img, pft_out;
    clmh::DevBuf pft_grid, pft_base, pft_wt, pft_dec;
    void* fft_plan[2] = {nullptr, nullptr};
    int fft_side[2] = {0, 0}, fft_batch[2] = {0, 0};
    // shift least squares: packed lines, u, per-tile sums, 2K vectors and the tiles' partial
    // A'u sums, tables, scalar state, pair tiles
    clmh::DevBuf ls_pack, ls_u, ls_part, ls_vec, ls_tab, ls_state, ls_tiles;
    clmh::DevBuf es_rot, es_x;  // host-buffer form of cryo_estimate_shifts: rotations, solution
    int ls_tiles_K = 0, ls_ntiles = 0;  // the pair tiles held in ls_tiles are for this K
    // voting: clstack as 16-bit tables, row-major and transposed; per-pair histogram
    // peaks; cos and sin of the line-difference angles
    clmh::DevBuf vt_m, vt_mt, vt_pk, vt_cs;
    // sync3n: symmetric copies of R (9 planes), of v and of the pair scores;
    // pair tiles (for s3_nb blocks per side), histogram thresholds and counts
    clmh::DevBuf s3_rs, s3_vs, s3_ss, s3_tiles, s3_thr, s3_hist;
    int s3_nb = 0;
    // autotuned variants: shape key -> (ct, W, lag)
    std::vector<std::pair<unsigned long long, int4>> tuned;
    clmh::DevBuf tune_tiles;
    // captured device job of the grouped form, replayed while the signature repeats
    std::vector<unsigned long long> gsig;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // verbose reports and warnings (clm_set_log)
    clm_log_fn log_fn = nullptr;
    void* log_user = nullptr;
    // allocation-free calls: this device's copy of the library's static scratch
    // array (from hipGetSymbolAddress at clm_create) and, while such a call
    // runs, the bump region every device buffer is carved from
    void* scratch = nullptr;
    size_t scratch_cap = 0;
    char* arena = nullptr;
    size_t arena_cap = 0, arena_off = 0;
};

namespace clmh {

// Releases the hipFFT plans of the polar Fourier transform (pft.hip).
void pft_release(clm_ctx* c);

inline int fail(clm_ctx* c, int code, const char* fmt, ...) {
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
            return ::clmh::fail((ctx), CLM_E_HIP, "%s failed: %s (%s:%d)", #expr,      \
                                hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

inline int ensure(clm_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.n >= bytes) return CLM_OK;
    if (c->arena) {
        // allocation-free call: carve from the scratch array, never hipMalloc/hipFree
        const size_t off = (c->arena_off + 255) & ~(size_t)255;
        if (off > c->arena_cap || bytes > c->arena_cap - off)
            return fail(c, CLM_E_UNSUP,
                        "allocation-free call: the shape needs more device memory than the %zu-byte scratch "
                        "array (%zu bytes carved, %zu more requested)",
                        c->arena_cap, off, bytes);
        b.p = c->arena + off;
        b.n = bytes;
        c->arena_off = off + bytes;
        return CLM_OK;
    }
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

inline void release(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

}  // namespace clmh
