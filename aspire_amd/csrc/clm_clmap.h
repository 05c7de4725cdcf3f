// clm_clmap.h -- the common-lines search with map_filter_radius > 0
// (cryo_clmatrix_cpu.m:383-385): every shift's correlation map is filtered
// by cryo_average_clmap (sinograms/cryo_average_clmap.m:20-48) before the
// first-maximum search, and the filtered value is what the pair keeps.
//
// The filter replaces each entry by a weighted sum over its (2d+1)^2
// neighbourhood in the map extended to 2L rows (rows L..2L-1 repeat the map
// with its column halves swapped, :30), circularly:
//   out(l1, c) = sum_{i,j=0..2d} W(i,j) map_ext((l1-d+i) mod 2L, (c-d+j) mod 2L)
// A neighbourhood sum cannot be bounded from one cell's maximum, so this
// path does not use the MFMA engine's certified pruning: it evaluates every
// map entry in FP64 with the oracle's arithmetic and order (as exact64_kernel
// does) and sums the stencil in a fixed order, so its results equal the
// FP64 oracle's bit for bit.
//
// One workgroup per pair (persistent over pairs).  Per shift, the map is
// computed in 48 x 80 tiles of the extended map (output tiles of
// (48 - 2d) x (80 - 2d) with their halo), r in chunks of 32 staged in LDS:
// each thread owns a 4 x 4 block of the tile.  map_ext(r', c) is
// C_{h(r') xor h(c)}(r' mod L, c mod L) with h(x) = x >= L, so the row
// operand is P1 .* ph or conj(P1) .* ph by that parity.
#pragma once

namespace clm {

constexpr int CM_R = 48, CM_C = 80, CM_RC = 32;

struct ClmapArgs {
    const StackDev* stacks;
    int stack0, nstacks;
    long long n_all;
};

__global__ __launch_bounds__(256) void clmap_kernel(ClmapArgs a) {
    __shared__ double2 sA[CM_R][CM_RC];   // P1 .* ph
    __shared__ double2 sF[CM_R][CM_RC];   // conj(P1) .* ph
    __shared__ double2 sB[CM_C][CM_RC];   // P2
    __shared__ double sM[CM_R][CM_C + 1];  // map tile
    __shared__ double s_v[256];
    __shared__ long long s_k[256];
    const int tid = threadIdx.x;
    const int bi = tid / 20, bj = tid % 20;  // 12 x 20 blocks of 4 x 4 (threads 240..255 idle)
    for (long long item = blockIdx.x; item < a.n_all; item += gridDim.x) {
        const int stk = a.stack0 + (a.nstacks > 1 ? find_stack(a.stacks + a.stack0, a.nstacks, item) : 0);
        const StackDev& sd = a.stacks[stk];
        const Geom& g = sd.g;
        const int slot = (int)(item - sd.pair_base);
        int k1, k2;
        slot_to_pair(slot, g.K, k1, k2);
        const int L = g.L, L2 = 2 * L, R = g.R, d = sd.filt, w = 2 * d + 1;
        const int OR = CM_R - 2 * d, OC = CM_C - 2 * d;
        const double2* P1 = sd.P64 + (size_t)k1 * L * R;
        const double2* P2 = sd.P64 + (size_t)k2 * L * R;
        double bv = -INFINITY;
        long long bk = 0x7fffffffffffffffLL;
        for (int j = 0; j < g.S; j++) {
            const double2* ph = sd.ph64 + (size_t)j * R;
            for (int row0 = 0; row0 < L; row0 += OR) {
                for (int col0 = 0; col0 < L2; col0 += OC) {
                    double acc[4][4];
#pragma unroll
                    for (int x = 0; x < 4; x++)
#pragma unroll
                        for (int y = 0; y < 4; y++) acc[x][y] = 0.0;
                    // parity of each owned entry: flipped row operand?
                    bool flip[4][4];
#pragma unroll
                    for (int x = 0; x < 4; x++) {
                        const int rp = ((row0 - d + 4 * bi + x) % L2 + L2) % L2;
#pragma unroll
                        for (int y = 0; y < 4; y++) {
                            const int cc = ((col0 - d + 4 * bj + y) % L2 + L2) % L2;
                            flip[x][y] = (rp >= L) != (cc >= L);
                        }
                    }
                    for (int r0 = 0; r0 < R; r0 += CM_RC) {
                        const int nr = R - r0 < CM_RC ? R - r0 : CM_RC;
                        __syncthreads();
                        for (int q = tid; q < CM_R * CM_RC; q += 256) {
                            const int ti = q / CM_RC, rr = q % CM_RC;
                            if (rr < nr) {
                                const int rp = ((row0 - d + ti) % L2 + L2) % L2;
                                const double2 p = P1[(size_t)(rp % L) * R + r0 + rr];
                                const double2 f = ph[r0 + rr];
                                sA[ti][rr] = make_double2(p.x * f.x - p.y * f.y, p.x * f.y + p.y * f.x);
                                sF[ti][rr] = make_double2(p.x * f.x + p.y * f.y, p.x * f.y - p.y * f.x);
                            }
                        }
                        for (int q = tid; q < CM_C * CM_RC; q += 256) {
                            const int tj = q / CM_RC, rr = q % CM_RC;
                            if (rr < nr) {
                                const int cc = ((col0 - d + tj) % L2 + L2) % L2;
                                sB[tj][rr] = P2[(size_t)(cc % L) * R + r0 + rr];
                            }
                        }
                        __syncthreads();
                        if (tid < 240) {
                            for (int rr = 0; rr < nr; rr++) {
                                double2 xa[4], xf[4], xb[4];
#pragma unroll
                                for (int x = 0; x < 4; x++) {
                                    xa[x] = sA[4 * bi + x][rr];
                                    xf[x] = sF[4 * bi + x][rr];
                                    xb[x] = sB[4 * bj + x][rr];
                                }
#pragma unroll
                                for (int x = 0; x < 4; x++)
#pragma unroll
                                    for (int y = 0; y < 4; y++) {
                                        const double2 o = flip[x][y] ? xf[x] : xa[x];
                                        const double t = o.x * xb[y].x;
                                        const double u = o.y * xb[y].y;
                                        acc[x][y] = acc[x][y] + (t + u);
                                    }
                            }
                        }
                    }
                    __syncthreads();
                    if (tid < 240) {
#pragma unroll
                        for (int x = 0; x < 4; x++)
#pragma unroll
                            for (int y = 0; y < 4; y++) sM[4 * bi + x][4 * bj + y] = 2.0 * acc[x][y];
                    }
                    __syncthreads();
                    // stencil over this tile's outputs
                    for (int o = tid; o < OR * OC; o += 256) {
                        const int oi = o / OC, oj = o % OC;
                        const int l1 = row0 + oi, c = col0 + oj;
                        if (l1 >= L || c >= L2) continue;
                        double v = 0.0;
                        for (int i = 0; i < w; i++)
                            for (int jx = 0; jx < w; jx++) v = v + sd.wts[i * w + jx] * sM[oi + i][oj + jx];
                        if (isnan(v)) continue;
                        const long long key = ((long long)j * L2 + c) * L + l1;
                        if (better64(v, key, bv, bk)) {
                            bv = v;
                            bk = key;
                        }
                    }
                }
            }
        }
        // block-wide best (value, then smallest key)
        s_v[tid] = bv;
        s_k[tid] = bk;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s && better64(s_v[tid + s], s_k[tid + s], s_v[tid], s_k[tid])) {
                s_v[tid] = s_v[tid + s];
                s_k[tid] = s_k[tid + s];
            }
            __syncthreads();
        }
        if (tid == 0) write_result(sd, slot, s_v[0], s_k[0]);
        __syncthreads();
    }
}

}  // namespace clm
