// clm_device.h -- device-side types shared by the HIP kernels of libclm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace clm {

// Per-pair record of the first pass.  A "cell" is what one lane of a wave
// accumulates for one 32x32 output tile: 16 rows (e,l1) x 1 column l2, each
// row holding both maps (C1 and C2) through t = X + |Y|.
constexpr int MAXC = 58;          // cells kept per pair (record = 256 B)
constexpr int MAXE = 1024;        // candidates kept per pair in the second pass

constexpr int F_UNCERTAIN = 1;    // a wave's cell list overflowed: a cell may be missing
constexpr int F_OVERFLOW = 2;     // more than MAXC cells reached the final threshold
constexpr int F_EMPTY = 4;        // no finite candidate at all

struct alignas(16) PairRec {
    float M;         // max of t over all candidates (operand precision)
    float T;         // certified threshold used for this pair
    int32_t k1, k2;  // images of the pair (0-based)
    int32_t flags;
    int32_t ncells;
    uint32_t cell[MAXC];  // (cell_id << 6) | lane, cell_id = at*NCT + ct
};
static_assert(sizeof(PairRec) == 256, "PairRec layout");

struct EmitRec {
    int32_t n;
    int32_t overflow;
    uint32_t cand[MAXE];  // row * Lp + col
};

// Geometry of one run (all device kernels read it by value).
struct Geom {
    int K, n_r, T, L, R, Rp, Lp, NCT;
    int S;        // shifts of the reference grid
    int NE;       // evaluation shifts (S if the grid is symmetric, else 2S)
    int sym;      // 1: grid symmetric, t = X + |Y|;  0: t = X + Y, C2 rows negated
    int NROWS;    // NE * L
    int NAT;      // ceil(NROWS / 32)
};

}  // namespace clm
