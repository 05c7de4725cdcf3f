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

// Everything the kernels need about one stack (one cryo_clmatrix call).  A
// launch serves one or more stacks of the same kernel variant; tiles and
// list entries name their stack.
struct StackDev {
    Geom g;
    const void* astk;         // [K][NAT] A-tiles of the operand type (or [n][NAT], see amap)
    const int* amap;          // null, or image -> its A-tiles' index in astk (A-stacks built
                              // for some images only: an FP32 twin's, for its uncertain pairs)
    const void* Bop;          // [K][2][Lp][Rp] operand planes
    const float* nrm;         // [K] largest half-line norm per image
    PairRec* recs;            // [P] pass-1 records
    const float2* P32;        // [K][L][Rp]
    const double2* P64;       // [K][L][R]
    const float2* ph32;       // [NE+1][Rp], row NE zero
    const double2* ph64;      // [S][R]
    const double* shiftv;     // [S]
    int32_t* cl1;             // outputs, [P] in equation order
    int32_t* cl2;
    int32_t* jidx;
    double* corr;
    double* shift1d;
    float coef, sub;          // error bound: delta = coef*n1*n2 + sub*(n1+n2)
    float tau_t;              // tie margin in t units (tau/2)
    float tau;                // FP64 re-score margin (correlation units)
    int fp64;                 // re-score near ties in FP64
    int CPP;                  // waves per pair
    long long pair_base;      // first flat pair index of this stack within its group
    long long npairs;
    int filt;                 // map_filter_radius d (0: no map filter)
    const double* wts;        // (2d+1)^2 filter weights, row-major
};

// Pass-2 / fallback list entry: (stack, slot)
struct SlotRef {
    int32_t stack;
    int32_t slot;
};

}  // namespace clm
