// pair_variants.hip -- instantiates the pair kernel and A-stack templates of
// clm_pair.h for one operand type (CLM_VARIANT_OP: f32, bf16 or f16) and one
// shift-grid symmetry (CLM_VARIANT_SYM: 0 or 1).  The build compiles this
// file once per combination, in parallel.
#include "clm_pair.h"

#if !defined(CLM_VARIANT_OP) || !defined(CLM_VARIANT_SYM)
#error "compile with -DCLM_VARIANT_OP=<f32|bf16|f16> -DCLM_VARIANT_SYM=<0|1>"
#endif

#define CLM_CAT2(a, b, c) a##b##_##c
#define CLM_CAT(a, b, c) CLM_CAT2(a, b, c)

namespace clm {
namespace {

using f32 = float;
using bf16 = __bf16;
using f16 = _Float16;
using VOp = CLM_VARIANT_OP;
constexpr bool VSYM = CLM_VARIANT_SYM != 0;

template <int CT, int NK>
PairFn pk(bool emit) {
    return emit ? pair_kernel<VOp, CT, NK, VSYM, true> : pair_kernel<VOp, CT, NK, VSYM, false>;
}

}  // namespace

#define CLM_OPNAME CLM_VARIANT_OP

PairFn CLM_CAT(pick_pair_, CLM_OPNAME, CLM_VARIANT_SYM)(int ct, int NK, bool emit) {
    if constexpr (sizeof(VOp) == 4) {
        switch (ct * 100 + NK) {
            case 108: return pk<1, 8>(emit);
            case 116: return pk<1, 16>(emit);
            case 124: return pk<1, 24>(emit);
            case 132: return pk<1, 32>(emit);
            case 148: return pk<1, 48>(emit);
            case 164: return pk<1, 64>(emit);
            case 208: return pk<2, 8>(emit);
            case 216: return pk<2, 16>(emit);
        }
    } else {
        switch (ct * 100 + NK) {
            case 101: return pk<1, 1>(emit);
            case 102: return pk<1, 2>(emit);
            case 103: return pk<1, 3>(emit);
            case 104: return pk<1, 4>(emit);
            case 106: return pk<1, 6>(emit);
            case 108: return pk<1, 8>(emit);
            case 201: return pk<2, 1>(emit);
            case 202: return pk<2, 2>(emit);
            case 203: return pk<2, 3>(emit);
            case 204: return pk<2, 4>(emit);
            case 206: return pk<2, 6>(emit);
            case 208: return pk<2, 8>(emit);
            case 301: return pk<3, 1>(emit);
            case 302: return pk<3, 2>(emit);
            case 303: return pk<3, 3>(emit);
            case 304: return pk<3, 4>(emit);
        }
    }
    return nullptr;
}

AstkFn CLM_CAT(pick_astk_, CLM_OPNAME, CLM_VARIANT_SYM)(int NK) {
    if constexpr (sizeof(VOp) == 4) {
        switch (NK) {
            case 8: return astack_kernel<VOp, 8, VSYM>;
            case 16: return astack_kernel<VOp, 16, VSYM>;
            case 24: return astack_kernel<VOp, 24, VSYM>;
            case 32: return astack_kernel<VOp, 32, VSYM>;
            case 48: return astack_kernel<VOp, 48, VSYM>;
            case 64: return astack_kernel<VOp, 64, VSYM>;
        }
    } else {
        switch (NK) {
            case 1: return astack_kernel<VOp, 1, VSYM>;
            case 2: return astack_kernel<VOp, 2, VSYM>;
            case 3: return astack_kernel<VOp, 3, VSYM>;
            case 4: return astack_kernel<VOp, 4, VSYM>;
            case 6: return astack_kernel<VOp, 6, VSYM>;
            case 8: return astack_kernel<VOp, 8, VSYM>;
        }
    }
    return nullptr;
}

}  // namespace clm
