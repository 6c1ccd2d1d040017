// Instantiations of the fused limb kernel with 2^15 words per workgroup (two
// workgroups per CU), sizes 2^10 .. 2^15, window widths 6 and 7, canonical
// and lazy (q < 2^30) arithmetic in both directions.
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

namespace {
template <int EB>
FusedFn pick_w15(int n, bool inv, bool lazy) {
    if (lazy) {
        if (inv) {
            switch (n) {
                case 10: return fused_kernel<10, true, true, EB, false, false, 15>;
                case 11: return fused_kernel<11, true, true, EB, false, false, 15>;
                case 12: return fused_kernel<12, true, true, EB, false, false, 15>;
                case 13: return fused_kernel<13, true, true, EB, false, false, 15>;
                case 14: return fused_kernel<14, true, true, EB, false, false, 15>;
                case 15: return fused_kernel<15, true, true, EB, false, false, 15>;
                default: return nullptr;
            }
        }
        switch (n) {
            case 10: return fused_kernel<10, false, true, EB, false, false, 15>;
            case 11: return fused_kernel<11, false, true, EB, false, false, 15>;
            case 12: return fused_kernel<12, false, true, EB, false, false, 15>;
            case 13: return fused_kernel<13, false, true, EB, false, false, 15>;
            case 14: return fused_kernel<14, false, true, EB, false, false, 15>;
            case 15: return fused_kernel<15, false, true, EB, false, false, 15>;
            default: return nullptr;
        }
    }
    switch (n * 2 + (inv ? 1 : 0)) {
        case 20: return fused_kernel<10, false, false, EB, false, false, 15>;
        case 21: return fused_kernel<10, true, false, EB, false, false, 15>;
        case 22: return fused_kernel<11, false, false, EB, false, false, 15>;
        case 23: return fused_kernel<11, true, false, EB, false, false, 15>;
        case 24: return fused_kernel<12, false, false, EB, false, false, 15>;
        case 25: return fused_kernel<12, true, false, EB, false, false, 15>;
        case 26: return fused_kernel<13, false, false, EB, false, false, 15>;
        case 27: return fused_kernel<13, true, false, EB, false, false, 15>;
        case 28: return fused_kernel<14, false, false, EB, false, false, 15>;
        case 29: return fused_kernel<14, true, false, EB, false, false, 15>;
        case 30: return fused_kernel<15, false, false, EB, false, false, 15>;
        case 31: return fused_kernel<15, true, false, EB, false, false, 15>;
        default: return nullptr;
    }
}
}  // namespace

FusedFn fused_fn_w15(int n, bool inv, int eb, bool lazy) {
    return eb == 6 ? pick_w15<6>(n, inv, lazy) : eb == 7 ? pick_w15<7>(n, inv, lazy) : nullptr;
}

}  // namespace ntt
