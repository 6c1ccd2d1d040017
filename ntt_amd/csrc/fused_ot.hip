// Instantiations of the fused limb kernel with on-the-fly twiddles in the
// low-bit round: forward with lazy reduction at window width 7 (every
// q < 2^30), forward canonical and inverse canonical at width 6.  A
// translation unit of its own, so it compiles in parallel with the others.
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

FusedFn fused_fn_ot(int n, bool inv, bool lazy) {
    if (inv) {
        switch (n) {
            case 10: return fused_kernel<10, true, false, 6, false, true>;
            case 11: return fused_kernel<11, true, false, 6, false, true>;
            case 12: return fused_kernel<12, true, false, 6, false, true>;
            case 13: return fused_kernel<13, true, false, 6, false, true>;
            case 14: return fused_kernel<14, true, false, 6, false, true>;
            case 15: return fused_kernel<15, true, false, 6, false, true>;
            case 16: return fused_kernel<16, true, false, 6, false, true>;
            default: return nullptr;
        }
    }
    if (lazy) {
        switch (n) {
            case 10: return fused_kernel<10, false, true, 7, false, true>;
            case 11: return fused_kernel<11, false, true, 7, false, true>;
            case 12: return fused_kernel<12, false, true, 7, false, true>;
            case 13: return fused_kernel<13, false, true, 7, false, true>;
            case 14: return fused_kernel<14, false, true, 7, false, true>;
            case 15: return fused_kernel<15, false, true, 7, false, true>;
            case 16: return fused_kernel<16, false, true, 7, false, true>;
            default: return nullptr;
        }
    }
    switch (n) {
        case 10: return fused_kernel<10, false, false, 6, false, true>;
        case 11: return fused_kernel<11, false, false, 6, false, true>;
        case 12: return fused_kernel<12, false, false, 6, false, true>;
        case 13: return fused_kernel<13, false, false, 6, false, true>;
        case 14: return fused_kernel<14, false, false, 6, false, true>;
        case 15: return fused_kernel<15, false, false, 6, false, true>;
        case 16: return fused_kernel<16, false, false, 6, false, true>;
        default: return nullptr;
    }
}

}  // namespace ntt
