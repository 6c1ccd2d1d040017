// Instantiations of the fused limb kernel's forward mod-down form, which
// stores (sub - NTT(src)) * post per limb: window width 7 with lazy
// reduction (every q < 2^30) and window width 6 with canonical arithmetic
// (any q < 2^31).  A translation unit of its own, so it compiles in parallel
// with the plain shapes.
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

FusedFn fused_fn_md(int n, bool lazy) {
    if (lazy) {
        switch (n) {
            case 10: return fused_kernel<10, false, true, 7, true>;
            case 11: return fused_kernel<11, false, true, 7, true>;
            case 12: return fused_kernel<12, false, true, 7, true>;
            case 13: return fused_kernel<13, false, true, 7, true>;
            case 14: return fused_kernel<14, false, true, 7, true>;
            case 15: return fused_kernel<15, false, true, 7, true>;
            case 16: return fused_kernel<16, false, true, 7, true>;
            default: return nullptr;
        }
    }
    switch (n) {
        case 10: return fused_kernel<10, false, false, 6, true>;
        case 11: return fused_kernel<11, false, false, 6, true>;
        case 12: return fused_kernel<12, false, false, 6, true>;
        case 13: return fused_kernel<13, false, false, 6, true>;
        case 14: return fused_kernel<14, false, false, 6, true>;
        case 15: return fused_kernel<15, false, false, 6, true>;
        case 16: return fused_kernel<16, false, false, 6, true>;
        default: return nullptr;
    }
}

}  // namespace ntt
