// Instantiations of the fused limb kernel's forward mod-down form, which
// stores (sub - NTT(src)) * post per limb: window width 6 (1024 threads x 64
// words; from N = 2^12, one limb per wave, the stores go through the in-wave
// transpose) with lazy reduction (every q < 2^30) or canonical arithmetic
// (any q < 2^31), except lazy N = 2^10, 2^11 at width 7 (512 threads x 128
// words), where width 6 spills.  A translation unit of
// its own, so it compiles in parallel with the plain shapes.
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

int fused_md_eb(int n, bool lazy) { return lazy && n < 12 ? 7 : 6; }

FusedFn fused_fn_md(int n, bool lazy) {
    if (lazy) {
        switch (n) {
            case 10: return fused_kernel<10, false, true, 7, true>;
            case 11: return fused_kernel<11, false, true, 7, true>;
            case 12: return fused_kernel<12, false, true, 6, true>;
            case 13: return fused_kernel<13, false, true, 6, true>;
            case 14: return fused_kernel<14, false, true, 6, true>;
            case 15: return fused_kernel<15, false, true, 6, true>;
            case 16: return fused_kernel<16, false, true, 6, true>;
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
