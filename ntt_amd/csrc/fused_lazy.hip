// Instantiations of the fused limb kernel, lazy-reduction arithmetic (one translation
// unit per arithmetic variant, so the two compile in parallel).
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

FusedFn fused_fn_lazy(int n, bool inv) {
    switch (n * 2 + (inv ? 1 : 0)) {
        case 20: return fused_kernel<10, false, true>;
        case 21: return fused_kernel<10, true, true>;
        case 22: return fused_kernel<11, false, true>;
        case 23: return fused_kernel<11, true, true>;
        case 24: return fused_kernel<12, false, true>;
        case 25: return fused_kernel<12, true, true>;
        case 26: return fused_kernel<13, false, true>;
        case 27: return fused_kernel<13, true, true>;
        case 28: return fused_kernel<14, false, true>;
        case 29: return fused_kernel<14, true, true>;
        case 30: return fused_kernel<15, false, true>;
        case 31: return fused_kernel<15, true, true>;
        case 32: return fused_kernel<16, false, true>;
        case 33: return fused_kernel<16, true, true>;
        default: return nullptr;
    }
}

}  // namespace ntt
