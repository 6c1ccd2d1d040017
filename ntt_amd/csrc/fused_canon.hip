// Instantiations of the fused limb kernel, canonical arithmetic (one translation
// unit per arithmetic variant, so the two compile in parallel).
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

FusedFn fused_fn_canon(int n, bool inv) {
    switch (n * 2 + (inv ? 1 : 0)) {
        case 20: return fused_kernel<10, false, false>;
        case 21: return fused_kernel<10, true, false>;
        case 22: return fused_kernel<11, false, false>;
        case 23: return fused_kernel<11, true, false>;
        case 24: return fused_kernel<12, false, false>;
        case 25: return fused_kernel<12, true, false>;
        case 26: return fused_kernel<13, false, false>;
        case 27: return fused_kernel<13, true, false>;
        case 28: return fused_kernel<14, false, false>;
        case 29: return fused_kernel<14, true, false>;
        case 30: return fused_kernel<15, false, false>;
        case 31: return fused_kernel<15, true, false>;
        case 32: return fused_kernel<16, false, false>;
        case 33: return fused_kernel<16, true, false>;
        default: return nullptr;
    }
}

}  // namespace ntt
