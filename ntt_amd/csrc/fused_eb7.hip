// Instantiations of the fused limb kernel with window width 7
// (512 threads x 128 words per thread): canonical arithmetic in
// both directions, and lazy reduction (q < 2^30) in both directions.  One
// translation unit per shape, so the two compile in parallel.
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

FusedFn fused_fn_eb7(int n, bool inv, bool lazy) {
    constexpr int EB = 7;
    if (lazy) {
        if (inv) {
            switch (n) {
                case 10: return fused_kernel<10, true, true, EB>;
                case 11: return fused_kernel<11, true, true, EB>;
                case 12: return fused_kernel<12, true, true, EB>;
                case 13: return fused_kernel<13, true, true, EB>;
                case 14: return fused_kernel<14, true, true, EB>;
                case 15: return fused_kernel<15, true, true, EB>;
                case 16: return fused_kernel<16, true, true, EB>;
                default: return nullptr;
            }
        }
        switch (n * 2) {
        case 20: return fused_kernel<10, false, true, EB>;
        case 22: return fused_kernel<11, false, true, EB>;
        case 24: return fused_kernel<12, false, true, EB>;
        case 26: return fused_kernel<13, false, true, EB>;
        case 28: return fused_kernel<14, false, true, EB>;
        case 30: return fused_kernel<15, false, true, EB>;
        case 32: return fused_kernel<16, false, true, EB>;
            default: return nullptr;
        }
    }
    switch (n * 2 + (inv ? 1 : 0)) {
        case 20: return fused_kernel<10, false, false, EB>;
        case 21: return fused_kernel<10, true, false, EB>;
        case 22: return fused_kernel<11, false, false, EB>;
        case 23: return fused_kernel<11, true, false, EB>;
        case 24: return fused_kernel<12, false, false, EB>;
        case 25: return fused_kernel<12, true, false, EB>;
        case 26: return fused_kernel<13, false, false, EB>;
        case 27: return fused_kernel<13, true, false, EB>;
        case 28: return fused_kernel<14, false, false, EB>;
        case 29: return fused_kernel<14, true, false, EB>;
        case 30: return fused_kernel<15, false, false, EB>;
        case 31: return fused_kernel<15, true, false, EB>;
        case 32: return fused_kernel<16, false, false, EB>;
        case 33: return fused_kernel<16, true, false, EB>;
        default: return nullptr;
    }
}

}  // namespace ntt
