// Instantiations of the fused limb kernel with on-the-fly twiddles in the
// low-bit round: window width 6 (1024 threads x 64 words), lazy reduction
// (every q < 2^30) or canonical arithmetic, both directions.  A translation
// unit of its own, so it compiles in parallel with the others.
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

namespace {
template <bool INV, bool LAZY>
FusedFn pick_ot(int n) {
    switch (n) {
        case 10: return fused_kernel<10, INV, LAZY, 6, false, true>;
        case 11: return fused_kernel<11, INV, LAZY, 6, false, true>;
        case 12: return fused_kernel<12, INV, LAZY, 6, false, true>;
        case 13: return fused_kernel<13, INV, LAZY, 6, false, true>;
        case 14: return fused_kernel<14, INV, LAZY, 6, false, true>;
        case 15: return fused_kernel<15, INV, LAZY, 6, false, true>;
        case 16: return fused_kernel<16, INV, LAZY, 6, false, true>;
        default: return nullptr;
    }
}
}  // namespace

FusedFn fused_fn_ot(int n, bool inv, bool lazy) {
    if (inv) return lazy ? pick_ot<true, true>(n) : pick_ot<true, false>(n);
    return lazy ? pick_ot<false, true>(n) : pick_ot<false, false>(n);
}

}  // namespace ntt
