// Instantiations of the fused limb kernel for limbs of 2^17 words, each
// processed as two 2^16-word halves by two workgroups (fused_kernel.hpp,
// SPLIT): window width 6, canonical or lazy arithmetic, table twiddles or
// on-the-fly twiddles staged in LDS.
#include <hip/hip_runtime.h>

#include "fused_kernel.hpp"

namespace ntt {

FusedFn fused_fn_split(bool inv, bool lazy, bool ot) {
    if (inv && ot) return lazy ? fused_split_kernel<true, true, true> : fused_split_kernel<true, false, true>;
    if (inv) return lazy ? fused_split_kernel<true, true> : fused_split_kernel<true, false>;
    if (ot) return lazy ? fused_split_kernel<false, true, true> : fused_split_kernel<false, false, true>;
    return lazy ? fused_split_kernel<false, true> : fused_split_kernel<false, false>;
}

}  // namespace ntt
