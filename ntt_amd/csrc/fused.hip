// Host side of the fused limb kernel (fused_kernel.hpp): launch geometry and
// the uniform-batch entry used by plans.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/ntt_amd.h"
#include "context.hpp"
#include "fused_types.hpp"
#include "host_math.hpp"
#include "launch.hpp"

namespace ntt {

static std::once_flag g_cu_once;
static int g_cus = 256;

int device_cus() {
    std::call_once(g_cu_once, [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            g_cus = v;
    });
    return g_cus;
}

bool fused_supported(int logN) { return logN >= FUSED_MIN_LOG_N && logN <= FUSED_MAX_LOG_N; }

uint32_t fused_items(int logN, size_t limbs) {
    const size_t lpw = (size_t)1 << (16 - logN);
    return (uint32_t)((limbs + lpw - 1) / lpw);
}


FusedFn fused_fn_eb6(int n, bool inv, bool lazy);   // fused_eb6.hip
FusedFn fused_fn_eb7(int n, bool inv, bool lazy);   // fused_eb7.hip

static FusedFn fused_fn(int n, bool inv, int eb, bool lazy) {
    return eb == 6 ? fused_fn_eb6(n, inv, lazy) : fused_fn_eb7(n, inv, lazy);
}

// Raises the dynamic-LDS limit of every fused kernel once per process (before
// any stream capture that launches them).
int fused_prepare() {
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [] {
        const int lds = (int)(FUSED_LDS_WORDS * sizeof(uint32_t));
        for (int m = FUSED_MIN_LOG_N; m <= FUSED_MAX_LOG_N && attr_err == hipSuccess; ++m)
            for (int d = 0; d < 8 && attr_err == hipSuccess; ++d) {
                FusedFn f = fused_fn(m, d & 1, 6 + ((d >> 1) & 1), d >= 4);
                if (f)
                    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(f),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            }
    });
    return attr_err == hipSuccess ? NTT_OK : record_hip_error(attr_err, "hipFuncSetAttribute(fused)");
}

// Launches the persistent kernel of size 2^n / direction `inv`; every job in
// p has that size and direction.
int launch_fused(int n, bool inv, int eb, bool lazy, const FusedParams& p, hipStream_t stream,
                 unsigned grid_limit) {
    if (p.total_items == 0) return NTT_OK;
    if (eb != 6 && eb != 7) return NTT_ERR_INVALID_ARGUMENT;
    FusedFn fn = fused_fn(n, inv, eb, lazy);
    if (!fn) return NTT_ERR_CONFIG;
    int rc = fused_prepare();
    if (rc) return rc;
    const size_t lds = FUSED_LDS_WORDS * sizeof(uint32_t);
    unsigned grid = (unsigned)device_cus();
    if (grid_limit && grid_limit < grid) grid = grid_limit;
    if (p.total_items < grid) grid = p.total_items;
    hipError_t stale = hipGetLastError();
    if (stale != hipSuccess) record_hip_error(stale, "error pending before launch (cleared)");
    hipLaunchKernelGGL(fn, dim3(grid), dim3(1u << (16 - eb)), lds, stream, p);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? NTT_OK : record_hip_error(e, "fused kernel launch");
}

// `count` operands of a plan as one job (limb l uses table l % limbs).
int plan_launch_fused(const ntt_plan* P, uint32_t* A, size_t count, bool inverse, hipStream_t stream, int eb,
                      bool lazy, int stagger) {
    if (!fused_supported(P->logN)) return NTT_ERR_CONFIG;
    if (reinterpret_cast<uintptr_t>(A) & 15) return NTT_ERR_INVALID_ARGUMENT;
    const size_t limbs = P->limbs * count;
    if (limbs > 0xffffffffu) return NTT_ERR_INVALID_ARGUMENT;
    FusedParams p{};
    p.job0.data = A;
    p.job0.tw = inverse ? P->d_tw_inv : P->d_tw_fwd;
    p.job0.lc = inverse ? P->d_lc_inv : P->d_lc_fwd;
    p.job0.limbs = (uint32_t)limbs;
    p.job0.period = (uint32_t)P->limbs;
    p.job0.item_begin = 0;
    p.job0.n = (uint16_t)P->logN;
    p.job0.inverse = inverse ? 1 : 0;
    p.jobs = nullptr;
    p.njobs = 0;
    p.total_items = fused_items(P->logN, limbs);
    p.counter = nullptr;
    p.stagger = stagger < 0 ? 0u : (uint32_t)stagger;
    // lazy reduction needs every q < 2^30 and exists for the forward kernels
    return launch_fused(P->logN, inverse, eb, lazy && P->lazy && !inverse, p, stream, 0);
}

}  // namespace ntt
