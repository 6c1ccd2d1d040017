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
bool fused_plan_supported(int logN) { return fused_supported(logN) || logN == FUSED_SPLIT_LOG_N; }

uint32_t fused_items(int logN, size_t limbs, int logw) {
    if (logN == FUSED_SPLIT_LOG_N) return (uint32_t)limbs;   // one limb per workgroup
    const size_t lpw = (size_t)1 << (logw - logN);
    return (uint32_t)((limbs + lpw - 1) / lpw);
}


FusedFn fused_fn_eb6(int n, bool inv, bool lazy);   // fused_eb6.hip
FusedFn fused_fn_eb7(int n, bool inv, bool lazy);   // fused_eb7.hip
FusedFn fused_fn_md(int n, bool lazy);               // fused_md.hip
FusedFn fused_fn_ot(int n, bool inv, bool lazy);     // fused_ot.hip
FusedFn fused_fn_w15(int n, bool inv, int eb, bool lazy);   // fused_w15.hip
FusedFn fused_fn_split(bool inv, bool lazy, bool ot);        // fused_split.hip

// Window width of the on-the-fly twiddle kernels.
constexpr int OT_EB = 6;

int fused_md_eb(int n, bool lazy);                  // fused_md.hip: window width of fused_fn_md(n, lazy)

static FusedFn fused_fn(int n, bool inv, int eb, bool lazy, int logw = 16) {
    if (logw == 15) return fused_fn_w15(n, inv, eb, lazy);
    if (logw != 16) return nullptr;
    return eb == 6 ? fused_fn_eb6(n, inv, lazy) : fused_fn_eb7(n, inv, lazy);
}

// Raises the dynamic-LDS limit of every fused kernel once per process (before
// any stream capture that launches them).
int fused_prepare() {
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [] {
        const int lds = (int)fused_lds_bytes(16);
        for (int m = FUSED_MIN_LOG_N; m <= FUSED_MAX_LOG_N && attr_err == hipSuccess; ++m)
            for (int d = 0; d < 8 && attr_err == hipSuccess; ++d) {
                FusedFn f = fused_fn(m, d & 1, 6 + ((d >> 1) & 1), d >= 4);
                if (f)
                    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(f),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            }
        for (int m = FUSED_MIN_LOG_N; m <= FUSED_MAX_LOG_N && attr_err == hipSuccess; ++m)
            for (int d = 0; d < 4 && attr_err == hipSuccess; ++d) {
                FusedFn f = fused_fn_ot(m, d & 1, d >= 2);
                if (f)
                    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(f),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            }
        const int lds15 = (int)fused_lds_bytes(15);
        for (int m = FUSED_MIN_LOG_N; m <= 15 && attr_err == hipSuccess; ++m)
            for (int d = 0; d < 8 && attr_err == hipSuccess; ++d) {
                FusedFn f = fused_fn(m, d & 1, 6 + ((d >> 1) & 1), d >= 4, 15);
                if (f)
                    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(f),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, lds15);
            }
        for (int d = 0; d < 8 && attr_err == hipSuccess; ++d)
            attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(fused_fn_split(d & 1, d & 2, d & 4)),
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)fused_split_lds_bytes(d & 1, d & 4));
        for (int m = FUSED_MIN_LOG_N; m <= FUSED_MAX_LOG_N && attr_err == hipSuccess; ++m)
            for (int lz = 0; lz < 2 && attr_err == hipSuccess; ++lz) {
                FusedFn f = fused_fn_md(m, lz == 1);
                if (f)
                    attr_err = hipFuncSetAttribute(reinterpret_cast<const void*>(f),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            }
    });
    return attr_err == hipSuccess ? NTT_OK : record_hip_error(attr_err, "hipFuncSetAttribute(fused)");
}

// One launch of fused kernel `fn` (window width eb, 2^logw words per
// workgroup, 2^(16-logw) workgroups per CU) over the items of p.
static int launch_fused_fn(FusedFn fn, int eb, const FusedParams& p, hipStream_t stream, unsigned grid_limit,
                           int logw = 16, size_t lds_bytes = 0) {
    int rc = fused_prepare();
    if (rc) return rc;
    const size_t lds = lds_bytes ? lds_bytes : fused_lds_bytes(logw);
    unsigned grid = (unsigned)device_cus() << (16 - logw);
    if (grid_limit && grid_limit < grid) grid = grid_limit;
    if (p.total_items < grid) grid = p.total_items;
    hipError_t stale = hipGetLastError();
    if (stale != hipSuccess) record_hip_error(stale, "error pending before launch (cleared)");
    hipLaunchKernelGGL(fn, dim3(grid), dim3(1u << (logw - eb)), lds, stream, p);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? NTT_OK : record_hip_error(e, "fused kernel launch");
}

// Launches the persistent kernel of size 2^n / direction `inv`; every job in
// p has that size and direction.
int launch_fused(int n, bool inv, int eb, bool lazy, const FusedParams& p, hipStream_t stream,
                 unsigned grid_limit, int logw, bool ot) {
    if (p.total_items == 0) return NTT_OK;
    if (eb != 6 && eb != 7) return NTT_ERR_INVALID_ARGUMENT;
    if (n == FUSED_SPLIT_LOG_N) {
        // split limbs: width 6, one limb per workgroup
        return launch_fused_fn(fused_fn_split(inv, lazy, ot), 6, p, stream, grid_limit, 16,
                               fused_split_lds_bytes(inv, ot));
    }
    if (ot) {
        // on-the-fly twiddles: tables with the msb appendix, 1024 x 64 words per CU
        FusedFn f = fused_fn_ot(n, inv, lazy);
        if (!f) return NTT_ERR_CONFIG;
        return launch_fused_fn(f, OT_EB, p, stream, grid_limit);
    }
    FusedFn fn = fused_fn(n, inv, eb, lazy, logw);
    if (!fn) return NTT_ERR_CONFIG;
    return launch_fused_fn(fn, eb, p, stream, grid_limit, logw);
}

// `count` operands of a plan as one job (limb l uses table l % limbs), read
// from `src` (null: in place) and written to A; with `sub` non-null, the
// forward mod-down form (the plan's per-limb multipliers).
int plan_launch_fused(const ntt_plan* P, uint32_t* A, size_t count, bool inverse, hipStream_t stream,
                      const FusedVariant& v, int stagger, const uint32_t* src, const uint32_t* sub,
                      const LimbMap& map) {
    if (!fused_plan_supported(P->logN)) return NTT_ERR_CONFIG;
    if ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(sub)) & 15)
        return NTT_ERR_INVALID_ARGUMENT;
    if (sub && (inverse || !P->d_post)) return NTT_ERR_INVALID_ARGUMENT;
    const size_t limbs = (P->limbs - map.skip_len) * count;
    if (limbs > 0xffffffffu) return NTT_ERR_INVALID_ARGUMENT;
    FusedParams p{};
    p.job0 = fused_job_plain((uint32_t)P->limbs);
    p.job0.skip_start = map.skip_start;
    p.job0.skip_len = map.skip_len;
    if (map.src_period) p.job0.src_period = map.src_period;
    p.job0.src_first = map.src_first;
    if (sub && P->sub_range) {
        p.job0.sub_start = P->sub_start;
        p.job0.sub_end = P->sub_end;
    }
    p.job0.subf = sub ? P->d_subf : nullptr;
    p.job0.data = A;
    p.job0.src = src ? src : A;
    p.job0.sub = sub;
    p.job0.post = sub ? P->d_post : nullptr;
    p.job0.tw = inverse ? P->d_tw_inv : P->d_tw_fwd;
    p.job0.lc = inverse ? P->d_lc_inv : P->d_lc_fwd;
    p.job0.limbs = (uint32_t)limbs;
    p.job0.item_begin = 0;
    p.job0.n = (uint16_t)P->logN;
    p.job0.inverse = inverse ? 1 : 0;
    p.jobs = nullptr;
    p.njobs = 0;
    const bool w15 = v.logw == 15 && !v.ot && !sub && P->logN <= 15;
    p.total_items = fused_items(P->logN, limbs, w15 ? 15 : 16);
    p.counter = nullptr;
    p.stagger = stagger < 0 ? 0u : (uint32_t)stagger;
    if (P->logN == FUSED_SPLIT_LOG_N) {
        // split limbs: width 6; no mod-down form
        if (sub) return NTT_ERR_CONFIG;
        if (p.total_items == 0) return NTT_OK;
        // An item is a whole limb, so a launch of L limbs runs ceil(L / CUs)
        // rounds.  When the whole operands that fit one round fewer leave a
        // small tail (at most half a round of limbs), the tail runs on the
        // two-pass kernels after them instead of taking a round of its own
        // (17 x 48 x 43: 2016 limbs in 8 rounds, then one operand of 48 limbs).
        size_t head = count;
        if (!map.skip_len && !map.src_period && !map.src_first) {
            const size_t cus = (size_t)device_cus(), lp = P->limbs;
            const size_t rounds = (limbs + cus - 1) / cus;
            const size_t k = rounds > 1 ? ((rounds - 1) * cus) / lp : 0;
            if (k > 0 && k < count && (count - k) * lp <= cus / 2) head = k;
        }
        if (head != count) {   // only without a limb map: every operand has P->limbs limbs
            p.job0.limbs = (uint32_t)(head * P->limbs);
            p.total_items = fused_items(P->logN, p.job0.limbs, 16);
        }
        int rc = launch_fused(P->logN, inverse, 6, v.lazy && P->lazy, p, stream, 0, 16, v.ot);
        if (rc || head == count) return rc;
        const size_t off = (head * P->limbs) << P->logN;
        TransformLaunch L = plan_launch(P, A + off, count - head, inverse);
        L.src = src ? src + off : nullptr;
        return launch_transform(L, stream);
    }
    if (v.ot && !sub) {
        // on-the-fly twiddles: 1024 x 64 words, lazy when every q < 2^30 (and
        // the variant asks for it)
        FusedFn fn = fused_fn_ot(P->logN, inverse, v.lazy && P->lazy);
        if (!fn) return NTT_ERR_CONFIG;
        if (p.total_items == 0) return NTT_OK;
        return launch_fused_fn(fn, OT_EB, p, stream, 0);
    }
    if (sub) {
        const bool lz = P->lazy;   // the mod-down shapes: lazy when every q < 2^30
        FusedFn fn = fused_fn_md(P->logN, lz);
        if (!fn) return NTT_ERR_CONFIG;
        if (p.total_items == 0) return NTT_OK;
        return launch_fused_fn(fn, fused_md_eb(P->logN, lz), p, stream, 0);
    }
    // lazy reduction needs every q < 2^30
    return launch_fused(P->logN, inverse, v.eb, v.lazy && P->lazy, p, stream, 0, w15 ? 15 : 16, false);
}

}  // namespace ntt
