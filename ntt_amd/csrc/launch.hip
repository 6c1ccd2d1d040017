// Kernel instantiation, launch geometry and the drop-in launchers of the
// ntt_amd C ABI.
//
// Reference launch geometry (src/v04_montgomery.cu:582-601): workgroup
// 2^(R - S + W) threads, LDS (workgroup << S) words; here plus one padding
// word per 32.  The grid is flat: limbs x operands x tiles workgroups with
// tiles = 2^(n - R - W), so limb counts are not bounded by gridDim.y.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/ntt_amd.h"
#include "host_math.hpp"
#include "launch.hpp"
#include "phase_kernels.hpp"
#include "trace.hpp"

namespace ntt {

using KernelFn = void (*)(PhaseParams);

template <int S, int MODE, int RF>
static KernelFn pick_r(bool inv, bool col) {
    if (!inv && col) return phase_kernel<S, false, true, MODE, RF>;
    if (!inv && !col) return phase_kernel<S, false, false, MODE, RF>;
    if (inv && col) return phase_kernel<S, true, true, MODE, RF>;
    return phase_kernel<S, true, false, MODE, RF>;
}

// Column passes with the whole geometry fixed at compile time: logN 17
// (where the two-pass kernels compete with the split fused kernels, one
// workgroup per 2^17-word limb), R1 = 8 or 9, W = 5 or 6 within 1024 threads
// per workgroup.
template <int S, int MODE, int R>
static KernelFn pick_col17(bool inv, int W) {
    if (W == 5) return inv ? phase_kernel<S, true, true, MODE, R, 17, 5> : phase_kernel<S, false, true, MODE, R, 17, 5>;
    if constexpr (R - S + 6 <= 10)
        return inv ? phase_kernel<S, true, true, MODE, R, 17, 6> : phase_kernel<S, false, true, MODE, R, 17, 6>;
    return nullptr;
}

template <int S, int MODE>
static KernelFn pick_m(bool inv, bool col, int n, int R, int W) {
    // Compile-time radix for the row-pass splits the tuner picks at logN
    // 16-17 (measured 8-15 % faster there).  A column pass with only its
    // radix fixed measured no better; with its whole geometry fixed
    // (pick_col17) its LDS and global offsets become constants.
    if (MODE == MODE_LAZY && S >= 4 && col && n == 17 && (W == 5 || W == 6) && R - S + W <= 10) {
        if (R == 8) return pick_col17<S, MODE, 8>(inv, W);
        if (R == 9) return pick_col17<S, MODE, 9>(inv, W);
    }
    if (MODE != MODE_TWCONV && S >= 4 && !col) {
        if (R == 8) return pick_r<S, MODE, 8>(inv, col);
        if (R == 9) return pick_r<S, MODE, 9>(inv, col);
        if (R == 10) return pick_r<S, MODE, 10>(inv, col);
    }
    return pick_r<S, MODE, 0>(inv, col);
}

template <int S>
static KernelFn pick_s(bool inv, bool col, int mode, int n, int R, int W) {
    switch (mode) {
        case MODE_CANON: return pick_m<S, MODE_CANON>(inv, col, n, R, W);
        case MODE_TWCONV: return pick_m<S, MODE_TWCONV>(inv, col, n, R, W);
        default: return pick_m<S, MODE_LAZY>(inv, col, n, R, W);
    }
}

static KernelFn pick(int S, bool inv, bool col, int mode, int n, int R, int W) {
    switch (S) {
        case 1: return pick_s<1>(inv, col, mode, n, R, W);
        case 2: return pick_s<2>(inv, col, mode, n, R, W);
        case 3: return pick_s<3>(inv, col, mode, n, R, W);
        case 4: return pick_s<4>(inv, col, mode, n, R, W);
        case 5: return pick_s<5>(inv, col, mode, n, R, W);
        default: return nullptr;
    }
}

static thread_local char g_last_error[256];

int record_hip_error(hipError_t e, const char* where) {
    snprintf(g_last_error, sizeof g_last_error, "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
    return NTT_ERR_HIP;
}

static std::once_flag g_lds_once;
static size_t g_max_lds = 64 * 1024;

size_t device_max_lds() {
    std::call_once(g_lds_once, [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            g_max_lds = (size_t)v;
        if (g_max_lds > NTT_LDS_BYTES_GFX950) g_max_lds = NTT_LDS_BYTES_GFX950;
    });
    return g_max_lds;
}

static int launch_pass(const ntt_phase_config& pc, bool inv, bool col, int mode, const PhaseParams& base,
                       hipStream_t stream) {
    if (pc.radix_stages == 0) return NTT_OK;
    const int R = pc.radix_stages, S = pc.stage_merging, W = pc.warp_batching;
    PhaseParams p = base;
    p.R = R;
    p.W = W;
    p.tiles = 1u << (p.logN - R - W);
    p.tiles_log2 = (uint32_t)(p.logN - R - W);
    const size_t blocks = (size_t)p.tiles * (p.limbs - p.skip_len) * p.count;
    if (blocks == 0 || blocks > 0x7fffffffu) return NTT_ERR_INVALID_ARGUMENT;
    div_magic(p.count, &p.count_magic, &p.count_s1, &p.count_s2);
    p.xcd_remap = (base.xcd_remap && blocks % 8 == 0) ? 1 : 0;
    const unsigned threads = 1u << (R - S + W);
    const size_t lds = lds_bytes_for((size_t)1 << (R + W));
    KernelFn fn = pick(S, inv, col, mode, p.logN, R, W);
    if (!fn) return NTT_ERR_CONFIG;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return record_hip_error(e, "hipFuncSetAttribute");
    }
    hipError_t stale = hipGetLastError();   // clear an error left by an unrelated earlier call
    if (stale != hipSuccess) record_hip_error(stale, "error pending before launch (cleared)");
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(threads), lds, stream, p);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? NTT_OK : record_hip_error(e, "phase kernel launch");
}

int launch_transform(const TransformLaunch& L, hipStream_t stream) {
    if (L.A_ptrs || L.chunk == 0 || L.chunk >= L.count) return launch_transform_passes(L, stream, 3);
    const size_t words = L.limbs << L.logN;
    const size_t src_words = (L.src_period ? (size_t)L.src_period : L.limbs) << L.logN;
    const size_t sub_words = (L.sub_range ? (size_t)(L.sub_end - L.sub_start) : L.limbs) << L.logN;
    for (size_t first = 0; first < L.count; first += L.chunk) {
        TransformLaunch c = L;
        c.A = L.A + first * words;
        if (L.src) c.src = L.src + first * src_words;
        if (L.sub) c.sub = L.sub + first * sub_words;
        c.count = L.count - first < L.chunk ? L.count - first : L.chunk;
        int rc = launch_transform_passes(c, stream, 3);
        if (rc) return rc;
    }
    return NTT_OK;
}

int launch_transform_passes(const TransformLaunch& L, hipStream_t stream, int pass_mask) {
    if ((!L.A && !L.A_ptrs) || (!L.tw && !L.tw_ptrs) || !L.moduli || L.limbs == 0 || L.count == 0)
        return NTT_ERR_INVALID_ARGUMENT;
    if (L.A_ptrs && (L.count != 1 || !L.tw_ptrs)) return NTT_ERR_INVALID_ARGUMENT;
    if (L.logN < 1 || L.logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    if (L.scale_mode != SCALE_NONE && !L.scale) return NTT_ERR_INVALID_ARGUMENT;
    if ((reinterpret_cast<uintptr_t>(L.A) | reinterpret_cast<uintptr_t>(L.src) | reinterpret_cast<uintptr_t>(L.sub)) &
        15)
        return NTT_ERR_INVALID_ARGUMENT;   // 16 B vector access
    // out-of-place and mod-down launches run whole transforms on contiguous operands
    if ((L.src || L.sub) && (L.A_ptrs || pass_mask != 3)) return NTT_ERR_INVALID_ARGUMENT;
    if (L.sub && (L.inverse || !L.post)) return NTT_ERR_INVALID_ARGUMENT;
    if (L.A_ptrs && (L.skip_len || L.src_period || L.src_first || L.sub_range)) return NTT_ERR_INVALID_ARGUMENT;
    if ((size_t)L.skip_start + L.skip_len > L.limbs || L.skip_len >= L.limbs) return NTT_ERR_INVALID_ARGUMENT;
    if (L.sub_range && (L.sub_start > L.sub_end || L.sub_end > L.limbs)) return NTT_ERR_INVALID_ARGUMENT;
    int rc = validate_config(L.logN, &L.cfg, device_max_lds());
    if (rc) return rc;

    PhaseParams p{};
    p.A = L.A;
    p.A_ptrs = L.A_ptrs;
    p.tw = L.tw;
    p.tw_ptrs = L.tw_ptrs;
    p.moduli = L.moduli;
    p.q_inv = L.q_inv;
    p.logN = L.logN;
    p.limbs = (uint32_t)L.limbs;
    p.count = (uint32_t)L.count;
    p.xcd_remap = L.xcd_remap;
    p.skip_start = L.skip_start;
    p.skip_len = L.skip_len;
    p.src_period = L.src_period ? L.src_period : (uint32_t)L.limbs;
    p.src_first = L.src_first;
    p.sub_start = L.sub_range ? L.sub_start : 0;
    p.sub_end = L.sub_range ? L.sub_end : (uint32_t)L.limbs;
    p.subf = L.subf;

    // The scale multiplies on the first load of a forward transform and on
    // the last store of an inverse one, in whichever pass that is.
    PhaseParams none = p;
    none.scale = nullptr;
    none.scale_mode = SCALE_NONE;
    PhaseParams scaled = p;
    scaled.scale = L.scale;
    scaled.scale_mode = L.scale_mode;

    const int mode = L.twconv ? MODE_TWCONV : (L.lazy ? MODE_LAZY : MODE_CANON);
    if (!L.inverse) {
        const ntt_phase_config &c1 = L.cfg.ntt_phase1, &c2 = L.cfg.ntt_phase2;
        const bool first_is_col = c1.radix_stages > 0;
        PhaseParams a = first_is_col ? scaled : none, b = first_is_col ? none : scaled;
        a.final_pass = c2.radix_stages == 0;
        b.final_pass = 1;
        (first_is_col ? a : b).src = L.src;
        // the mod-down store is in the row pass, where the forward transform ends
        if (L.sub && c2.radix_stages == 0) return NTT_ERR_CONFIG;
        b.sub = L.sub;
        b.post = L.post;
        if ((pass_mask & 1) && (rc = launch_pass(c1, false, true, mode, a, stream))) return rc;
        if ((pass_mask & 2) && (rc = launch_pass(c2, false, false, mode, b, stream))) return rc;
    } else {
        const ntt_phase_config &c1 = L.cfg.intt_phase1, &c2 = L.cfg.intt_phase2;
        const bool last_is_col = c2.radix_stages > 0;
        PhaseParams a = last_is_col ? none : scaled, b = last_is_col ? scaled : none;
        a.final_pass = c2.radix_stages == 0;
        b.final_pass = 1;
        (c1.radix_stages > 0 ? a : b).src = L.src;
        if ((pass_mask & 1) && (rc = launch_pass(c1, true, false, mode, a, stream))) return rc;
        if ((pass_mask & 2) && (rc = launch_pass(c2, true, true, mode, b, stream))) return rc;
    }
    return NTT_OK;
}

static int resolve_cfg(int logN, const ntt_transform_config* cfg, ntt_transform_config* out) {
    if (logN < 1 || logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    if (cfg) *out = *cfg; else default_config(logN, out);
    return NTT_OK;
}

}  // namespace ntt

using namespace ntt;

extern "C" {

const char* ntt_last_hip_error(void) { return g_last_error; }

int ntt_launch_ntt_montgomery(uint32_t* A_dev, const uint32_t* twiddle_dev, const uint32_t* moduli_dev,
                              const uint32_t* q_inv_dev, const uint32_t* r2_dev, size_t limb_count, int logN,
                              const ntt_transform_config* cfg, void* stream) {
    TraceRange range("custom_ntt");
    TransformLaunch L{};
    int rc = resolve_cfg(logN, cfg, &L.cfg);
    if (rc) return rc;
    if (!q_inv_dev || !r2_dev) return NTT_ERR_INVALID_ARGUMENT;
    L.A = A_dev; L.tw = twiddle_dev; L.moduli = moduli_dev; L.q_inv = q_inv_dev;
    L.scale = r2_dev; L.scale_mode = SCALE_AS_IS;       // mont_mul(x, R^2) = x R (v04:229-232)
    L.logN = logN; L.limbs = limb_count; L.count = 1; L.inverse = false; L.twconv = false; L.xcd_remap = 1;
    return launch_transform(L, (hipStream_t)stream);
}

int ntt_launch_intt_montgomery(uint32_t* A_dev, const uint32_t* inv_twiddle_dev, const uint32_t* moduli_dev,
                               const uint32_t* q_inv_dev, const uint32_t* degree_inv_dev, size_t limb_count,
                               int logN, const ntt_transform_config* cfg, void* stream) {
    TraceRange range("custom_intt");
    TransformLaunch L{};
    int rc = resolve_cfg(logN, cfg, &L.cfg);
    if (rc) return rc;
    if (!q_inv_dev || !degree_inv_dev) return NTT_ERR_INVALID_ARGUMENT;
    L.A = A_dev; L.tw = inv_twiddle_dev; L.moduli = moduli_dev; L.q_inv = q_inv_dev;
    // NormalizeKernel (v04:567-580) is from_mont(mont_mul(x, to_mont(N^-1))) = mont_mul(x, N^-1).
    L.scale = degree_inv_dev; L.scale_mode = SCALE_FROM_MONT;
    L.logN = logN; L.limbs = limb_count; L.count = 1; L.inverse = true; L.twconv = false; L.xcd_remap = 1;
    return launch_transform(L, (hipStream_t)stream);
}

int ntt_launch_ntt_plain(uint32_t* A_dev, const uint32_t* twiddle_dev, const uint32_t* moduli_dev,
                         size_t limb_count, int logN, const ntt_transform_config* cfg, void* stream) {
    TraceRange range("custom_ntt");
    TransformLaunch L{};
    int rc = resolve_cfg(logN, cfg, &L.cfg);
    if (rc) return rc;
    L.A = A_dev; L.tw = twiddle_dev; L.moduli = moduli_dev; L.q_inv = nullptr;
    L.scale = nullptr; L.scale_mode = SCALE_NONE;
    L.logN = logN; L.limbs = limb_count; L.count = 1; L.inverse = false; L.twconv = true; L.xcd_remap = 1;
    return launch_transform(L, (hipStream_t)stream);
}

int ntt_launch_intt_plain(uint32_t* A_dev, const uint32_t* inv_twiddle_dev, const uint32_t* moduli_dev,
                          const uint32_t* degree_inv_dev, size_t limb_count, int logN,
                          const ntt_transform_config* cfg, void* stream) {
    TraceRange range("custom_intt");
    TransformLaunch L{};
    int rc = resolve_cfg(logN, cfg, &L.cfg);
    if (rc) return rc;
    if (!degree_inv_dev) return NTT_ERR_INVALID_ARGUMENT;
    L.A = A_dev; L.tw = inv_twiddle_dev; L.moduli = moduli_dev; L.q_inv = nullptr;
    L.scale = degree_inv_dev; L.scale_mode = SCALE_TO_MONT;   // x * N^-1 (v03:516-527)
    L.logN = logN; L.limbs = limb_count; L.count = 1; L.inverse = true; L.twconv = true; L.xcd_remap = 1;
    return launch_transform(L, (hipStream_t)stream);
}

}  // extern "C"
