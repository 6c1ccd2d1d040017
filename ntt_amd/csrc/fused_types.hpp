// Work descriptors of the fused limb kernel (fused_kernel.hpp), shared with
// the host code that builds them.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace ntt {

// LDS of the fused kernel: half of the workgroup's 2^logw words plus one
// padding word per 32.
constexpr uint32_t FUSED_LDS_WORDS = 32768 + 1024;   // logw = 16
constexpr uint32_t fused_lds_words(int logw) { return (1u << (logw - 1)) + (1u << (logw - 6)); }
// After the exchange buffer of a 2^16-word workgroup, the on-the-fly
// twiddle kernels for N = 2^16 stage the twiddle words their later rounds
// read: tw[0 .. 2047] and the msb appendix msb[0 .. 1023] (12 KiB; the CU
// has 160 KiB of LDS, the exchange buffer takes 132 KiB).
constexpr uint32_t FUSED_TWL_TW = 2048, FUSED_TWL_WORDS = 2048 + 1024;
// Dynamic LDS bytes of a fused launch with 2^logw words per workgroup.
constexpr uint32_t fused_lds_bytes(int logw) {
    return (fused_lds_words(logw) + (logw == 16 ? FUSED_TWL_WORDS : 0u)) * 4u;
}
// Default window widths (7: 512 threads x 128 words per CU; 6: 1024 threads
// x 64 words) per direction, the measured faster ones at N = 2^16 on MI355X
// (with lazy arithmetic in both directions where every q < 2^30); the
// autotuner overrides them per launch size, together with the arithmetic and
// the words per workgroup.
constexpr int FUSED_DEFAULT_EB_FWD = 6;
constexpr int FUSED_DEFAULT_EB_INV = 6;

// The fused kernel variant of one launch direction.
struct FusedVariant {
    int eb = 6;          // window width: 6 or 7
    bool lazy = false;   // lazy reduction (every q < 2^30)
    int logw = 16;       // words per workgroup (log2): 16, or 15 (logN <= 15)
    bool ot = false;     // on-the-fly twiddles in the low-bit round (eb 6, logw 16; tables with the msb appendix)
};

// On-the-fly twiddles: a cached table of N = 2^n words (n >= 10) is followed
// by N / 64 words msb[h] = tw[64 h], so tw[i] = mont_mul(tw[i % 64], msb[i / 64])
// (the exponent brv(i) splits over the disjoint bit fields of i).
constexpr int OT_LSB_BITS = 6;
constexpr int OT_MIN_LOG_N = 10;
inline size_t ot_msb_words(int logN) { return logN >= OT_MIN_LOG_N ? ((size_t)1 << logN) >> OT_LSB_BITS : 0; }

struct FusedParams;
using FusedFn = void (*)(FusedParams);

struct LimbConst {
    uint32_t q;
    uint32_t q_inv;
    uint32_t scale;    // Montgomery multiplier: forward input / inverse output
    // bit 0: apply scale (forward; the inverse always applies it).  With bit 0
    // clear, scale must be R mod q (mont_mul by it is the identity): the lazy
    // forward kernels fold scale into their first stage and apply it always.
    uint32_t flags;
};

// Operands of `period` limbs of size 2^n stored back to back at `data`.  The
// job transforms `limbs` limbs: limb v of the job is limb p of operand
// o = v / (period - skip_len), where p runs over [0, period) except the
// skipped range [skip_start, skip_start + skip_len).  Limb p uses table
// tw[p] and constants lc[p], is read from src + (o * src_period + src_first
// + p) * 2^n (src == data, src_period == period, src_first == 0 in place)
// and written to data + (o * period + p) * 2^n.
// The mod-down kernels also read `sub`, which holds, per operand,
// sub_end - sub_start limbs for plan limbs [sub_start, sub_end), optionally
// multiplied by the Montgomery multipliers subf[p - sub_start]; limbs
// outside the range read 0.  They store (sub - transform) * post[p].
struct FusedJob {
    uint32_t* data;
    const uint32_t* src;
    const uint32_t* sub;
    const uint32_t* post;
    const uint32_t* const* tw;
    const LimbConst* lc;
    const uint32_t* subf;
    uint32_t limbs;
    uint32_t period;
    uint32_t item_begin;   // first work item of this job (prefix sum)
    uint16_t n;
    uint16_t inverse;      // descriptive: the kernel's direction is its template parameter
    uint32_t skip_start, skip_len;
    uint32_t src_period, src_first;
    uint32_t sub_start, sub_end;
};

// A job over `limbs` whole operands' limbs of `period` limbs each, in place,
// no skipped limbs, full mod-down range.
inline FusedJob fused_job_plain(uint32_t period) {
    FusedJob j{};
    j.period = period;
    j.src_period = period;
    j.sub_end = period;
    return j;
}

struct FusedParams {
    FusedJob job0;            // used when jobs == nullptr
    const FusedJob* jobs;     // device array, item_begin ascending
    uint32_t njobs;
    uint32_t total_items;
    // Dynamic scheduling: workgroups claim items from counter[0]; the last
    // workgroup to leave resets counter[0] and counter[1] to zero.  Two words
    // of device memory, zero before the first launch, one pair per launch in
    // flight.  Null: static round-robin (item b + k * grid), which assumes the
    // whole grid resident at once.
    uint32_t* counter;
    // Start of workgroup b is delayed by (b % 4) * stagger sleeps of 64*127
    // cycles each, so that the load/store phases of different CUs interleave
    // with the compute phases of others instead of all hitting HBM at once.
    uint32_t stagger;
};
// The split kernels: limbs of 2^17 words, one per workgroup, transformed as
// two 2^16-word halves in turn.
constexpr int FUSED_SPLIT_LOG_N = 17;
// The split kernels with on-the-fly twiddles stage the table words their
// later rounds read, tw[0 .. 2047] (inverse) or tw[0 .. 4095] (forward), and
// the 2048-word msb appendix of the 2^17-word table after the exchange buffer
// (16 or 24 KiB; 148 or 156 KiB of the CU's 160 KiB in all).
constexpr uint32_t FUSED_SPLIT_MSB_WORDS = 2048;
constexpr uint32_t fused_split_twl_tw(bool inv) { return inv ? 2048u : 4096u; }
constexpr uint32_t fused_split_lds_bytes(bool inv, bool ot) {
    return (FUSED_LDS_WORDS + (ot ? fused_split_twl_tw(inv) + FUSED_SPLIT_MSB_WORDS : FUSED_TWL_WORDS)) * 4u;
}

}  // namespace ntt
