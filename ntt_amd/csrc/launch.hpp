// Internal launch description shared by the launchers, plans and batcher.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/ntt_amd.h"

namespace ntt {

struct TransformLaunch {
    uint32_t* A;               // operand(s), A[(op * limbs + l) * N + i] ...
    uint32_t* const* A_ptrs;   // ... or a device array of per-limb operand pointers (count == 1)
    const uint32_t* src;       // input, same layout as A (null: A, in place); needs A
    const uint32_t* sub;       // forward mod-down: A = (sub - NTT(src)) * post, same layout as A
    const uint32_t* post;      // per-limb Montgomery multipliers of the mod-down form
    const uint32_t* tw;        // Montgomery-form (or, with twconv, normal-domain) table per limb
    const uint32_t* const* tw_ptrs;  // device array of per-limb table pointers (overrides tw)
    const uint32_t* moduli;    // q per limb
    const uint32_t* q_inv;     // -q^-1 mod 2^32 per limb, or null
    const uint32_t* scale;     // per-limb scale source (see ScaleMode)
    int scale_mode;
    int logN;
    size_t limbs;
    size_t count;              // operands laid out back to back
    bool inverse;
    bool twconv;               // table is normal domain: convert each twiddle on load
    bool lazy;                 // every q < 2^30: lazy-reduction arithmetic
    int xcd_remap;
    // Operands per chunk (0 = all): both passes run chunk by chunk, so the
    // intermediate between the passes of one chunk can stay in the caches.
    size_t chunk;
    ntt_transform_config cfg;
    // Limb mapping (contiguous operands only; see PhaseParams): limbs
    // [skip_start, skip_start + skip_len) of each operand are not
    // transformed; src operand stride src_period limbs (0: limbs) starting
    // at limb src_first; mod-down sub range [sub_start, sub_end) when
    // sub_range (else every limb) with optional multipliers subf.
    uint32_t skip_start, skip_len;
    uint32_t src_period, src_first;
    bool sub_range;
    uint32_t sub_start, sub_end;
    const uint32_t* subf;
};

int launch_transform(const TransformLaunch& L, hipStream_t stream);
// pass_mask bit 0: the first pass, bit 1: the second.
int launch_transform_passes(const TransformLaunch& L, hipStream_t stream, int pass_mask);
// Records the HIP error text (ntt_last_hip_error) and returns NTT_ERR_HIP.
int record_hip_error(hipError_t e, const char* where);
size_t device_max_lds();

// One seeded fill (ntt_rt_fill_uniform semantics) of a batched fill launch.
struct FillJob {
    uint32_t* dst;
    const uint32_t* moduli;    // q per limb (device)
    uint64_t word_begin;       // prefix over the launch's jobs
    uint64_t seed;
    uint32_t limbs;
    int32_t logN;
};
constexpr uint64_t FILL_CHUNK = 16384;   // words per workgroup
// d_jobs: device array, word_begin ascending, starting at 0.
int launch_fill_jobs(const FillJob* d_jobs, uint32_t njobs, uint64_t total_words, hipStream_t stream);

}  // namespace ntt
