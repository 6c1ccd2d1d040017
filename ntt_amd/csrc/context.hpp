// Context and plan objects behind the opaque handles of include/ntt_amd.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/ntt_amd.h"
#include "fused_types.hpp"
#include "launch.hpp"

namespace ntt {

// Montgomery-form bit-reversed power tables, one per (q, root, logN,
// direction), resident in HBM for the life of the context.  The table is a
// function of the 2N-th root psi (forward) or psi^-1 (inverse), not of q
// alone: two primitive roots g of one q give different psi = g^((q-1)/2N)
// (src/v04_montgomery.cu:681-727), so the root is part of the key.
struct TwiddleCache {
    mutable std::mutex mu;
    std::map<std::tuple<uint32_t, uint32_t, int, int>, uint32_t*> tables;
    size_t bytes = 0;
    int get(uint32_t q, uint32_t root, int logN, bool inverse, const uint32_t** out);
    size_t entries() const;
    size_t total_bytes() const;
    ~TwiddleCache();
};

struct LimbConst;

// Which limbs of `count` contiguous operands a plan launch transforms and
// where it reads them: limbs [skip_start, skip_start + skip_len) of every
// operand are skipped; the input operands have src_period limbs (0: the
// plan's) and the plan's limb p is read from their limb src_first + p.
struct LimbMap {
    uint32_t skip_start = 0, skip_len = 0;
    uint32_t src_period = 0, src_first = 0;
};

// Autotuner result for one (logN, words per launch) bucket.
struct TunedEntry {
    ntt_transform_config cfg;
    int engine = NTT_ENGINE_TWO_PASS;   // NTT_ENGINE_TWO_PASS or NTT_ENGINE_FUSED (forward)
    int engine_inv = NTT_ENGINE_TWO_PASS;   // the inverse's
    int engine_for(bool inverse) const { return inverse ? engine_inv : engine; }
    int fused_eb_fwd = FUSED_DEFAULT_EB_FWD, fused_eb_inv = FUSED_DEFAULT_EB_INV;   // fused window width (6 or 7)
    bool fused_lazy = true;             // fused forward arithmetic: lazy (when every q < 2^30)
    bool fused_lazy_inv = true;         // fused inverse arithmetic: lazy (when every q < 2^30)
    int fused_logw_fwd = 16, fused_logw_inv = 16;   // fused words per workgroup (log2) per direction
    bool fused_ot_fwd = false, fused_ot_inv = true;  // fused on-the-fly twiddles per direction
    int stagger = 0;                    // fused engine: start stagger (FusedParams::stagger)
    size_t chunk = 0;          // operands per chunk, 0 = whole launch
    double fwd_ms = 0, inv_ms = 0;
};

}  // namespace ntt

struct ntt_context {
    int device = 0;
    ntt::TwiddleCache twiddles;
    mutable std::mutex tune_mu;
    // (logN, log2 of total limbs per launch) -> tuned entry
    std::map<std::pair<int, int>, ntt::TunedEntry> tuned;
    bool tuned_lookup(int logN, size_t total_limbs, ntt::TunedEntry* out) const;
};

struct ntt_plan {
    ntt_context* ctx = nullptr;
    int logN = 0;
    size_t limbs = 0;
    int domain = NTT_DOMAIN_MONTGOMERY;
    bool cfg_fixed = false;                        // config given by the caller
    ntt_transform_config cfg{};
    int engine = NTT_ENGINE_AUTO;
    int fused_eb = 0;                              // fused workgroup shape: 6, 7, or 0 = tuned/default
    int fused_arith = -1;                          // fused forward arithmetic: 1 lazy, 0 canonical, -1 tuned/default
    int fused_logw = 0;                            // fused words per workgroup (log2): 15, 16, or 0 = tuned/default
    bool lazy = false;                             // every q < 2^30: lazy-reduction kernels
    size_t chunk = 0;
    void* d_const = nullptr;                       // one allocation holding the arrays below
    const uint32_t* const* d_tw_fwd = nullptr;     // per-limb forward table pointers
    const uint32_t* const* d_tw_inv = nullptr;
    uint32_t* d_moduli = nullptr;
    uint32_t* d_q_inv = nullptr;
    uint32_t* d_scale_fwd = nullptr;               // Montgomery multiplier on the forward input
    uint32_t* d_scale_inv = nullptr;               // Montgomery multiplier on the inverse output
    const ntt::LimbConst* d_lc_fwd = nullptr;      // per-limb constants for the fused kernel
    const ntt::LimbConst* d_lc_inv = nullptr;
    // host copies for rebuilding the scale words (ntt_plan_set_constants)
    std::vector<uint32_t> h_q, h_q_inv, h_r2, h_degree_inv;
    std::vector<const uint32_t*> h_tw_fwd, h_tw_inv;   // host copies of the table pointers
    std::vector<uint32_t> h_scale_fwd, h_scale_inv;    // host copies of the scale words
    bool prologue = false;                         // forward input multiplied by per-limb constants
    uint32_t* d_post = nullptr;                    // mod-down multipliers to_mont(f_l), own allocation
    int twiddle_mode = NTT_TWIDDLE_AUTO;           // fused engine twiddles: ntt_twiddle_mode
    // mod-down sub operand: limbs [sub_start, sub_end) only (when sub_range),
    // each multiplied by d_subf[l - sub_start] = to_mont(factor) when non-null
    bool sub_range = false;
    uint32_t sub_start = 0, sub_end = 0;
    uint32_t* d_subf = nullptr;                    // own allocation
};

namespace ntt {
int plan_build(ntt_context* ctx, int logN, const ntt_modulus_config* moduli, size_t limbs, int domain,
               const ntt_transform_config* cfg, ntt_plan* P);
// The launch description of `count` operands; unless the caller fixed the
// config, the tuned entry for this launch size (if any) supplies config and
// chunking.
TransformLaunch plan_launch(const ntt_plan* P, uint32_t* A, size_t count, bool inverse);
int log2_bucket(size_t v);
// Engine a launch of `count` operands runs on (resolves NTT_ENGINE_AUTO), per
// direction (the autotuner picks each direction's engine).
int plan_engine(const ntt_plan* P, size_t count, bool inverse = false);
int plan_launch_fused(const ntt_plan* P, uint32_t* A, size_t count, bool inverse, hipStream_t stream,
                      const FusedVariant& v, int stagger, const uint32_t* src = nullptr, const uint32_t* sub = nullptr,
                      const LimbMap& map = LimbMap());
// The fused kernel variant a run of `count` operands uses in `inverse` direction.
void plan_fused_variant(const ntt_plan* P, size_t count, bool inverse, FusedVariant* v, int* stagger);
// Fused variant for a launch of `total_limbs` limbs of size 2^logN from the
// context's tuning (or the defaults); lazy only if the caller's moduli allow.
FusedVariant fused_variant_for(const ntt_context* ctx, int logN, size_t total_limbs, bool inverse,
                               bool lazy_capable);
bool fused_supported(int logN);
// Sizes a plan can run on the fused engine: fused_supported, and 2^17-word
// limbs as two halves (the split kernels).
bool fused_plan_supported(int logN);
int device_cus();
}  // namespace ntt
