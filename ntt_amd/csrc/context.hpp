// Context and plan objects behind the opaque handles of include/ntt_amd.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "../../include/ntt_amd.h"
#include "launch.hpp"

namespace ntt {

// Montgomery-form bit-reversed power tables, one per (q, logN, direction),
// resident in HBM for the life of the context.
struct TwiddleCache {
    std::mutex mu;
    std::map<std::tuple<uint32_t, int, int>, uint32_t*> tables;
    size_t bytes = 0;
    int get(uint32_t q, uint32_t root, int logN, bool inverse, const uint32_t** out);
    ~TwiddleCache();
};

}  // namespace ntt

struct ntt_context {
    int device = 0;
    ntt::TwiddleCache twiddles;
    mutable std::mutex tune_mu;
    std::map<std::pair<int, size_t>, ntt_transform_config> tuned;   // (logN, limbs) -> config
    bool tuned_config(int logN, size_t limbs, ntt_transform_config* out) const;
};

struct ntt_plan {
    ntt_context* ctx = nullptr;
    int logN = 0;
    size_t limbs = 0;
    int domain = NTT_DOMAIN_MONTGOMERY;
    ntt_transform_config cfg{};
    void* d_const = nullptr;                       // one allocation holding the arrays below
    const uint32_t* const* d_tw_fwd = nullptr;     // per-limb forward table pointers
    const uint32_t* const* d_tw_inv = nullptr;
    uint32_t* d_moduli = nullptr;
    uint32_t* d_q_inv = nullptr;
    uint32_t* d_scale_fwd = nullptr;               // Montgomery multiplier on the forward input
    uint32_t* d_scale_inv = nullptr;               // Montgomery multiplier on the inverse output
};

namespace ntt {
int plan_build(ntt_context* ctx, int logN, const ntt_modulus_config* moduli, size_t limbs, int domain,
               const ntt_transform_config* cfg, ntt_plan* P);
TransformLaunch plan_launch(const ntt_plan* P, uint32_t* A, size_t count, bool inverse);
}  // namespace ntt
