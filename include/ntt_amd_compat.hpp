// ntt_amd_compat.hpp — header-only C++ mirror of the reference's
// configuration types and launchers, forwarding to the ntt_amd C ABI.
//
// A caller of the reference includes this header instead of
// src/ntt_config.h and keeps its code: the same type names and members
// (ModulusConfig, PhaseConfig, TransformConfig, VersionConfig,
// PhaseLaunchConfig, NTTLaunchConfig, DefaultNTTLaunchConfig,
// default_moduli, make_version_config, default_ntt_config;
// src/ntt_config.h:6-107) and the same launcher signatures:
//   Montgomery  launch_ntt<LC>(A, tw, moduli, q_inv, r2, limbs)
//               launch_intt<LC>(A, inv_tw, moduli, q_inv, degree_inv, limbs)
//               (src/v04_montgomery.cu:603-658)
//   plain       launch_ntt<LC>(A, tw, moduli, limbs)
//               launch_intt<LC>(A, inv_tw, moduli, degree_inv, limbs)
//               (src/v03_template_config.cu:549-603)
//   runtime     launch_ntt(A, tw, moduli, const VersionConfig&, limbs)
//               launch_intt(A, inv_tw, moduli, degree_inv, const VersionConfig&, limbs)
//               (src/v02_vectorized_io.cu:625-674, plain domain)
// Device buffers, their layout (A[limb * N + i]) and their domains are the
// reference's.  Launches go to the legacy default stream, as the reference's
// do, and are asynchronous.  The reference returns void and reports nothing;
// here a failed launch (for example an invalid phase config) leaves its
// ntt_status in ntt_compat_last_status(), which the caller may check where it
// would check cudaGetLastError().
//
// DefaultNTTLaunchConfig<LogN> keeps the reference's radix and stage-merging
// split with warp_batching re-derived for 64-lane wavefronts: the column
// phases take 64 consecutive columns per tile (256 B rows), capped so a
// workgroup stays within 1024 threads; below logN 12 the limb is one row
// phase.  The literal CUDA values are ReferenceNTTLaunchConfig<LogN>; both
// are accepted by the launchers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "ntt_amd.h"

struct ModulusConfig {
    uint32_t qi;
    uint32_t primitive_root;
};

struct PhaseConfig {
    int radix_stages;
    int stage_merging;
    int warp_batching;
};

struct TransformConfig {
    PhaseConfig ntt_phase1;
    PhaseConfig ntt_phase2;
    PhaseConfig intt_phase1;
    PhaseConfig intt_phase2;
};

struct VersionConfig {
    int logN;
    std::vector<ModulusConfig> moduli;
    TransformConfig transform;
};

template <int RadixStages, int StageMerging, int WarpBatching>
struct PhaseLaunchConfig {
    static constexpr int radix_stages = RadixStages;
    static constexpr int stage_merging = StageMerging;
    static constexpr int warp_batching = WarpBatching;
    static constexpr PhaseConfig runtime() { return {RadixStages, StageMerging, WarpBatching}; }
};

template <int LogN, typename NttPhase1Config, typename NttPhase2Config, typename InttPhase1Config,
          typename InttPhase2Config>
struct NTTLaunchConfig {
    static constexpr int logN = LogN;
    using NttPhase1 = NttPhase1Config;
    using NttPhase2 = NttPhase2Config;
    using InttPhase1 = InttPhase1Config;
    using InttPhase2 = InttPhase2Config;
    static constexpr TransformConfig runtime_transform() {
        return {NttPhase1::runtime(), NttPhase2::runtime(), InttPhase1::runtime(), InttPhase2::runtime()};
    }
};

namespace ntt_amd_compat_detail {
// The reference's split (src/ntt_config.h:62-73).
constexpr int ref_r1(int n) { return n == 16 ? 7 : n - 9; }
constexpr int ref_s1(int n) { return n == 16 ? 4 : 3; }
constexpr int min3(int a, int b, int c) { return a < b ? (a < c ? a : c) : (b < c ? b : c); }
// Wave64 values (ntt_default_transform_config): 64 columns per column tile,
// at most 1024 threads; logN < 12 is a single row phase of min(3, logN)
// merged stages.
constexpr int r1(int n) { return n >= 12 ? ref_r1(n) : 0; }
constexpr int s1(int n) { return n >= 12 ? ref_s1(n) : 0; }
constexpr int w1(int n) { return n >= 12 ? min3(6, n - ref_r1(n), 10 - (ref_r1(n) - ref_s1(n))) : 0; }
constexpr int r2(int n) { return n >= 12 ? 9 : n; }
constexpr int s2(int n) { return n >= 12 ? 3 : (n < 3 ? n : 3); }

inline thread_local int last_status = NTT_OK;

inline ntt_transform_config to_c(const TransformConfig& t) {
    auto p = [](const PhaseConfig& c) { return ntt_phase_config{c.radix_stages, c.stage_merging, c.warp_batching}; };
    return {p(t.ntt_phase1), p(t.ntt_phase2), p(t.intt_phase1), p(t.intt_phase2)};
}
}  // namespace ntt_amd_compat_detail

template <int LogN>
struct DefaultNTTLaunchConfig
    : NTTLaunchConfig<LogN,
                      PhaseLaunchConfig<ntt_amd_compat_detail::r1(LogN), ntt_amd_compat_detail::s1(LogN),
                                        ntt_amd_compat_detail::w1(LogN)>,
                      PhaseLaunchConfig<ntt_amd_compat_detail::r2(LogN), ntt_amd_compat_detail::s2(LogN), 0>,
                      PhaseLaunchConfig<ntt_amd_compat_detail::r2(LogN), ntt_amd_compat_detail::s2(LogN), 0>,
                      PhaseLaunchConfig<ntt_amd_compat_detail::r1(LogN), ntt_amd_compat_detail::s1(LogN),
                                        ntt_amd_compat_detail::w1(LogN)>> {
    static_assert(LogN >= 10, "default launch config requires logN >= 10");
};

template <int LogN>
struct ReferenceNTTLaunchConfig
    : NTTLaunchConfig<LogN,
                      PhaseLaunchConfig<ntt_amd_compat_detail::ref_r1(LogN), ntt_amd_compat_detail::ref_s1(LogN), 4>,
                      PhaseLaunchConfig<9, 3, 0>, PhaseLaunchConfig<9, 3, 0>,
                      PhaseLaunchConfig<ntt_amd_compat_detail::ref_r1(LogN), ntt_amd_compat_detail::ref_s1(LogN), 4>> {
    static_assert(LogN >= 12, "the reference split is valid for logN >= 12");
};

inline std::vector<ModulusConfig> default_moduli() {
    ntt_modulus_config m[16];
    const int n = ntt_default_moduli(m, 16);
    std::vector<ModulusConfig> out;
    for (int i = 0; i < n; ++i) out.push_back({m[i].qi, m[i].primitive_root});
    return out;
}

template <typename LaunchConfig>
inline VersionConfig make_version_config() {
    return {LaunchConfig::logN, default_moduli(), LaunchConfig::runtime_transform()};
}

inline VersionConfig default_ntt_config() { return make_version_config<DefaultNTTLaunchConfig<16>>(); }

// Status of the calling thread's last launch through this header.
inline int ntt_compat_last_status() { return ntt_amd_compat_detail::last_status; }

// ---- Montgomery variant (src/v04_montgomery.cu:603-658) ----
template <typename LaunchConfig>
void launch_ntt(uint32_t* A_dev, const uint32_t* twiddle_dev, const uint32_t* moduli_dev,
                const uint32_t* q_inv_dev, const uint32_t* r2_dev, size_t limb_count) {
    const ntt_transform_config c = ntt_amd_compat_detail::to_c(LaunchConfig::runtime_transform());
    ntt_amd_compat_detail::last_status = ntt_launch_ntt_montgomery(
        A_dev, twiddle_dev, moduli_dev, q_inv_dev, r2_dev, limb_count, LaunchConfig::logN, &c, nullptr);
}

template <typename LaunchConfig>
void launch_intt(uint32_t* A_dev, const uint32_t* inv_twiddle_dev, const uint32_t* moduli_dev,
                 const uint32_t* q_inv_dev, const uint32_t* degree_inv_dev, size_t limb_count) {
    const ntt_transform_config c = ntt_amd_compat_detail::to_c(LaunchConfig::runtime_transform());
    ntt_amd_compat_detail::last_status = ntt_launch_intt_montgomery(
        A_dev, inv_twiddle_dev, moduli_dev, q_inv_dev, degree_inv_dev, limb_count, LaunchConfig::logN, &c, nullptr);
}

// ---- plain variant (src/v03_template_config.cu:549-603) ----
template <typename LaunchConfig>
void launch_ntt(uint32_t* A_dev, const uint32_t* twiddle_dev, const uint32_t* moduli_dev, size_t limb_count) {
    const ntt_transform_config c = ntt_amd_compat_detail::to_c(LaunchConfig::runtime_transform());
    ntt_amd_compat_detail::last_status =
        ntt_launch_ntt_plain(A_dev, twiddle_dev, moduli_dev, limb_count, LaunchConfig::logN, &c, nullptr);
}

template <typename LaunchConfig>
void launch_intt(uint32_t* A_dev, const uint32_t* inv_twiddle_dev, const uint32_t* moduli_dev,
                 const uint32_t* degree_inv_dev, size_t limb_count) {
    const ntt_transform_config c = ntt_amd_compat_detail::to_c(LaunchConfig::runtime_transform());
    ntt_amd_compat_detail::last_status = ntt_launch_intt_plain(A_dev, inv_twiddle_dev, moduli_dev, degree_inv_dev,
                                                               limb_count, LaunchConfig::logN, &c, nullptr);
}

// ---- runtime config, plain domain (src/v02_vectorized_io.cu:625-674) ----
inline void launch_ntt(uint32_t* A_dev, const uint32_t* twiddle_dev, const uint32_t* moduli_dev,
                       const VersionConfig& config, size_t limb_count) {
    const ntt_transform_config c = ntt_amd_compat_detail::to_c(config.transform);
    ntt_amd_compat_detail::last_status =
        ntt_launch_ntt_plain(A_dev, twiddle_dev, moduli_dev, limb_count, config.logN, &c, nullptr);
}

inline void launch_intt(uint32_t* A_dev, const uint32_t* inv_twiddle_dev, const uint32_t* moduli_dev,
                        const uint32_t* degree_inv_dev, const VersionConfig& config, size_t limb_count) {
    const ntt_transform_config c = ntt_amd_compat_detail::to_c(config.transform);
    ntt_amd_compat_detail::last_status = ntt_launch_intt_plain(A_dev, inv_twiddle_dev, moduli_dev, degree_inv_dev,
                                                               limb_count, config.logN, &c, nullptr);
}
