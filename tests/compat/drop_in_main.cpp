// A caller written against the reference's interface (src/ntt_config.h and
// the launchers of src/v04_montgomery.cu / src/v03_template_config.cu /
// src/v02_vectorized_io.cu), compiled unchanged against
// include/ntt_amd_compat.hpp.  Used by tests/test_compat.py.
//
//   drop_in_main info
//       prints the configs and default moduli the header exposes (no GPU).
//   drop_in_main run <variant> <in> <fwd_out> <back_out>
//       variant: mont-default | mont-reference | plain | plain-runtime;
//       reads 16 x 2^16 words (default_moduli, logN 16) from <in>, runs the
//       forward transform and then the inverse, as the reference's main()
//       does (v04:882-1062), and writes both results.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ntt_amd_compat.hpp"
#include "ntt_amd_runtime.h"

static void print_phase(const char* name, const PhaseConfig& p) {
    std::printf("%s %d %d %d\n", name, p.radix_stages, p.stage_merging, p.warp_batching);
}

static void print_transform(const char* tag, const TransformConfig& t) {
    std::printf("config %s\n", tag);
    print_phase("ntt_phase1", t.ntt_phase1);
    print_phase("ntt_phase2", t.ntt_phase2);
    print_phase("intt_phase1", t.intt_phase1);
    print_phase("intt_phase2", t.intt_phase2);
}

static int fail(const char* what, int rc) {
    std::fprintf(stderr, "%s failed: %d\n", what, rc);
    return 1;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "info") == 0) {
        VersionConfig v = default_ntt_config();
        std::printf("logN %d limbs %zu\n", v.logN, v.moduli.size());
        for (const ModulusConfig& m : v.moduli) std::printf("modulus %u %u\n", m.qi, m.primitive_root);
        print_transform("default16", v.transform);
        print_transform("default12", DefaultNTTLaunchConfig<12>::runtime_transform());
        print_transform("default17", DefaultNTTLaunchConfig<17>::runtime_transform());
        print_transform("default10", DefaultNTTLaunchConfig<10>::runtime_transform());
        print_transform("reference16", ReferenceNTTLaunchConfig<16>::runtime_transform());
        return 0;
    }
    if (argc != 6 || std::strcmp(argv[1], "run") != 0) {
        std::fprintf(stderr, "usage: %s info | run <variant> <in> <fwd_out> <back_out>\n", argv[0]);
        return 2;
    }
    const std::string variant = argv[2];
    const bool mont = variant.rfind("mont", 0) == 0;
    const int logN = 16;
    const size_t N = size_t{1} << logN;
    const std::vector<ModulusConfig> moduli = default_moduli();
    const size_t limbs = moduli.size();
    const size_t words = limbs * N, bytes = words * sizeof(uint32_t);

    // host tables exactly as the reference's main() builds them
    std::vector<uint32_t> tw(words), itw(words), q(limbs), qinv(limbs), r2(limbs), dinv(limbs);
    std::vector<ntt_modulus_config> mc(limbs);
    for (size_t l = 0; l < limbs; ++l) mc[l] = {moduli[l].qi, moduli[l].primitive_root};
    int rc = ntt_make_device_tables_host(logN, mc.data(), limbs, mont ? NTT_DOMAIN_MONTGOMERY : NTT_DOMAIN_PLAIN,
                                         tw.data(), itw.data(), q.data(), qinv.data(), r2.data(), dinv.data());
    if (rc) return fail("ntt_make_device_tables_host", rc);

    std::vector<uint32_t> a(words);
    FILE* f = std::fopen(argv[3], "rb");
    if (!f || std::fread(a.data(), 4, words, f) != words) return fail("read input", -1);
    std::fclose(f);

    uint32_t *A = nullptr, *tw_d = nullptr, *itw_d = nullptr, *q_d = nullptr, *qinv_d = nullptr, *r2_d = nullptr,
             *dinv_d = nullptr;
    if ((rc = ntt_rt_malloc((void**)&A, bytes)) || (rc = ntt_rt_malloc((void**)&tw_d, bytes)) ||
        (rc = ntt_rt_malloc((void**)&itw_d, bytes)) || (rc = ntt_rt_malloc((void**)&q_d, limbs * 4)) ||
        (rc = ntt_rt_malloc((void**)&qinv_d, limbs * 4)) || (rc = ntt_rt_malloc((void**)&r2_d, limbs * 4)) ||
        (rc = ntt_rt_malloc((void**)&dinv_d, limbs * 4)))
        return fail("ntt_rt_malloc", rc);
    ntt_rt_memcpy_htod(A, a.data(), bytes);
    ntt_rt_memcpy_htod(tw_d, tw.data(), bytes);
    ntt_rt_memcpy_htod(itw_d, itw.data(), bytes);
    ntt_rt_memcpy_htod(q_d, q.data(), limbs * 4);
    ntt_rt_memcpy_htod(qinv_d, qinv.data(), limbs * 4);
    ntt_rt_memcpy_htod(r2_d, r2.data(), limbs * 4);
    ntt_rt_memcpy_htod(dinv_d, dinv.data(), limbs * 4);

    if (variant == "mont-default") {
        using LC = DefaultNTTLaunchConfig<16>;
        launch_ntt<LC>(A, tw_d, q_d, qinv_d, r2_d, limbs);
    } else if (variant == "mont-reference") {
        using LC = ReferenceNTTLaunchConfig<16>;
        launch_ntt<LC>(A, tw_d, q_d, qinv_d, r2_d, limbs);
    } else if (variant == "plain") {
        launch_ntt<DefaultNTTLaunchConfig<16>>(A, tw_d, q_d, limbs);
    } else if (variant == "plain-runtime") {
        launch_ntt(A, tw_d, q_d, default_ntt_config(), limbs);
    } else {
        return fail("unknown variant", -1);
    }
    if ((rc = ntt_compat_last_status())) return fail("launch_ntt", rc);
    if ((rc = ntt_rt_synchronize())) return fail("synchronize", rc);
    std::vector<uint32_t> fwd(words), back(words);
    ntt_rt_memcpy_dtoh(fwd.data(), A, bytes);

    if (variant == "mont-default") launch_intt<DefaultNTTLaunchConfig<16>>(A, itw_d, q_d, qinv_d, dinv_d, limbs);
    else if (variant == "mont-reference")
        launch_intt<ReferenceNTTLaunchConfig<16>>(A, itw_d, q_d, qinv_d, dinv_d, limbs);
    else if (variant == "plain") launch_intt<DefaultNTTLaunchConfig<16>>(A, itw_d, q_d, dinv_d, limbs);
    else launch_intt(A, itw_d, q_d, dinv_d, default_ntt_config(), limbs);
    if ((rc = ntt_compat_last_status())) return fail("launch_intt", rc);
    if ((rc = ntt_rt_synchronize())) return fail("synchronize", rc);
    ntt_rt_memcpy_dtoh(back.data(), A, bytes);

    f = std::fopen(argv[4], "wb");
    if (!f || std::fwrite(fwd.data(), 4, words, f) != words) return fail("write fwd", -1);
    std::fclose(f);
    f = std::fopen(argv[5], "wb");
    if (!f || std::fwrite(back.data(), 4, words, f) != words) return fail("write back", -1);
    std::fclose(f);
    for (uint32_t* p : {A, tw_d, itw_d, q_d, qinv_d, r2_d, dinv_d}) ntt_rt_free(p);
    std::printf("ok %s\n", variant.c_str());
    return 0;
}
