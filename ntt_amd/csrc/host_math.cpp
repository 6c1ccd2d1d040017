// Host-side number theory and configuration for the ntt_amd library.
// Pure C++ (no HIP): moduli, Montgomery constants, twiddle tables, phase
// configs.  Exposed through the C ABI in include/ntt_amd.h.
//
// Reference semantics followed (file:line in the reference repository):
//   ModulusRuntime / make_modulus_runtime  src/v04_montgomery.cu:17-26, 681-718
//   mont_q_inv / to_mont / mont_r2 / from_mont  src/v04_montgomery.cu:125-146
//   make_twiddle_table                     src/v04_montgomery.cu:720-727
//   default_moduli / DefaultNTTLaunchConfig  src/ntt_config.h:62-94
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/ntt_amd.h"
#include "host_math.hpp"

namespace ntt {

uint32_t powmod(uint32_t b, uint64_t e, uint32_t m) {
    uint64_t r = 1 % m, x = b % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (uint32_t)r;
}

uint32_t mont_q_inv_host(uint32_t q) {
    uint32_t inv = 1;
    for (int i = 0; i < 5; ++i) inv *= 2u - q * inv;
    return 0u - inv;
}

uint32_t to_mont_host(uint32_t x, uint32_t q) { return (uint32_t)(((uint64_t)x << 32) % q); }

uint32_t r2_host(uint32_t q) {
    uint64_t r = (UINT64_C(1) << 32) % q;
    return (uint32_t)(r * r % q);
}

uint32_t from_mont_host(uint32_t x, uint32_t q, uint32_t q_inv) {
    uint64_t t = x;
    uint32_t m = (uint32_t)t * q_inv;
    uint64_t u = (t + (uint64_t)m * q) >> 32;
    return (uint32_t)(u >= q ? u - q : u);
}

// Deterministic Miller-Rabin for n < 2^32 (bases 2, 7, 61 are sufficient).
bool is_prime_u32(uint32_t n) {
    if (n < 2) return false;
    for (uint32_t p : {2u, 3u, 5u, 7u, 11u, 13u, 61u})
        if (n % p == 0) return n == p;
    uint32_t d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (uint32_t a : {2u, 7u, 61u}) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < s; ++i) {
            x = x * x % n;
            if (x == n - 1) { comp = false; break; }
        }
        if (comp) return false;
    }
    return true;
}

uint32_t smallest_primitive_root(uint32_t q) {
    std::vector<uint32_t> f;
    uint32_t m = q - 1;
    for (uint32_t p = 2; (uint64_t)p * p <= m; p += (p == 2 ? 1 : 2)) {
        if (m % p == 0) {
            f.push_back(p);
            while (m % p == 0) m /= p;
        }
    }
    if (m > 1) f.push_back(m);
    for (uint32_t g = 2; g < q; ++g) {
        bool ok = true;
        for (uint32_t p : f)
            if (powmod(g, (q - 1) / p, q) == 1) { ok = false; break; }
        if (ok) return g;
    }
    return 0;
}

int check_modulus(int logN, uint32_t q) {
    if (q <= 1 || !(q & 1u) || q >= (1u << 31)) return NTT_ERR_MODULUS_RANGE;
    if (((uint64_t)q - 1) % (UINT64_C(2) << logN) != 0) return NTT_ERR_MODULUS_CONGRUENCE;
    return NTT_OK;
}

int make_runtime(int logN, uint32_t q, uint32_t g, ntt_modulus_runtime* out) {
    if (logN < 1 || logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    int rc = check_modulus(logN, q);
    if (rc) return rc;
    const uint64_t N = UINT64_C(1) << logN;
    const uint32_t psi = powmod(g, ((uint64_t)q - 1) / (2 * N), q);
    if (powmod(psi, 2 * N, q) != 1 || powmod(psi, N, q) != q - 1) return NTT_ERR_PRIMITIVE_ROOT;
    out->qi = q;
    out->primitive_root = g;
    out->q_inv = mont_q_inv_host(q);
    out->r2 = r2_host(q);
    out->psi = psi;
    out->psi_inv = powmod(psi, q - 2, q);
    out->omega = (uint32_t)((uint64_t)psi * psi % q);
    out->degree_inv = powmod((uint32_t)N, q - 2, q);
    return NTT_OK;
}

// tw[i] = root^brv(i): successive powers, then a bit-reversal scatter
// (N multiplications instead of the reference's N modular exponentiations).
void twiddles_bitrev(int logN, uint32_t q, uint32_t root, bool mont, uint32_t* out) {
    const size_t N = (size_t)1 << logN;
    uint64_t p = 1;
    const uint64_t rm = mont ? ((UINT64_C(1) << 32) % q) : 1;
    for (size_t k = 0; k < N; ++k) {
        size_t r = 0, v = k;
        for (int b = 0; b < logN; ++b) { r = (r << 1) | (v & 1); v >>= 1; }
        out[r] = (uint32_t)(mont ? (p * rm % q) : p);
        p = p * root % q;
    }
}

void reference_config(int logN, ntt_transform_config* c) {
    const int r1 = logN == 16 ? 7 : logN - 9;
    const int s1 = logN == 16 ? 4 : 3;
    c->ntt_phase1 = {r1, s1, 4};
    c->ntt_phase2 = {9, 3, 0};
    c->intt_phase1 = {9, 3, 0};
    c->intt_phase2 = {r1, s1, 4};
}

// Wave64 default: the reference's radix/merging split where it is valid
// (logN >= 12), with 64 consecutive columns per column-pass tile (one wave
// per lane-row: 256 B coalesced rows, wave-uniform twiddles).  Below
// logN 12 the whole limb fits one workgroup: a single row pass.
void default_config(int logN, ntt_transform_config* c) {
    if (logN >= 12) {
        reference_config(logN, c);
        const int r = c->ntt_phase1.radix_stages, s = c->ntt_phase1.stage_merging;
        const int w = std::min({6, logN - r, 10 - (r - s)});   // <= 1024 threads
        c->ntt_phase1.warp_batching = w;
        c->intt_phase2.warp_batching = w;
    } else {
        const int s = std::min(3, logN);
        c->ntt_phase1 = {0, 0, 0};
        c->ntt_phase2 = {logN, s, 0};
        c->intt_phase1 = {logN, s, 0};
        c->intt_phase2 = {0, 0, 0};
    }
}

static int check_pass(int logN, const ntt_phase_config& p, int other_radix, size_t max_lds_bytes) {
    const int R = p.radix_stages, S = p.stage_merging, W = p.warp_batching;
    if (R == 0) return NTT_OK;   // empty pass (the other pass does all logN stages)
    if (R < 0 || S < 1 || S > NTT_MAX_STAGE_MERGING || S > R) return NTT_ERR_CONFIG;
    if (W < 0 || W > other_radix) return NTT_ERR_CONFIG;   // tiles = 2^(other - W) >= 1
    const int block_log = R - S + W;
    if (block_log > 10) return NTT_ERR_CONFIG;            // <= 1024 threads
    const size_t words = (size_t)1 << (R + W);
    if (lds_bytes_for(words) > max_lds_bytes) return NTT_ERR_CONFIG;
    (void)logN;
    return NTT_OK;
}

int validate_config(int logN, const ntt_transform_config* c, size_t max_lds_bytes) {
    if (!c) return NTT_ERR_INVALID_ARGUMENT;
    if (logN < 1 || logN > NTT_MAX_LOG_N) return NTT_ERR_CONFIG;
    const int f1 = c->ntt_phase1.radix_stages, f2 = c->ntt_phase2.radix_stages;
    const int i1 = c->intt_phase1.radix_stages, i2 = c->intt_phase2.radix_stages;
    if (f1 < 0 || f2 < 0 || i1 < 0 || i2 < 0) return NTT_ERR_CONFIG;
    if (f1 + f2 != logN || i1 + i2 != logN) return NTT_ERR_CONFIG;
    int rc;
    if ((rc = check_pass(logN, c->ntt_phase1, f2, max_lds_bytes))) return rc;
    if ((rc = check_pass(logN, c->ntt_phase2, f1, max_lds_bytes))) return rc;
    if ((rc = check_pass(logN, c->intt_phase1, i2, max_lds_bytes))) return rc;
    if ((rc = check_pass(logN, c->intt_phase2, i1, max_lds_bytes))) return rc;
    return NTT_OK;
}

}  // namespace ntt

using namespace ntt;

extern "C" {

const char* ntt_status_string(int s) {
    switch (s) {
        case NTT_OK: return "ok";
        case NTT_ERR_INVALID_ARGUMENT: return "invalid argument";
        case NTT_ERR_MODULUS_CONGRUENCE: return "qi must satisfy qi == 1 mod 2N";
        case NTT_ERR_PRIMITIVE_ROOT: return "primitive_root does not generate a primitive 2N-th root";
        case NTT_ERR_MODULUS_RANGE: return "qi must be odd and in (1, 2^31)";
        case NTT_ERR_CONFIG: return "invalid phase configuration";
        case NTT_ERR_HIP: return "HIP runtime error";
        case NTT_ERR_OUT_OF_MEMORY: return "out of device memory";
        case NTT_ERR_BUDGET: return "request exceeds the HBM budget";
        case NTT_ERR_NOT_FOUND: return "unknown request";
        case NTT_ERR_BUSY: return "busy";
        case NTT_ERR_NO_DEVICE: return "no HIP device";
        case NTT_ERR_SCRATCH_LIMIT: return "more jobs or limbs than the allocation-free launch's fixed tables hold";
        default: return "unknown status";
    }
}

const char* ntt_version(void) { return "ntt_amd 0.1 (gfx950)"; }

int ntt_max_log_n(void) { return NTT_MAX_LOG_N; }

int ntt_default_moduli(ntt_modulus_config* out, int cap) {
    static const ntt_modulus_config d[16] = {
        {974258177u, 3u},   {1081212929u, 6u}, {1196556289u, 7u}, {993263617u, 5u},
        {989986817u, 3u},   {1074266113u, 5u}, {1168900097u, 5u}, {1010565121u, 7u},
        {957349889u, 6u},   {1073872897u, 7u}, {1209139201u, 31u}, {994705409u, 3u},
        {998244353u, 3u},   {1073479681u, 11u}, {1158676481u, 3u}, {1016463361u, 38u},
    };
    if (out)
        for (int i = 0; i < cap && i < 16; ++i) out[i] = d[i];
    return 16;
}

int ntt_default_transform_config(int logN, ntt_transform_config* out) {
    if (!out || logN < 1 || logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    default_config(logN, out);
    return NTT_OK;
}

int ntt_reference_transform_config(int logN, ntt_transform_config* out) {
    if (!out || logN < 10 || logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    reference_config(logN, out);
    return NTT_OK;
}

int ntt_validate_transform_config(int logN, const ntt_transform_config* cfg) {
    return validate_config(logN, cfg, NTT_LDS_BYTES_GFX950);
}

int ntt_generate_moduli(int logN, uint32_t bound, int count, ntt_modulus_config* out) {
    if (logN < 1 || logN > 30 || count < 0 || (!out && count > 0)) return -NTT_ERR_INVALID_ARGUMENT;
    const uint64_t step = UINT64_C(2) << logN;
    if (bound > (1u << 31)) bound = 1u << 31;
    if (bound <= step + 1) return 0;
    int k = 0;
    for (uint64_t q = ((uint64_t)bound - 2) / step * step + 1; q > 2 && k < count; q -= step) {
        if (q >= bound) continue;
        if (is_prime_u32((uint32_t)q)) {
            out[k].qi = (uint32_t)q;
            out[k].primitive_root = smallest_primitive_root((uint32_t)q);
            ++k;
        }
        if (q <= step) break;
    }
    return k;
}

int ntt_make_modulus_runtime(int logN, ntt_modulus_config cfg, ntt_modulus_runtime* out) {
    if (!out) return NTT_ERR_INVALID_ARGUMENT;
    return make_runtime(logN, cfg.qi, cfg.primitive_root, out);
}

int ntt_make_twiddle_table(int logN, uint32_t q, uint32_t psi, uint32_t* tw_out) {
    if (!tw_out || logN < 0 || logN > NTT_MAX_LOG_N || q < 2) return NTT_ERR_INVALID_ARGUMENT;
    twiddles_bitrev(logN, q, psi, false, tw_out);
    return NTT_OK;
}

int ntt_make_device_tables_host(int logN, const ntt_modulus_config* moduli, size_t limbs, int domain,
                                uint32_t* twiddles, uint32_t* inv_twiddles, uint32_t* moduli_out,
                                uint32_t* q_inv, uint32_t* r2, uint32_t* degree_inv) {
    if (!moduli || limbs == 0) return NTT_ERR_INVALID_ARGUMENT;
    const size_t N = (size_t)1 << logN;
    const bool mont = domain == NTT_DOMAIN_MONTGOMERY;
    for (size_t l = 0; l < limbs; ++l) {
        ntt_modulus_runtime rt;
        int rc = make_runtime(logN, moduli[l].qi, moduli[l].primitive_root, &rt);
        if (rc) return rc;
        if (twiddles) twiddles_bitrev(logN, rt.qi, rt.psi, mont, twiddles + l * N);
        if (inv_twiddles) twiddles_bitrev(logN, rt.qi, rt.psi_inv, mont, inv_twiddles + l * N);
        if (moduli_out) moduli_out[l] = rt.qi;
        if (q_inv) q_inv[l] = rt.q_inv;
        if (r2) r2[l] = rt.r2;
        if (degree_inv) degree_inv[l] = mont ? to_mont_host(rt.degree_inv, rt.qi) : rt.degree_inv;
    }
    return NTT_OK;
}

uint32_t ntt_mont_q_inv(uint32_t q) { return mont_q_inv_host(q); }
uint32_t ntt_to_mont(uint32_t x, uint32_t q) { return q ? to_mont_host(x, q) : 0; }
uint32_t ntt_mont_r2(uint32_t q) { return q ? r2_host(q) : 0; }
uint32_t ntt_from_mont(uint32_t x, uint32_t q, uint32_t q_inv) { return from_mont_host(x, q, q_inv); }

}  // extern "C"
