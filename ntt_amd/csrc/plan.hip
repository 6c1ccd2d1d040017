// Contexts (device-resident twiddle cache) and plans of the ntt_amd C ABI.
//
// The reference builds every table on the host and uploads it inside main()
// for one fixed modulus set (src/v04_montgomery.cu:882-916, 947-967).  Here a
// context owns one Montgomery-form table per (q, root, logN, direction), shared by
// every plan and batch that uses that modulus, and kept in HBM until the
// context is destroyed.  A plan binds a modulus list to device arrays of
// per-limb constants and table pointers.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/ntt_amd.h"
#include "context.hpp"
#include "fused_types.hpp"
#include "host_math.hpp"
#include "launch.hpp"
#include "phase_kernels.hpp"
#include "trace.hpp"

namespace ntt {

int TwiddleCache::get(uint32_t q, uint32_t psi_root, int logN, bool inverse, const uint32_t** out) {
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_tuple(q, psi_root, logN, inverse ? 1 : 0);
    auto it = tables.find(key);
    if (it != tables.end()) {
        *out = it->second;
        return NTT_OK;
    }
    const size_t N = (size_t)1 << logN, M = ot_msb_words(logN);
    std::vector<uint32_t> host(N + M);
    twiddles_bitrev(logN, q, psi_root, true, host.data());
    for (size_t h = 0; h < M; ++h) host[N + h] = host[h << OT_LSB_BITS];   // on-the-fly appendix
    uint32_t* d = nullptr;
    if (hipMalloc(&d, (N + M) * sizeof(uint32_t)) != hipSuccess) return NTT_ERR_OUT_OF_MEMORY;
    if (hipMemcpy(d, host.data(), (N + M) * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return NTT_ERR_HIP;
    }
    tables[key] = d;
    bytes += (N + M) * sizeof(uint32_t);
    *out = d;
    return NTT_OK;
}

size_t TwiddleCache::entries() const {
    std::lock_guard<std::mutex> g(mu);
    return tables.size();
}

size_t TwiddleCache::total_bytes() const {
    std::lock_guard<std::mutex> g(mu);
    return bytes;
}

TwiddleCache::~TwiddleCache() {
    for (auto& kv : tables) (void)hipFree(kv.second);
}

int plan_build(ntt_context* ctx, int logN, const ntt_modulus_config* moduli, size_t limbs, int domain,
               const ntt_transform_config* cfg, ntt_plan* P) {
    P->ctx = ctx;
    P->logN = logN;
    P->limbs = limbs;
    P->domain = domain;
    P->cfg_fixed = cfg != nullptr;
    if (cfg) P->cfg = *cfg;
    else default_config(logN, &P->cfg);
    int rc = validate_config(logN, &P->cfg, device_max_lds());
    if (rc) return rc;

    std::vector<uint32_t> q(limbs), qi(limbs), sf(limbs), si(limbs);
    std::vector<const uint32_t*> twf(limbs), twi(limbs);
    for (size_t l = 0; l < limbs; ++l) {
        ntt_modulus_runtime rt;
        if ((rc = make_runtime(logN, moduli[l].qi, moduli[l].primitive_root, &rt))) return rc;
        q[l] = rt.qi;
        qi[l] = rt.q_inv;
        // Forward input scale: the Montgomery variant multiplies by R^2 (-> x R);
        // the plain variant skips the multiply (see plan_launch).
        sf[l] = rt.r2;
        // Inverse output scale: mont_mul(v, K) with K = N^-1 for the
        // Montgomery variant (drops the R factor too) and to_mont(N^-1) for plain.
        si[l] = domain == NTT_DOMAIN_MONTGOMERY ? rt.degree_inv : to_mont_host(rt.degree_inv, rt.qi);
        if ((rc = ctx->twiddles.get(rt.qi, rt.psi, logN, false, &twf[l]))) return rc;
        if ((rc = ctx->twiddles.get(rt.qi, rt.psi_inv, logN, true, &twi[l]))) return rc;
    }
    P->lazy = true;
    for (size_t l = 0; l < limbs; ++l) P->lazy = P->lazy && q[l] < (1u << 30);
    P->h_q = q;
    P->h_q_inv = qi;
    P->h_tw_fwd = twf;
    P->h_tw_inv = twi;
    P->h_scale_fwd = sf;
    P->h_scale_inv = si;
    P->h_r2.resize(limbs);
    P->h_degree_inv.resize(limbs);
    for (size_t l = 0; l < limbs; ++l) {
        ntt_modulus_runtime rt;
        make_runtime(logN, moduli[l].qi, moduli[l].primitive_root, &rt);
        P->h_r2[l] = rt.r2;
        P->h_degree_inv[l] = rt.degree_inv;
    }
    const size_t u32b = limbs * sizeof(uint32_t), ptrb = limbs * sizeof(void*);
    const size_t lcb = limbs * sizeof(LimbConst);
    if (hipMalloc(&P->d_const, 2 * lcb + 4 * u32b + 2 * ptrb) != hipSuccess) return NTT_ERR_OUT_OF_MEMORY;
    char* base = reinterpret_cast<char*>(P->d_const);
    P->d_tw_fwd = reinterpret_cast<const uint32_t* const*>(base);
    P->d_tw_inv = reinterpret_cast<const uint32_t* const*>(base + ptrb);
    P->d_moduli = reinterpret_cast<uint32_t*>(base + 2 * ptrb);
    P->d_q_inv = P->d_moduli + limbs;
    P->d_scale_fwd = P->d_q_inv + limbs;
    P->d_scale_inv = P->d_scale_fwd + limbs;
    P->d_lc_fwd = reinterpret_cast<const LimbConst*>(base + 2 * ptrb + 4 * u32b);
    P->d_lc_inv = P->d_lc_fwd + limbs;
    std::vector<LimbConst> lcf(limbs), lci(limbs);
    for (size_t l = 0; l < limbs; ++l) {
        // plain: no input scale, so the fused kernels' multiplier is R mod q (the identity)
        lcf[l] = domain == NTT_DOMAIN_MONTGOMERY ? LimbConst{q[l], qi[l], sf[l], 1u}
                                                 : LimbConst{q[l], qi[l], to_mont_host(1u, q[l]), 0u};
        lci[l] = {q[l], qi[l], si[l], 1u};
    }
    std::vector<char> h(2 * lcb + 4 * u32b + 2 * ptrb);
    std::memcpy(h.data() + 2 * ptrb + 4 * u32b, lcf.data(), lcb);
    std::memcpy(h.data() + 2 * ptrb + 4 * u32b + lcb, lci.data(), lcb);
    std::memcpy(h.data(), twf.data(), ptrb);
    std::memcpy(h.data() + ptrb, twi.data(), ptrb);
    std::memcpy(h.data() + 2 * ptrb, q.data(), u32b);
    std::memcpy(h.data() + 2 * ptrb + u32b, qi.data(), u32b);
    std::memcpy(h.data() + 2 * ptrb + 2 * u32b, sf.data(), u32b);
    std::memcpy(h.data() + 2 * ptrb + 3 * u32b, si.data(), u32b);
    if (hipMemcpy(P->d_const, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess) return NTT_ERR_HIP;
    return NTT_OK;
}

TransformLaunch plan_launch(const ntt_plan* P, uint32_t* A, size_t count, bool inverse) {
    TransformLaunch L{};
    L.A = A;
    L.tw = nullptr;
    L.tw_ptrs = inverse ? P->d_tw_inv : P->d_tw_fwd;
    L.moduli = P->d_moduli;
    L.q_inv = P->d_q_inv;
    L.scale = inverse ? P->d_scale_inv : P->d_scale_fwd;
    // The plain forward transform needs no input multiply.
    L.scale_mode = (!inverse && P->domain == NTT_DOMAIN_PLAIN && !P->prologue) ? SCALE_NONE : SCALE_AS_IS;
    L.logN = P->logN;
    L.limbs = P->limbs;
    L.count = count;
    L.inverse = inverse;
    L.twconv = false;
    L.lazy = P->lazy;
    L.xcd_remap = 1;
    L.cfg = P->cfg;
    L.chunk = P->chunk;
    TunedEntry t;
    if (!P->cfg_fixed && P->ctx->tuned_lookup(P->logN, P->limbs * count, &t)) {
        L.cfg = t.cfg;
        L.chunk = t.chunk;
    }
    return L;
}

int plan_engine(const ntt_plan* P, size_t count, bool inverse) {
    if (P->engine != NTT_ENGINE_AUTO) return P->engine;
    TunedEntry t;
    if (!P->cfg_fixed && P->ctx->tuned_lookup(P->logN, P->limbs * count, &t)) return t.engine_for(inverse);
    if (P->cfg_fixed) return NTT_ENGINE_TWO_PASS;
    return fused_plan_supported(P->logN) ? NTT_ENGINE_FUSED : NTT_ENGINE_TWO_PASS;
}

// Tuned (or default) variant of one direction; 32K-word workgroups only
// where whole limbs fit them (logN <= 15).
static FusedVariant variant_of(const TunedEntry* t, int logN, bool inverse, bool lazy_capable) {
    FusedVariant v;
    v.eb = t ? (inverse ? t->fused_eb_inv : t->fused_eb_fwd) : (inverse ? FUSED_DEFAULT_EB_INV : FUSED_DEFAULT_EB_FWD);
    v.lazy = (t ? (inverse ? t->fused_lazy_inv : t->fused_lazy) : true) && lazy_capable;
    v.logw = t ? (inverse ? t->fused_logw_inv : t->fused_logw_fwd) : 16;
    if (logN > 15) v.logw = 16;
    // untuned: on-the-fly twiddles for the inverse (its staged twiddle words
    // made it the faster one), table twiddles for the forward (equal or
    // faster since its stores go through the in-wave transpose)
    v.ot = t ? (inverse ? t->fused_ot_inv : t->fused_ot_fwd) : inverse;
    if (v.ot) {   // the on-the-fly twiddle kernels: 1024 x 64 words per CU
        v.eb = 6;
        v.logw = 16;
    }
    if (logN == FUSED_SPLIT_LOG_N) {   // the split kernels: width 6
        v.eb = 6;
        v.logw = 16;
    }
    return v;
}

FusedVariant fused_variant_for(const ntt_context* ctx, int logN, size_t total_limbs, bool inverse,
                               bool lazy_capable) {
    TunedEntry t;
    const bool tuned = ctx->tuned_lookup(logN, total_limbs, &t);
    return variant_of(tuned ? &t : nullptr, logN, inverse, lazy_capable);
}

void plan_fused_variant(const ntt_plan* P, size_t count, bool inverse, FusedVariant* v, int* stagger) {
    TunedEntry t;
    const bool tuned = !P->cfg_fixed && P->ctx->tuned_lookup(P->logN, P->limbs * count, &t);
    *v = variant_of(tuned ? &t : nullptr, P->logN, inverse, P->lazy);
    if (P->twiddle_mode != NTT_TWIDDLE_AUTO) v->ot = P->twiddle_mode == NTT_TWIDDLE_OT;
    // an explicit shape or word count other than the on-the-fly kernels' one
    // selects table twiddles unless the caller asked for on-the-fly ones
    if ((P->fused_eb && P->fused_eb != 6) || (P->fused_logw && P->fused_logw != 16 && P->logN <= 15))
        if (P->twiddle_mode != NTT_TWIDDLE_OT) v->ot = false;
    if (P->fused_eb) v->eb = P->fused_eb;
    if (P->fused_arith >= 0) v->lazy = P->fused_arith == 1 && P->lazy;
    if (P->fused_logw) v->logw = P->logN <= 15 ? P->fused_logw : 16;
    *stagger = tuned ? t.stagger : 0;
    if (v->ot) {   // the on-the-fly twiddle kernels: 1024 x 64 words per CU
        v->eb = 6;
        v->logw = 16;
    }
    if (P->logN == FUSED_SPLIT_LOG_N) {   // the split kernels: width 6
        v->eb = 6;
        v->logw = 16;
    }
}

int log2_bucket(size_t v) {
    int b = 0;
    while ((size_t(2) << b) <= v) ++b;
    return b;
}

}  // namespace ntt

using namespace ntt;

bool ntt_context::tuned_lookup(int logN, size_t total_limbs, TunedEntry* out) const {
    std::lock_guard<std::mutex> g(tune_mu);
    const int b = log2_bucket(total_limbs);
    const TunedEntry* best = nullptr;
    int bestd = 1 << 30;
    for (auto& kv : tuned) {
        if (kv.first.first != logN) continue;
        const int d = kv.first.second > b ? kv.first.second - b : b - kv.first.second;
        if (d < bestd) { bestd = d; best = &kv.second; }
    }
    if (!best) return false;
    *out = *best;
    return true;
}

extern "C" {

int ntt_context_create(int device, ntt_context** out) {
    if (!out) return NTT_ERR_INVALID_ARGUMENT;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return NTT_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return NTT_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return NTT_ERR_HIP;
    ntt_context* c = new ntt_context();
    c->device = device;
    *out = c;
    return NTT_OK;
}

int ntt_context_destroy(ntt_context* ctx) {
    if (!ctx) return NTT_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    delete ctx;
    return NTT_OK;
}

size_t ntt_context_twiddle_bytes(const ntt_context* ctx) { return ctx ? ctx->twiddles.total_bytes() : 0; }
size_t ntt_context_twiddle_entries(const ntt_context* ctx) { return ctx ? ctx->twiddles.entries() : 0; }

int ntt_plan_create(ntt_context* ctx, int logN, const ntt_modulus_config* moduli, size_t limbs, int domain,
                    const ntt_transform_config* cfg, ntt_plan** out) {
    if (!ctx || !moduli || !out || limbs == 0 || limbs > 0xffffffu) return NTT_ERR_INVALID_ARGUMENT;
    if (logN < 1 || logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    if (domain != NTT_DOMAIN_PLAIN && domain != NTT_DOMAIN_MONTGOMERY) return NTT_ERR_INVALID_ARGUMENT;
    ntt_plan* P = new ntt_plan();
    int rc = plan_build(ctx, logN, moduli, limbs, domain, cfg, P);
    if (rc) {
        ntt_plan_destroy(P);
        return rc;
    }
    *out = P;
    return NTT_OK;
}

int ntt_plan_destroy(ntt_plan* plan) {
    if (!plan) return NTT_ERR_INVALID_ARGUMENT;
    if (plan->d_const) (void)hipFree(plan->d_const);
    if (plan->d_post) (void)hipFree(plan->d_post);
    if (plan->d_subf) (void)hipFree(plan->d_subf);
    delete plan;
    return NTT_OK;
}

int ntt_plan_get_config(const ntt_plan* plan, ntt_transform_config* out) {
    if (!plan || !out) return NTT_ERR_INVALID_ARGUMENT;
    *out = plan->cfg;
    return NTT_OK;
}

// Two device ranges of `bytes` bytes are the same range or disjoint.
static bool same_or_disjoint(const void* a, const void* b, size_t bytes) {
    const uintptr_t x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x == y || x + bytes <= y || y + bytes <= x;
}

// `count` operands: in place (src null), out of place from src, or (forward,
// sub non-null) the mod-down form A = (sub - NTT(src or A)) * f.
// Two device ranges overlap.
static bool overlap(const void* a, size_t abytes, const void* b, size_t bbytes) {
    const uintptr_t x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x < y + bbytes && y < x + abytes;
}

static int plan_run(ntt_plan* plan, uint32_t* A_dev, size_t count, bool inverse, void* stream,
                    const uint32_t* src = nullptr, const uint32_t* sub = nullptr, const LimbMap& map = LimbMap()) {
    if (!plan || !A_dev || count == 0) return NTT_ERR_INVALID_ARGUMENT;
    TraceRange range(sub ? "ntt_forward_mod_down" : inverse ? "ntt_inverse" : "ntt_forward");
    const size_t N = (size_t)1 << plan->logN;
    const size_t bytes = plan->limbs * N * count * sizeof(uint32_t);
    if (src == A_dev && !map.src_period && !map.src_first) src = nullptr;
    if (map.src_period || map.src_first) {
        // a limb range of wider input operands: a separate buffer
        if (!src || overlap(src, (size_t)map.src_period * N * count * sizeof(uint32_t), A_dev, bytes))
            return NTT_ERR_INVALID_ARGUMENT;
    } else if (src && !same_or_disjoint(src, A_dev, bytes)) {
        return NTT_ERR_INVALID_ARGUMENT;
    }
    if (sub && (inverse || !plan->d_post)) return NTT_ERR_INVALID_ARGUMENT;
    // sub is read in the last pass, after A already holds intermediate values
    const size_t sub_bytes = (plan->sub_range ? plan->sub_end - plan->sub_start : plan->limbs) * N * count * 4;
    if (sub && (sub == A_dev || overlap(sub, sub_bytes, A_dev, bytes))) return NTT_ERR_INVALID_ARGUMENT;
    // the split (2^17-word) fused kernels have no mod-down form: at that size
    // it runs on the two-pass kernels
    if (plan_engine(plan, count, inverse) == NTT_ENGINE_FUSED && !(sub && plan->logN == FUSED_SPLIT_LOG_N)) {
        int stagger = 0;
        FusedVariant v;
        plan_fused_variant(plan, count, inverse, &v, &stagger);
        return plan_launch_fused(plan, A_dev, count, inverse, (hipStream_t)stream, v, stagger, src, sub, map);
    }
    TransformLaunch L = plan_launch(plan, A_dev, count, inverse);
    L.src = src;
    L.sub = sub;
    L.post = sub ? plan->d_post : nullptr;
    L.skip_start = map.skip_start;
    L.skip_len = map.skip_len;
    L.src_period = map.src_period;
    L.src_first = map.src_first;
    L.sub_range = sub && plan->sub_range;
    L.sub_start = plan->sub_start;
    L.sub_end = plan->sub_end;
    L.subf = sub ? plan->d_subf : nullptr;
    return launch_transform(L, (hipStream_t)stream);
}

int ntt_plan_forward_many(ntt_plan* plan, uint32_t* A_dev, size_t count, void* stream) {
    return plan_run(plan, A_dev, count, false, stream);
}

int ntt_plan_inverse_many(ntt_plan* plan, uint32_t* A_dev, size_t count, void* stream) {
    return plan_run(plan, A_dev, count, true, stream);
}

int ntt_plan_set_engine(ntt_plan* plan, int engine) {
    if (!plan) return NTT_ERR_INVALID_ARGUMENT;
    if (engine != NTT_ENGINE_AUTO && engine != NTT_ENGINE_TWO_PASS && engine != NTT_ENGINE_FUSED)
        return NTT_ERR_INVALID_ARGUMENT;
    if (engine == NTT_ENGINE_FUSED && !fused_plan_supported(plan->logN)) return NTT_ERR_CONFIG;
    plan->engine = engine;
    return NTT_OK;
}

int ntt_plan_set_fused_shape(ntt_plan* plan, int eb) {
    if (!plan) return NTT_ERR_INVALID_ARGUMENT;
    if (eb != 0 && eb != 6 && eb != 7) return NTT_ERR_CONFIG;
    plan->fused_eb = eb;
    return NTT_OK;
}

int ntt_plan_set_fused_words(ntt_plan* plan, int log2_words) {
    if (!plan) return NTT_ERR_INVALID_ARGUMENT;
    if (log2_words != 0 && log2_words != 15 && log2_words != 16) return NTT_ERR_CONFIG;
    plan->fused_logw = log2_words;
    return NTT_OK;
}

int ntt_plan_set_constants(ntt_plan* plan, const uint32_t* prologue, const uint32_t* epilogue) {
    if (!plan) return NTT_ERR_INVALID_ARGUMENT;
    const size_t L = plan->limbs;
    const bool mont = plan->domain == NTT_DOMAIN_MONTGOMERY;
    std::vector<uint32_t> sf(L), si(L);
    std::vector<LimbConst> lcf(L), lci(L);
    for (size_t l = 0; l < L; ++l) {
        const uint64_t q = plan->h_q[l];
        const uint64_t c = prologue ? prologue[l] % q : 1, e = epilogue ? epilogue[l] % q : 1;
        // forward: mont_mul(x, K) = x K / R must be x c R (Montgomery) or x c (plain)
        sf[l] = mont ? (uint32_t)(c * plan->h_r2[l] % q) : to_mont_host((uint32_t)c, (uint32_t)q);
        // inverse: mont_mul(v, K) must be v e N^-1 / R (Montgomery input) or v e N^-1 (plain)
        const uint32_t de = (uint32_t)((uint64_t)plan->h_degree_inv[l] * e % q);
        si[l] = mont ? de : to_mont_host(de, (uint32_t)q);
        lcf[l] = {(uint32_t)q, plan->h_q_inv[l], sf[l], (mont || prologue) ? 1u : 0u};
        lci[l] = {(uint32_t)q, plan->h_q_inv[l], si[l], 1u};
    }
    hipError_t err = hipMemcpy(plan->d_scale_fwd, sf.data(), L * 4, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(plan->d_scale_inv, si.data(), L * 4, hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(const_cast<LimbConst*>(plan->d_lc_fwd), lcf.data(), L * sizeof(LimbConst), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(const_cast<LimbConst*>(plan->d_lc_inv), lci.data(), L * sizeof(LimbConst), hipMemcpyHostToDevice);
    if (err != hipSuccess) return record_hip_error(err, "hipMemcpy(plan constants)");
    plan->prologue = prologue != nullptr;
    plan->h_scale_fwd = sf;
    plan->h_scale_inv = si;
    return NTT_OK;
}

int ntt_plan_set_twiddle_mode(ntt_plan* plan, int mode) {
    if (!plan || (mode != NTT_TWIDDLE_TABLE && mode != NTT_TWIDDLE_OT && mode != NTT_TWIDDLE_AUTO))
        return NTT_ERR_INVALID_ARGUMENT;
    plan->twiddle_mode = mode;
    return NTT_OK;
}

int ntt_plan_set_fused_arith(ntt_plan* plan, int lazy) {
    if (!plan || lazy < -1 || lazy > 1) return NTT_ERR_INVALID_ARGUMENT;
    plan->fused_arith = lazy;
    return NTT_OK;
}

int ntt_plan_get_fused_variant(const ntt_plan* plan, size_t count, int direction, int* eb, int* lazy,
                               int* log2_words, int* on_the_fly) {
    if (!plan || (direction != NTT_FORWARD && direction != NTT_INVERSE)) return NTT_ERR_INVALID_ARGUMENT;
    int st = 0;
    FusedVariant v;
    plan_fused_variant(plan, count, direction == NTT_INVERSE, &v, &st);
    if (eb) *eb = v.eb;
    if (lazy) *lazy = v.lazy ? 1 : 0;
    if (log2_words) *log2_words = v.logw;
    if (on_the_fly) *on_the_fly = v.ot ? 1 : 0;
    return NTT_OK;
}

int ntt_plan_get_engine(const ntt_plan* plan, size_t count) {
    if (!plan || count == 0) return -NTT_ERR_INVALID_ARGUMENT;
    return plan_engine(plan, count);
}

int ntt_plan_get_engine_for(const ntt_plan* plan, size_t count, int direction) {
    if (!plan || count == 0 || (direction != NTT_FORWARD && direction != NTT_INVERSE)) return -NTT_ERR_INVALID_ARGUMENT;
    return plan_engine(plan, count, direction == NTT_INVERSE);
}

int ntt_plan_run_pass(ntt_plan* plan, uint32_t* A_dev, size_t count, int direction, int pass, void* stream) {
    if (!plan || !A_dev || count == 0 || (pass != 0 && pass != 1)) return NTT_ERR_INVALID_ARGUMENT;
    if (direction != NTT_FORWARD && direction != NTT_INVERSE) return NTT_ERR_INVALID_ARGUMENT;
    return launch_transform_passes(plan_launch(plan, A_dev, count, direction == NTT_INVERSE), (hipStream_t)stream,
                                   1 << pass);
}

int ntt_plan_forward_to(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, size_t count, void* stream) {
    if (!src_dev) return NTT_ERR_INVALID_ARGUMENT;
    return plan_run(plan, dst_dev, count, false, stream, src_dev);
}

int ntt_plan_inverse_to(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, size_t count, void* stream) {
    if (!src_dev) return NTT_ERR_INVALID_ARGUMENT;
    return plan_run(plan, dst_dev, count, true, stream, src_dev);
}

int ntt_plan_set_mod_down_factors(ntt_plan* plan, const uint32_t* factors) {
    if (!plan) return NTT_ERR_INVALID_ARGUMENT;
    if (!factors) {
        if (plan->d_post) (void)hipFree(plan->d_post);
        plan->d_post = nullptr;
        return NTT_OK;
    }
    const size_t L = plan->limbs;
    std::vector<uint32_t> post(L);
    for (size_t l = 0; l < L; ++l) {
        // mont_mul(v, to_mont(f)) = v f mod q in either domain of v
        const uint32_t q = plan->h_q[l];
        post[l] = to_mont_host(factors[l] % q, q);
    }
    if (!plan->d_post) {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&plan->d_post), L * sizeof(uint32_t));
        if (e != hipSuccess) {
            plan->d_post = nullptr;
            return record_hip_error(e, "hipMalloc(mod-down factors)");
        }
    }
    hipError_t e = hipMemcpy(plan->d_post, post.data(), L * sizeof(uint32_t), hipMemcpyHostToDevice);
    return e == hipSuccess ? NTT_OK : record_hip_error(e, "hipMemcpy(mod-down factors)");
}

int ntt_plan_forward_mod_down(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, const uint32_t* sub_dev,
                              size_t count, void* stream) {
    if (!src_dev || !sub_dev) return NTT_ERR_INVALID_ARGUMENT;
    return plan_run(plan, dst_dev, count, false, stream, src_dev, sub_dev);
}

int ntt_plan_set_mod_down_sub(ntt_plan* plan, uint32_t sub_start, uint32_t sub_end, const uint32_t* sub_factors) {
    if (!plan || sub_start > sub_end || sub_end > plan->limbs) return NTT_ERR_INVALID_ARGUMENT;
    const uint32_t w = sub_end - sub_start;
    if (plan->d_subf) (void)hipFree(plan->d_subf);
    plan->d_subf = nullptr;
    plan->sub_range = !(sub_start == 0 && sub_end == plan->limbs);
    plan->sub_start = sub_start;
    plan->sub_end = sub_end;
    if (!sub_factors || w == 0) return NTT_OK;
    std::vector<uint32_t> g(w);
    for (uint32_t k = 0; k < w; ++k) {
        // mont_mul(s, to_mont(g)) = s g mod q
        const uint32_t q = plan->h_q[sub_start + k];
        g[k] = to_mont_host(sub_factors[k] % q, q);
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&plan->d_subf), w * sizeof(uint32_t));
    if (e != hipSuccess) {
        plan->d_subf = nullptr;
        return record_hip_error(e, "hipMalloc(mod-down sub factors)");
    }
    e = hipMemcpy(plan->d_subf, g.data(), w * sizeof(uint32_t), hipMemcpyHostToDevice);
    return e == hipSuccess ? NTT_OK : record_hip_error(e, "hipMemcpy(mod-down sub factors)");
}

int ntt_plan_forward_skip(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, uint32_t skip_start,
                          uint32_t skip_end, size_t count, void* stream) {
    if (!plan || skip_start > skip_end || skip_end > plan->limbs || skip_end - skip_start >= plan->limbs)
        return NTT_ERR_INVALID_ARGUMENT;
    LimbMap m;
    m.skip_start = skip_start;
    m.skip_len = skip_end - skip_start;
    return plan_run(plan, dst_dev, count, false, stream, src_dev, nullptr, m);
}

int ntt_plan_inverse_from(ntt_plan* plan, uint32_t* dst_dev, const uint32_t* src_dev, uint32_t src_limbs,
                          uint32_t src_first, size_t count, void* stream) {
    if (!plan || !src_dev || src_limbs == 0 || (size_t)src_first + plan->limbs > src_limbs)
        return NTT_ERR_INVALID_ARGUMENT;
    LimbMap m;
    m.src_period = src_limbs;
    m.src_first = src_first;
    return plan_run(plan, dst_dev, count, true, stream, src_dev, nullptr, m);
}

int ntt_plan_forward(ntt_plan* plan, uint32_t* A_dev, void* stream) {
    return ntt_plan_forward_many(plan, A_dev, 1, stream);
}

int ntt_plan_inverse(ntt_plan* plan, uint32_t* A_dev, void* stream) {
    return ntt_plan_inverse_many(plan, A_dev, 1, stream);
}

}  // extern "C"
