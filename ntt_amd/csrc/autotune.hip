// Phase-config autotuner.
//
// The reference picks its phase split by hand-editing
// DefaultNTTLaunchConfig<LogN> (src/ntt_config.h:62-73, README.md:98-163).
// Here the split R1 + R2 = logN, the per-pass stage merging S and the warp
// batching W are measured on the device for a given launch size (limbs x
// operands), each pass timed alone with HIP events, forward and inverse
// separately; then the operands-per-chunk of a multi-operand launch.  The
// winner is stored in the context (and can be saved to / loaded from a text
// file), where plans pick it up for launches of that size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ntt_amd.h"
#include "../../include/ntt_amd_runtime.h"
#include "context.hpp"
#include "fused_types.hpp"
#include "host_math.hpp"
#include "launch.hpp"

namespace ntt {
namespace {

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer() {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~Timer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

// Median duration (ms) of `reps` runs of `fn`, after one warm-up run.
template <class F>
int time_median(hipStream_t s, int reps, F fn, double* ms_out) {
    Timer t;
    int rc = fn();
    if (rc) return rc;
    std::vector<double> v;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(t.a, s);
        if ((rc = fn())) return rc;
        (void)hipEventRecord(t.b, s);
        if (hipEventSynchronize(t.b) != hipSuccess) return NTT_ERR_HIP;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        v.push_back(ms);
    }
    std::sort(v.begin(), v.end());
    *ms_out = v[v.size() / 2];
    return NTT_OK;
}

// Steady-state time: `reps` launches back to back between two events (the
// device never idles between them), three trials, the fastest average.  Used
// for the few finalists, where single synchronised launches are too noisy to
// separate candidates a few percent apart.
template <class F>
int time_steady(hipStream_t s, int reps, F fn, double* ms_out) {
    Timer t;
    int rc = fn();
    if (rc) return rc;
    double best = 1e30;
    for (int trial = 0; trial < 3; ++trial) {
        (void)hipEventRecord(t.a, s);
        for (int r = 0; r < reps; ++r)
            if ((rc = fn())) return rc;
        (void)hipEventRecord(t.b, s);
        if (hipEventSynchronize(t.b) != hipSuccess) return NTT_ERR_HIP;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        best = std::min(best, (double)ms / reps);
    }
    *ms_out = best;
    return NTT_OK;
}

struct PassChoice {
    ntt_phase_config pc{0, 0, 0};
    double ms = 0;
};

}  // namespace

int autotune(ntt_context* ctx, int logN, const ntt_modulus_config* moduli, size_t limbs, size_t count, int domain,
             int reps, ntt_tune_result* out) {
    const size_t max_lds = device_max_lds();
    ntt_plan P;
    int rc = plan_build(ctx, logN, moduli, limbs, domain, nullptr, &P);
    if (rc) {
        if (P.d_const) (void)hipFree(P.d_const);
        return rc;
    }
    P.cfg_fixed = true;
    const size_t words = (limbs * count) << logN;
    uint32_t* buf = nullptr;
    if (hipMalloc(&buf, words * sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(P.d_const);
        return NTT_ERR_OUT_OF_MEMORY;
    }
    hipStream_t s = nullptr;
    (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    rc = ntt_rt_fill_uniform(buf, P.d_moduli, limbs, logN, count, 0x5eed, s);
    int ncand = 0;

    // Best config of one pass for a given split.  pass 0 = forward column /
    // inverse row, pass 1 = forward row / inverse column.
    auto tune_pass = [&](bool inverse, int r1, bool column, PassChoice* best) -> int {
        const int r2 = logN - r1;
        const int R = column ? r1 : r2, other = column ? r2 : r1;
        best->ms = 1e30;
        if (R == 0) {
            best->pc = {0, 0, 0};
            best->ms = 0;
            return NTT_OK;
        }
        for (int S = std::min(R, 1); S <= std::min(R, NTT_MAX_STAGE_MERGING); ++S) {
            if (S == 1 && R > 1) continue;   // one stage per round is never competitive
            const int wlist_col[] = {3, 4, 5, 6};
            const int wlist_row[] = {0, 1, 2};
            const int* wl = column ? wlist_col : wlist_row;
            const int nw = column ? 4 : 3;
            for (int k = 0; k < nw; ++k) {
                const int W = std::min(wl[k], other);
                if (k > 0 && W == std::min(wl[k - 1], other)) continue;
                ntt_transform_config c{};
                ntt_phase_config pc{R, S, W};
                ntt_phase_config filler{other, std::min(other, NTT_MAX_STAGE_MERGING), 0};
                if (other == 0) filler = {0, 0, 0};
                if (column) {
                    c.ntt_phase1 = pc; c.intt_phase2 = pc;
                    c.ntt_phase2 = filler; c.intt_phase1 = filler;
                } else {
                    c.ntt_phase2 = pc; c.intt_phase1 = pc;
                    c.ntt_phase1 = filler; c.intt_phase2 = filler;
                    if (other) { c.ntt_phase1.warp_batching = 0; c.intt_phase2.warp_batching = 0; }
                }
                if (validate_config(logN, &c, max_lds) != NTT_OK) continue;
                TransformLaunch L = plan_launch(&P, buf, count, inverse);
                L.cfg = c;
                L.chunk = 0;
                const int mask = inverse ? (column ? 2 : 1) : (column ? 1 : 2);
                double ms = 0;
                int r = time_median(s, reps, [&] { return launch_transform_passes(L, s, mask); }, &ms);
                if (r) return r;
                ++ncand;
                if (ms < best->ms) { best->ms = ms; best->pc = pc; }
            }
        }
        return NTT_OK;
    };

    ntt_transform_config best{};
    double best_fwd = 1e30, best_inv = 1e30;
    const int r1_lo = std::max(0, logN - 15), r1_hi = std::min(logN, 15);
    for (int r1 = r1_lo; rc == NTT_OK && r1 <= r1_hi; ++r1) {
        PassChoice fc, fr, ic, ir;
        if ((rc = tune_pass(false, r1, true, &fc))) break;
        if ((rc = tune_pass(false, r1, false, &fr))) break;
        if ((rc = tune_pass(true, r1, false, &ir))) break;
        if ((rc = tune_pass(true, r1, true, &ic))) break;
        if (fc.ms + fr.ms < best_fwd) {
            best_fwd = fc.ms + fr.ms;
            best.ntt_phase1 = fc.pc;
            best.ntt_phase2 = fr.pc;
        }
        if (ir.ms + ic.ms < best_inv) {
            best_inv = ir.ms + ic.ms;
            best.intt_phase1 = ir.pc;
            best.intt_phase2 = ic.pc;
        }
    }

    size_t best_chunk = 0;
    double def_fwd = 0, def_inv = 0, fused_fwd = 0, fused_inv = 0;
    int engine = NTT_ENGINE_TWO_PASS, engine_inv = NTT_ENGINE_TWO_PASS;
    FusedVariant vf, vi;   // fused variants per direction
    int stagger = 0;
    if (rc == NTT_OK) {
        // chunking of multi-operand launches (whole transforms, both passes)
        double best_pair = 1e30;
        std::vector<size_t> chunks = {0};
        const size_t op_bytes = (limbs << logN) * sizeof(uint32_t);
        for (size_t c = 1; c < count; c *= 2)
            if (c * op_bytes >= ((size_t)8 << 20)) chunks.push_back(c);
        for (size_t ch : chunks) {
            TransformLaunch Lf = plan_launch(&P, buf, count, false), Li = plan_launch(&P, buf, count, true);
            Lf.cfg = best; Li.cfg = best;
            Lf.chunk = ch; Li.chunk = ch;
            double f = 0, i = 0;
            if ((rc = time_median(s, reps, [&] { return launch_transform(Lf, s); }, &f))) break;
            if ((rc = time_median(s, reps, [&] { return launch_transform(Li, s); }, &i))) break;
            ++ncand;
            if (f + i < best_pair) {
                best_pair = f + i;
                best_chunk = ch;
                best_fwd = f;
                best_inv = i;
            }
        }
        // the fused limb kernel where it applies: each direction on its own,
        // over both window widths, both arithmetics when every q < 2^30, for
        // logN <= 15 both workgroup sizes (2^16 words, one per CU; 2^15
        // words, two per CU), and table or on-the-fly twiddles (the latter
        // at width 6, 2^16 words)
        if (rc == NTT_OK && fused_plan_supported(logN)) {
            double bf = 0, bi = 0;
            for (int logw = 16; rc == NTT_OK && logw >= (logN <= 15 ? 15 : 16); --logw)
                for (int eb = 7; rc == NTT_OK && eb >= 6; --eb)
                    for (int ot = 0; rc == NTT_OK && ot <= (eb == 6 && logw == 16 ? 1 : 0); ++ot)
                    for (int lz = P.lazy ? 1 : 0; rc == NTT_OK && lz >= 0; --lz)
                        for (int inv = 0; rc == NTT_OK && inv < 2; ++inv) {
                            // split limbs (2^17 words): width 6
                            if (logN == FUSED_SPLIT_LOG_N && eb != 6) continue;
                            FusedVariant v;
                            v.eb = eb;
                            v.lazy = lz != 0;
                            v.logw = logw;
                            v.ot = ot != 0;
                            double t = 0;
                            rc = time_steady(s, reps + 2,
                                             [&] { return plan_launch_fused(&P, buf, count, inv != 0, s, v, 0); }, &t);
                            ++ncand;
                            double& b = inv ? bi : bf;
                            if (rc == NTT_OK && (b == 0 || t < b)) {
                                b = t;
                                (inv ? vi : vf) = v;
                            }
                        }
            fused_fwd = bf;
            fused_inv = bi;
            // the two-pass finalist, timed the same way
            if (rc == NTT_OK) {
                TransformLaunch Lf = plan_launch(&P, buf, count, false), Li = plan_launch(&P, buf, count, true);
                Lf.cfg = best; Li.cfg = best;
                Lf.chunk = best_chunk; Li.chunk = best_chunk;
                double f = 0, i = 0;
                rc = time_steady(s, reps + 2, [&] { return launch_transform(Lf, s); }, &f);
                if (rc == NTT_OK) rc = time_steady(s, reps + 2, [&] { return launch_transform(Li, s); }, &i);
                if (rc == NTT_OK) {
                    best_fwd = f;
                    best_inv = i;
                }
            }
            // each direction on the engine that measured faster (the
            // chunking applies to two-pass launches only)
            if (rc == NTT_OK && bf < best_fwd) {
                engine = NTT_ENGINE_FUSED;
                best_fwd = bf;
            }
            if (rc == NTT_OK && bi < best_inv) {
                engine_inv = NTT_ENGINE_FUSED;
                best_inv = bi;
            }
            if (engine == NTT_ENGINE_FUSED && engine_inv == NTT_ENGINE_FUSED) best_chunk = 0;
        }
        // the default config, for comparison
        if (rc == NTT_OK) {
            TransformLaunch Lf = plan_launch(&P, buf, count, false), Li = plan_launch(&P, buf, count, true);
            default_config(logN, &Lf.cfg);
            default_config(logN, &Li.cfg);
            Lf.chunk = Li.chunk = 0;
            rc = time_median(s, reps, [&] { return launch_transform(Lf, s); }, &def_fwd);
            if (rc == NTT_OK) rc = time_median(s, reps, [&] { return launch_transform(Li, s); }, &def_inv);
        }
    }

    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    (void)hipFree(buf);
    (void)hipFree(P.d_const);
    if (rc) return rc;

    TunedEntry e;
    e.cfg = best;
    e.engine = engine;
    e.engine_inv = engine_inv;
    e.fused_eb_fwd = vf.eb;
    e.fused_eb_inv = vi.eb;
    e.fused_lazy = vf.lazy;
    e.fused_lazy_inv = vi.lazy;
    e.fused_logw_fwd = vf.logw;
    e.fused_logw_inv = vi.logw;
    e.fused_ot_fwd = vf.ot;
    e.fused_ot_inv = vi.ot;
    e.stagger = stagger;
    e.chunk = best_chunk;
    e.fwd_ms = best_fwd;
    e.inv_ms = best_inv;
    {
        std::lock_guard<std::mutex> g(ctx->tune_mu);
        ctx->tuned[std::make_pair(logN, log2_bucket(limbs * count))] = e;
    }
    if (out) {
        out->cfg = best;
        out->chunk = best_chunk;
        out->fwd_ms = best_fwd;
        out->inv_ms = best_inv;
        out->default_fwd_ms = def_fwd;
        out->default_inv_ms = def_inv;
        out->candidates = ncand;
        out->engine = engine;
        out->engine_inv = engine_inv;
        out->fused_fwd_ms = fused_fwd;
        out->fused_inv_ms = fused_inv;
        out->fused_stagger = stagger;
        out->fused_eb = vf.eb;
        out->fused_eb_inv = vi.eb;
        out->fused_lazy = vf.lazy ? 1 : 0;
        out->fused_lazy_inv = vi.lazy ? 1 : 0;
        out->fused_words_log2 = vf.logw;
        out->fused_words_log2_inv = vi.logw;
        out->fused_ot = vf.ot ? 1 : 0;
        out->fused_ot_inv = vi.ot ? 1 : 0;
    }
    return NTT_OK;
}

}  // namespace ntt

using namespace ntt;

extern "C" {

int ntt_autotune(ntt_context* ctx, int logN, const ntt_modulus_config* moduli, size_t limbs, size_t count,
                 int domain, int reps, ntt_tune_result* out) {
    if (!ctx || !moduli || limbs == 0 || count == 0 || logN < 1 || logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    if (domain != NTT_DOMAIN_PLAIN && domain != NTT_DOMAIN_MONTGOMERY) return NTT_ERR_INVALID_ARGUMENT;
    return autotune(ctx, logN, moduli, limbs, count, domain, reps < 1 ? 3 : reps, out);
}

int ntt_context_set_tuning(ntt_context* ctx, int logN, size_t total_limbs, const ntt_transform_config* cfg,
                           size_t chunk, int engine) {
    if (!ctx || !cfg || total_limbs == 0) return NTT_ERR_INVALID_ARGUMENT;
    int rc = validate_config(logN, cfg, NTT_LDS_BYTES_GFX950);
    if (rc) return rc;
    if (engine == NTT_ENGINE_FUSED && !fused_plan_supported(logN)) return NTT_ERR_CONFIG;
    TunedEntry e;
    e.cfg = *cfg;
    e.chunk = chunk;
    e.engine = engine == NTT_ENGINE_FUSED ? NTT_ENGINE_FUSED : NTT_ENGINE_TWO_PASS;
    e.engine_inv = e.engine;
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    ctx->tuned[std::make_pair(logN, log2_bucket(total_limbs))] = e;
    return NTT_OK;
}

int ntt_context_clear_tuning(ntt_context* ctx) {
    if (!ctx) return NTT_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    ctx->tuned.clear();
    return NTT_OK;
}

// Text format, one entry per line:
//   logN limbs_log2 chunk  R S W (x4: ntt1 ntt2 intt1 intt2)  fwd_ms inv_ms
int ntt_context_save_tuning(const ntt_context* ctx, const char* path) {
    if (!ctx || !path) return NTT_ERR_INVALID_ARGUMENT;
    FILE* f = std::fopen(path, "w");
    if (!f) return NTT_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    std::fprintf(f, "# ntt_amd tuning v5: logN limbs_log2 chunk r s w x4 fwd_ms inv_ms engine fused_eb_fwd "
                    "fused_eb_inv fused_lazy stagger fused_lazy_inv fused_logw_fwd fused_logw_inv fused_ot_fwd "
                    "fused_ot_inv engine_inv\n");
    for (auto& kv : ctx->tuned) {
        const TunedEntry& e = kv.second;
        const ntt_phase_config* p = &e.cfg.ntt_phase1;
        std::fprintf(f, "%d %d %zu", kv.first.first, kv.first.second, e.chunk);
        for (int i = 0; i < 4; ++i) std::fprintf(f, " %d %d %d", p[i].radix_stages, p[i].stage_merging, p[i].warp_batching);
        std::fprintf(f, " %.6f %.6f %d %d %d %d %d %d %d %d %d %d %d\n", e.fwd_ms, e.inv_ms, e.engine,
                     e.fused_eb_fwd, e.fused_eb_inv, e.fused_lazy ? 1 : 0, e.stagger, e.fused_lazy_inv ? 1 : 0,
                     e.fused_logw_fwd, e.fused_logw_inv, e.fused_ot_fwd ? 1 : 0, e.fused_ot_inv ? 1 : 0,
                     e.engine_inv);
    }
    std::fclose(f);
    return NTT_OK;
}

int ntt_context_load_tuning(ntt_context* ctx, const char* path) {
    if (!ctx || !path) return NTT_ERR_INVALID_ARGUMENT;
    FILE* f = std::fopen(path, "r");
    if (!f) return NTT_ERR_NOT_FOUND;
    char line[512];
    int loaded = 0, rc = NTT_OK;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        int logN, lb, v[12];
        size_t chunk;
        double fm = 0, im = 0;
        int eng = NTT_ENGINE_TWO_PASS, febf = FUSED_DEFAULT_EB_FWD, febi = FUSED_DEFAULT_EB_INV, lz = 1, stag = 0;
        int lzi = 1, lwf = 16, lwi = 16, otf = 0, oti = 0, engi = -1;
        int n = std::sscanf(line,
                            "%d %d %zu %d %d %d %d %d %d %d %d %d %d %d %d %lf %lf %d %d %d %d %d %d %d %d %d %d %d",
                            &logN, &lb, &chunk, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9],
                            &v[10], &v[11], &fm, &im, &eng, &febf, &febi, &lz, &stag, &lzi, &lwf, &lwi, &otf, &oti,
                            &engi);
        if (engi < 0) engi = eng;   // files of one engine for both directions
        if (n < 15) { rc = NTT_ERR_INVALID_ARGUMENT; break; }
        TunedEntry e;
        ntt_phase_config* p = &e.cfg.ntt_phase1;
        for (int i = 0; i < 4; ++i) p[i] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        if (validate_config(logN, &e.cfg, NTT_LDS_BYTES_GFX950) != NTT_OK) { rc = NTT_ERR_CONFIG; break; }
        e.chunk = chunk;
        e.engine = (eng == NTT_ENGINE_FUSED && fused_plan_supported(logN)) ? NTT_ENGINE_FUSED : NTT_ENGINE_TWO_PASS;
        e.engine_inv = (engi == NTT_ENGINE_FUSED && fused_plan_supported(logN)) ? NTT_ENGINE_FUSED : NTT_ENGINE_TWO_PASS;
        e.fused_eb_fwd = febf == 6 ? 6 : 7;
        e.fused_eb_inv = febi == 6 ? 6 : 7;
        e.fused_lazy = lz != 0;
        e.fused_lazy_inv = lzi != 0;
        e.fused_logw_fwd = lwf == 15 && logN <= 15 ? 15 : 16;
        e.fused_logw_inv = lwi == 15 && logN <= 15 ? 15 : 16;
        e.fused_ot_fwd = otf != 0;
        e.fused_ot_inv = oti != 0;
        e.stagger = stag < 0 ? 0 : (stag > 64 ? 64 : stag);
        e.fwd_ms = fm;
        e.inv_ms = im;
        std::lock_guard<std::mutex> g(ctx->tune_mu);
        ctx->tuned[std::make_pair(logN, lb)] = e;
        ++loaded;
    }
    std::fclose(f);
    return rc ? rc : (loaded ? NTT_OK : NTT_ERR_NOT_FOUND);
}

}  // extern "C"
