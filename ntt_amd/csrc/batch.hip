// Batches of uneven transforms (ntt_batch_* in include/ntt_amd.h).
//
// Jobs are grouped by (logN, direction).  A group with a fused-capable size
// becomes one persistent fused launch over a device-resident table of
// FusedJob descriptors (one per job, each with its own limb count, table
// pointers and per-limb constants), so any number of jobs of that size costs
// one kernel.  Jobs of other sizes run, per (logN, direction, domain), as one
// two-pass transform over all their limbs (per-limb operand pointers).  Work units are
// balanced over the caller's stream plus up to three streams owned by the
// batch (fork/join through events), so groups overlap on the device.  The
// same enqueue sequence is captured into a hipGraph for replay; it contains
// kernel launches and event edges only (no memset or copy nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
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
int launch_fused(int n, bool inv, int eb, bool lazy, const FusedParams& p, hipStream_t stream, unsigned grid_limit,
                 int logw = 16, bool ot = false);
int fused_prepare();
uint32_t fused_items(int logN, size_t limbs, int logw = 16);
}  // namespace ntt

using namespace ntt;

namespace {

constexpr int BATCH_AUX_STREAMS = 3;

struct PlanDeleter {
    void operator()(ntt_plan* p) const { ntt_plan_destroy(p); }
};

struct FusedGroup {
    int n = 0;
    bool inverse = false;
    FusedVariant v;                 // kernel variant (lazy only when every job's q < 2^30)
    size_t desc_begin = 0;          // first descriptor in the batch's table
    size_t njobs = 0;
    uint32_t items = 0;
    double work = 0;                // butterflies, for stream balancing
};

// Jobs of one (logN, direction, domain, arithmetic) that the fused kernel
// does not cover run as ONE two-pass transform over all their limbs, through
// per-limb operand / table / constant arrays.
struct TwoPassGroup {
    int n = 0;
    bool inverse = false;
    int domain = NTT_DOMAIN_MONTGOMERY;
    bool lazy = false;
    size_t limb_begin = 0;          // first entry in the batch's per-limb arrays
    size_t limbs = 0;
    double work = 0;
};

struct Unit {                       // one enqueue: a fused group (group >= 0) or a two-pass group
    int group = -1;
    int tp_group = -1;
    double work = 0;
};

}  // namespace

struct ntt_batch {
    ntt_context* ctx = nullptr;
    std::vector<ntt_batch_job> jobs;            // moduli pointer cleared after create
    std::vector<ntt_plan*> job_plan;
    std::vector<int> job_group;                 // fused group index or -1
    std::vector<size_t> job_desc;               // descriptor index for fused jobs
    std::map<std::tuple<int, int, std::vector<uint32_t>>, std::unique_ptr<ntt_plan, PlanDeleter>> plans;
    std::vector<FusedGroup> groups;
    std::vector<FusedJob> desc;                 // host copy of the device job table
    FusedJob* d_desc = nullptr;
    uint32_t* d_counters = nullptr;             // two words per fused group (dynamic scheduling)
    std::vector<TwoPassGroup> tp_groups;
    std::vector<size_t> job_limb0;              // two-pass jobs: first entry in the per-limb arrays
    std::vector<uint32_t*> h_A;                 // per-limb operand pointers (host copy)
    void* d_tp = nullptr;                       // per-limb arrays: A, tw, q, q_inv, scale
    uint32_t* const* d_A = nullptr;
    const uint32_t* const* d_tw = nullptr;
    const uint32_t* d_q = nullptr;
    const uint32_t* d_qi = nullptr;
    const uint32_t* d_scale = nullptr;
    std::vector<std::vector<Unit>> lanes;       // lane 0 = caller's stream
    hipStream_t aux[BATCH_AUX_STREAMS] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[BATCH_AUX_STREAMS] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~ntt_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& s : aux)
            if (s) (void)hipStreamDestroy(s);
        for (auto& e : join)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (d_desc) (void)hipFree(d_desc);
        if (d_counters) (void)hipFree(d_counters);
        if (d_tp) (void)hipFree(d_tp);
    }
};

namespace {

int ck(hipError_t e, const char* where) { return e == hipSuccess ? NTT_OK : record_hip_error(e, where); }

int enqueue(ntt_batch* b, hipStream_t stream) {
    const size_t nl = b->lanes.size();
    int rc = NTT_OK;
    if (nl > 1) {
        if ((rc = ck(hipEventRecord(b->fork, stream), "hipEventRecord(fork)"))) return rc;
        for (size_t l = 1; l < nl; ++l)
            if ((rc = ck(hipStreamWaitEvent(b->aux[l - 1], b->fork, 0), "hipStreamWaitEvent(fork)"))) return rc;
    }
    for (size_t l = 0; l < nl && !rc; ++l) {
        hipStream_t s = l == 0 ? stream : b->aux[l - 1];
        for (const Unit& u : b->lanes[l]) {
            if (u.group >= 0) {
                const FusedGroup& g = b->groups[(size_t)u.group];
                FusedParams p{};
                p.jobs = b->d_desc + g.desc_begin;
                p.njobs = (uint32_t)g.njobs;
                p.total_items = g.items;
                // dynamic scheduling: groups on the batch's streams share the
                // CUs, so a group's grid is not resident all at once; the
                // kernel resets its counters itself, so graph replays need no
                // memset node
                p.counter = b->d_counters + 2 * (size_t)u.group;
                p.stagger = 0;
                rc = launch_fused(g.n, g.inverse, g.v.eb, g.v.lazy, p, s, 0, g.v.logw, g.v.ot);
            } else {
                const TwoPassGroup& g = b->tp_groups[(size_t)u.tp_group];
                TransformLaunch L{};
                L.A_ptrs = b->d_A + g.limb_begin;
                L.tw_ptrs = b->d_tw + g.limb_begin;
                L.moduli = b->d_q + g.limb_begin;
                L.q_inv = b->d_qi + g.limb_begin;
                L.scale = b->d_scale + g.limb_begin;
                L.scale_mode = (!g.inverse && g.domain == NTT_DOMAIN_PLAIN) ? SCALE_NONE : SCALE_AS_IS;
                L.logN = g.n;
                L.limbs = g.limbs;
                L.count = 1;
                L.inverse = g.inverse;
                L.lazy = g.lazy;
                L.xcd_remap = 1;
                TunedEntry t;
                if (b->ctx->tuned_lookup(g.n, g.limbs, &t)) L.cfg = t.cfg;
                else default_config(g.n, &L.cfg);
                rc = launch_transform(L, s);
            }
            if (rc) break;
        }
    }
    if (nl > 1) {
        // join even after an error so no stream is left forked from a capture
        for (size_t l = 1; l < nl; ++l) {
            int r1 = ck(hipEventRecord(b->join[l - 1], b->aux[l - 1]), "hipEventRecord(join)");
            int r2 = ck(hipStreamWaitEvent(stream, b->join[l - 1], 0), "hipStreamWaitEvent(join)");
            if (!rc) rc = r1 ? r1 : r2;
        }
    }
    return rc;
}

void drop_graph(ntt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

}  // namespace

extern "C" {

int ntt_batch_create(ntt_context* ctx, const ntt_batch_job* jobs, size_t njobs, ntt_batch** out) {
    if (!ctx || !jobs || !out || njobs == 0) return NTT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    std::unique_ptr<ntt_batch> b(new ntt_batch());
    b->ctx = ctx;
    b->jobs.assign(jobs, jobs + njobs);
    b->job_plan.assign(njobs, nullptr);
    b->job_group.assign(njobs, -1);
    b->job_desc.assign(njobs, 0);
    int rc = NTT_OK;
    std::map<std::tuple<int, int, int>, std::vector<size_t>> by_shape;   // (logN, inverse, lazy-capable) -> jobs
    std::map<std::tuple<int, int, int, int>, std::vector<size_t>> tp_shape;   // (logN, inverse, domain, lazy)
    b->job_limb0.assign(njobs, 0);
    for (size_t i = 0; i < njobs; ++i) {
        const ntt_batch_job& j = jobs[i];
        if (!j.data || !j.moduli || j.limbs == 0 || j.logN < 1 || j.logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
        if (j.direction != NTT_FORWARD && j.direction != NTT_INVERSE) return NTT_ERR_INVALID_ARGUMENT;
        if (j.domain != NTT_DOMAIN_PLAIN && j.domain != NTT_DOMAIN_MONTGOMERY) return NTT_ERR_INVALID_ARGUMENT;
        if (reinterpret_cast<uintptr_t>(j.data) & 15) return NTT_ERR_INVALID_ARGUMENT;
        std::vector<uint32_t> key_q(2 * (size_t)j.limbs);
        for (uint32_t l = 0; l < j.limbs; ++l) {
            key_q[2 * l] = j.moduli[l].qi;
            key_q[2 * l + 1] = j.moduli[l].primitive_root;
        }
        auto key = std::make_tuple(j.logN, j.domain, std::move(key_q));
        auto it = b->plans.find(key);
        if (it == b->plans.end()) {
            ntt_plan* P = nullptr;
            if ((rc = ntt_plan_create(ctx, j.logN, j.moduli, j.limbs, j.domain, nullptr, &P))) return rc;
            it = b->plans.emplace(std::move(key), std::unique_ptr<ntt_plan, PlanDeleter>(P)).first;
        }
        b->job_plan[i] = it->second.get();
        b->jobs[i].moduli = nullptr;
        if (fused_plan_supported(j.logN))
            by_shape[std::make_tuple(j.logN, j.direction == NTT_INVERSE ? 1 : 0, it->second->lazy ? 1 : 0)].push_back(i);
        else
            tp_shape[std::make_tuple(j.logN, j.direction == NTT_INVERSE ? 1 : 0, j.domain, it->second->lazy ? 1 : 0)]
                .push_back(i);
    }

    // a fused-capable group whose size the autotuner measured faster on the
    // two-pass kernels (small groups: a few hundred limbs of 2^12..2^15 words
    // fill too few workgroups of the persistent kernel) joins the two-pass path
    for (auto it = by_shape.begin(); it != by_shape.end();) {
        size_t total_limbs = 0;
        for (size_t i : it->second) total_limbs += b->jobs[i].limbs;
        TunedEntry t;
        if (ctx->tuned_lookup(std::get<0>(it->first), total_limbs, &t) &&
            t.engine_for(std::get<1>(it->first) != 0) == NTT_ENGINE_TWO_PASS) {
            for (size_t i : it->second) {
                const ntt_batch_job& j = b->jobs[i];
                tp_shape[std::make_tuple(j.logN, j.direction == NTT_INVERSE ? 1 : 0, j.domain,
                                         b->job_plan[i]->lazy ? 1 : 0)].push_back(i);
            }
            it = by_shape.erase(it);
        } else {
            ++it;
        }
    }

    // descriptor table, one contiguous run per group
    std::vector<Unit> units;
    for (auto& kv : by_shape) {
        FusedGroup g;
        g.n = std::get<0>(kv.first);
        g.inverse = std::get<1>(kv.first) != 0;
        const bool lazy_capable = std::get<2>(kv.first) != 0;
        g.desc_begin = b->desc.size();
        g.njobs = kv.second.size();
        size_t total_limbs = 0;
        for (size_t i : kv.second) total_limbs += b->jobs[i].limbs;
        g.v = fused_variant_for(ctx, g.n, total_limbs, g.inverse, lazy_capable);
        uint64_t items = 0;
        for (size_t i : kv.second) {
            const ntt_batch_job& j = b->jobs[i];
            const ntt_plan* P = b->job_plan[i];
            FusedJob fj = fused_job_plain(j.limbs);
            fj.data = j.data;
            fj.src = j.data;
            fj.tw = g.inverse ? P->d_tw_inv : P->d_tw_fwd;
            fj.lc = g.inverse ? P->d_lc_inv : P->d_lc_fwd;
            fj.limbs = j.limbs;
            fj.period = j.limbs;
            fj.item_begin = (uint32_t)items;
            fj.n = (uint16_t)g.n;
            fj.inverse = g.inverse ? 1 : 0;
            items += fused_items(g.n, j.limbs, g.v.logw);
            if (items > 0xffffffffull) return NTT_ERR_INVALID_ARGUMENT;
            b->job_group[i] = (int)b->groups.size();
            b->job_desc[i] = b->desc.size();
            b->desc.push_back(fj);
            g.work += 0.5 * (double)((size_t)1 << g.n) * g.n * j.limbs;
        }
        g.items = (uint32_t)items;
        units.push_back({(int)b->groups.size(), -1, g.work});
        b->groups.push_back(g);
    }
    std::vector<const uint32_t*> h_tw;
    std::vector<uint32_t> h_q, h_qi, h_scale;
    for (auto& kv : tp_shape) {
        TwoPassGroup g;
        g.n = std::get<0>(kv.first);
        g.inverse = std::get<1>(kv.first) != 0;
        g.domain = std::get<2>(kv.first);
        g.lazy = std::get<3>(kv.first) != 0;
        g.limb_begin = b->h_A.size();
        for (size_t i : kv.second) {
            const ntt_batch_job& j = b->jobs[i];
            const ntt_plan* P = b->job_plan[i];
            b->job_limb0[i] = b->h_A.size();
            for (uint32_t l = 0; l < j.limbs; ++l) {
                b->h_A.push_back(j.data + ((size_t)l << g.n));
                h_tw.push_back(g.inverse ? P->h_tw_inv[l] : P->h_tw_fwd[l]);
                h_q.push_back(P->h_q[l]);
                h_qi.push_back(P->h_q_inv[l]);
                h_scale.push_back(g.inverse ? P->h_scale_inv[l] : P->h_scale_fwd[l]);
            }
            g.work += 0.5 * (double)((size_t)1 << g.n) * g.n * j.limbs;
        }
        g.limbs = b->h_A.size() - g.limb_begin;
        if (g.limbs > 0xffffffu) return NTT_ERR_INVALID_ARGUMENT;
        units.push_back({-1, (int)b->tp_groups.size(), g.work});
        b->tp_groups.push_back(g);
    }

    // longest-processing-time-first over the lanes
    const size_t nl = std::min<size_t>(units.size(), 1 + BATCH_AUX_STREAMS);
    b->lanes.assign(nl, {});
    std::vector<double> load(nl, 0.0);
    std::stable_sort(units.begin(), units.end(), [](const Unit& a, const Unit& c) { return a.work > c.work; });
    for (const Unit& u : units) {
        const size_t l = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        b->lanes[l].push_back(u);
        load[l] += u.work;
    }
    for (size_t l = 1; l < nl; ++l) {
        if ((rc = ck(hipStreamCreate(&b->aux[l - 1]), "hipStreamCreate"))) return rc;
        if ((rc = ck(hipEventCreateWithFlags(&b->join[l - 1], hipEventDisableTiming), "hipEventCreate"))) return rc;
    }
    if ((rc = ck(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate"))) return rc;

    if (!b->groups.empty() && (rc = fused_prepare())) return rc;
    if (!b->h_A.empty()) {
        const size_t T = b->h_A.size();
        const size_t bytes = T * (2 * sizeof(void*) + 3 * sizeof(uint32_t));
        if (hipMalloc(&b->d_tp, bytes) != hipSuccess) return NTT_ERR_OUT_OF_MEMORY;
        std::vector<char> h(bytes);
        char* base = reinterpret_cast<char*>(b->d_tp);
        std::memcpy(h.data(), b->h_A.data(), T * sizeof(void*));
        std::memcpy(h.data() + T * sizeof(void*), h_tw.data(), T * sizeof(void*));
        std::memcpy(h.data() + 2 * T * sizeof(void*), h_q.data(), T * 4);
        std::memcpy(h.data() + 2 * T * sizeof(void*) + T * 4, h_qi.data(), T * 4);
        std::memcpy(h.data() + 2 * T * sizeof(void*) + 2 * T * 4, h_scale.data(), T * 4);
        b->d_A = reinterpret_cast<uint32_t* const*>(base);
        b->d_tw = reinterpret_cast<const uint32_t* const*>(base + T * sizeof(void*));
        b->d_q = reinterpret_cast<const uint32_t*>(base + 2 * T * sizeof(void*));
        b->d_qi = b->d_q + T;
        b->d_scale = b->d_q + 2 * T;
        if ((rc = ck(hipMemcpy(b->d_tp, h.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(limbs)"))) return rc;
    }
    if (!b->desc.empty()) {
        const size_t bytes = b->desc.size() * sizeof(FusedJob);
        if (hipMalloc(&b->d_desc, bytes) != hipSuccess) return NTT_ERR_OUT_OF_MEMORY;
        if ((rc = ck(hipMemcpy(b->d_desc, b->desc.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(jobs)")))
            return rc;
        const size_t cb = 2 * b->groups.size() * sizeof(uint32_t);
        if (hipMalloc(&b->d_counters, cb) != hipSuccess) return NTT_ERR_OUT_OF_MEMORY;
        if ((rc = ck(hipMemset(b->d_counters, 0, cb), "hipMemset(counters)"))) return rc;
    }
    *out = b.release();
    return NTT_OK;
}

int ntt_batch_destroy(ntt_batch* batch) {
    if (!batch) return NTT_ERR_INVALID_ARGUMENT;
    delete batch;
    return NTT_OK;
}

int ntt_batch_run(ntt_batch* batch, void* stream) {
    if (!batch) return NTT_ERR_INVALID_ARGUMENT;
    TraceRange range("ntt_batch_run");
    return enqueue(batch, (hipStream_t)stream);
}

int ntt_batch_capture(ntt_batch* batch, void* stream) {
    if (!batch) return NTT_ERR_INVALID_ARGUMENT;
    drop_graph(batch);
    hipStream_t s = (hipStream_t)stream;
    int rc = ck(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    if (rc) return rc;
    const int erc = enqueue(batch, s);
    hipGraph_t g = nullptr;
    rc = ck(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
    if (erc || rc) {
        if (g) (void)hipGraphDestroy(g);
        return erc ? erc : rc;
    }
    batch->graph = g;
    if ((rc = ck(hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate"))) {
        drop_graph(batch);
        return rc;
    }
    return NTT_OK;
}

int ntt_batch_replay(ntt_batch* batch, void* stream) {
    if (!batch) return NTT_ERR_INVALID_ARGUMENT;
    TraceRange range("ntt_batch_replay");
    int rc = NTT_OK;
    if (!batch->exec && (rc = ntt_batch_capture(batch, stream))) return rc;
    return ck(hipGraphLaunch(batch->exec, (hipStream_t)stream), "hipGraphLaunch");
}

int ntt_batch_set_data(ntt_batch* batch, size_t index, uint32_t* data) {
    if (!batch || index >= batch->jobs.size() || !data || (reinterpret_cast<uintptr_t>(data) & 15))
        return NTT_ERR_INVALID_ARGUMENT;
    batch->jobs[index].data = data;
    // the graph may still be running on another stream: copy after the device is idle
    int rc = ck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc) return rc;
    if (batch->job_group[index] < 0) {
        // two-pass job: its limbs' entries in the per-limb operand array
        const ntt_batch_job& j = batch->jobs[index];
        const size_t l0 = batch->job_limb0[index];
        for (uint32_t l = 0; l < j.limbs; ++l) batch->h_A[l0 + l] = data + ((size_t)l << j.logN);
        return ck(hipMemcpy(const_cast<uint32_t**>(batch->d_A) + l0, batch->h_A.data() + l0, j.limbs * sizeof(void*),
                            hipMemcpyHostToDevice),
                  "hipMemcpy(limb pointers)");
    }
    const size_t d = batch->job_desc[index];
    batch->desc[d].data = data;
    batch->desc[d].src = data;
    return ck(hipMemcpy(batch->d_desc + d, &batch->desc[d], sizeof(FusedJob), hipMemcpyHostToDevice),
              "hipMemcpy(job)");
}

int ntt_batch_info(const ntt_batch* batch, int* fused_groups, int* two_pass_jobs, int* streams) {
    if (!batch) return NTT_ERR_INVALID_ARGUMENT;
    int tp = 0;
    for (int g : batch->job_group) tp += g < 0 ? 1 : 0;
    if (fused_groups) *fused_groups = (int)batch->groups.size();
    if (two_pass_jobs) *two_pass_jobs = tp;
    if (streams) *streams = (int)batch->lanes.size();
    return NTT_OK;
}

}  // extern "C"
