// Request loop with continuous batching (ntt_server_* in include/ntt_amd.h).
//
// One scheduler thread per server.  Each pass of its loop:
//   1. retires launch slots whose completion event has fired (requests DONE,
//      auto-release operands freed back to the arena);
//   2. admits queued requests in arrival order while their operand fits the
//      arena (first-fit over a free list, 256-byte granules);
//   3. hands all admitted requests (up to max_jobs_per_launch) to a free
//      slot: on that slot's stream it places each operand (copy or seeded
//      fill), then launches one persistent fused kernel per (logN,
//      direction) group over a job table staged through pinned memory, and
//      two-pass kernels for sizes the fused kernel does not cover; then it
//      records the slot's completion event.
// Slots run on non-blocking streams, so launches of consecutive passes
// overlap on the device while new requests keep arriving.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/ntt_amd.h"
#include "../../include/ntt_amd_runtime.h"
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

using Clock = std::chrono::steady_clock;
constexpr int SERVER_MAX_SLOTS = 8;
// Poll interval of the scheduler while launches are in flight.
constexpr int SERVER_POLL_US = 20;
constexpr size_t ARENA_GRANULE = 256;

int ck(hipError_t e, const char* where) { return e == hipSuccess ? NTT_OK : record_hip_error(e, where); }

// First-fit allocator over [0, size) in ARENA_GRANULE units, coalescing frees.
class ArenaAllocator {
public:
    void reset(size_t bytes) {
        free_.clear();
        size_ = bytes / ARENA_GRANULE * ARENA_GRANULE;
        if (size_) free_[0] = size_;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + ARENA_GRANULE - 1) / ARENA_GRANULE * ARENA_GRANULE;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + ARENA_GRANULE - 1) / ARENA_GRANULE * ARENA_GRANULE;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0;
};

struct Req {
    uint64_t id = 0;
    ntt_request r{};
    ntt_plan* plan = nullptr;
    size_t bytes = 0;
    size_t off = 0;
    bool placed = false;            // holds arena space
    int pins = 0;                   // fetches copying out of the arena right now
    bool release_pending = false;   // released while pinned: freed by the last fetch
    int state = NTT_REQ_QUEUED;
    int status = NTT_OK;
    Clock::time_point arrival, done;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    FusedJob* h_desc = nullptr;     // pinned staging
    FusedJob* d_desc = nullptr;
    uint32_t* d_counter = nullptr;  // dynamic scheduling of the slot's fused launches (two words)
    // per-limb arrays of the two-pass groups (operand, table, q, q_inv, scale),
    // pinned staging and device copy, SLOT_LIMB_BYTES bytes per limb
    char* h_limb = nullptr;
    char* d_limb = nullptr;
    size_t limb_cap = 0;            // limbs
    FillJob* h_fill = nullptr;      // seeded-operand fills of a launch, pinned staging
    FillJob* d_fill = nullptr;
    bool busy = false;
    std::vector<std::shared_ptr<Req>> reqs;
};

using PlanKey = std::tuple<int, int, std::vector<uint32_t>>;
constexpr size_t SLOT_LIMB_BYTES = 2 * sizeof(void*) + 3 * sizeof(uint32_t);
constexpr size_t SLOT_LIMBS_PER_JOB = 64;   // per-limb capacity of a slot = max jobs x this

}  // namespace

struct ntt_server {
    ntt_context* ctx = nullptr;
    ntt_server_config cfg{};
    uint32_t* arena = nullptr;
    ArenaAllocator alloc;
    Slot slots[SERVER_MAX_SLOTS];
    int nslots = 0;

    // Plans by request shape.  Guarded by plan_mu alone, so that building a
    // new shape's tables (host work plus a blocking upload) never holds up
    // the scheduler; plans are only added, and live until the server ends.
    std::mutex plan_mu;
    std::map<PlanKey, ntt_plan*> plans;
    std::map<std::pair<int, uint32_t>, std::vector<ntt_modulus_config>> default_moduli;

    std::mutex mu;                  // guards everything below
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> queued;
    std::deque<std::shared_ptr<Req>> admitted;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> live;
    uint64_t next_id = 1;
    ntt_server_stats stats{};
    uint64_t jobs_launched = 0;
    bool stop = false;
    std::thread worker;

    ~ntt_server() {
        for (auto& kv : plans) ntt_plan_destroy(kv.second);
        for (int i = 0; i < nslots; ++i) {
            Slot& s = slots[i];
            if (s.stream) (void)hipStreamDestroy(s.stream);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.h_desc) (void)hipHostFree(s.h_desc);
            if (s.d_desc) (void)hipFree(s.d_desc);
            if (s.d_counter) (void)hipFree(s.d_counter);
            if (s.h_limb) (void)hipHostFree(s.h_limb);
            if (s.d_limb) (void)hipFree(s.d_limb);
            if (s.h_fill) (void)hipHostFree(s.h_fill);
            if (s.d_fill) (void)hipFree(s.d_fill);
        }
        if (arena) (void)hipFree(arena);
    }
};

namespace {

// Plan for a request's shape (takes plan_mu; the caller must not hold mu).
int plan_for(ntt_server* S, const ntt_request& r, ntt_plan** out) {
    std::lock_guard<std::mutex> g(S->plan_mu);
    const ntt_modulus_config* m = r.moduli;
    if (!m) {
        auto key = std::make_pair((int)r.logN, r.limbs);
        auto it = S->default_moduli.find(key);
        if (it == S->default_moduli.end()) {
            std::vector<ntt_modulus_config> v(r.limbs);
            const int got = ntt_generate_moduli(r.logN, 1u << 30, (int)r.limbs, v.data());
            if (got != (int)r.limbs) return got < 0 ? -got : NTT_ERR_MODULUS_RANGE;
            it = S->default_moduli.emplace(key, std::move(v)).first;
        }
        m = it->second.data();
    }
    std::vector<uint32_t> kq(2 * (size_t)r.limbs);
    for (uint32_t l = 0; l < r.limbs; ++l) {
        kq[2 * l] = m[l].qi;
        kq[2 * l + 1] = m[l].primitive_root;
    }
    PlanKey key(r.logN, r.domain, std::move(kq));
    auto it = S->plans.find(key);
    if (it != S->plans.end()) {
        *out = it->second;
        return NTT_OK;
    }
    ntt_plan* P = nullptr;
    int rc = ntt_plan_create(S->ctx, r.logN, m, r.limbs, r.domain, nullptr, &P);
    if (rc) return rc;
    S->plans.emplace(std::move(key), P);
    *out = P;
    return NTT_OK;
}

// Enqueues one launch of `reqs` on slot `s` (no lock held; requests are owned
// by this launch until the slot retires).
int enqueue_launch(ntt_server* S, Slot& s, const std::vector<std::shared_ptr<Req>>& reqs, uint64_t* fused,
                   uint64_t* two_pass) {
    TraceRange range("ntt_server_launch");
    int rc = NTT_OK;
    // 1. operands into the arena: copies, and every seeded fill in one launch
    uint32_t nfill = 0;
    uint64_t fill_words = 0;
    for (const auto& q : reqs) {
        uint32_t* dst = S->arena + q->off / sizeof(uint32_t);
        const ntt_request& r = q->r;
        if (r.src_dev) rc = ck(hipMemcpyAsync(dst, r.src_dev, q->bytes, hipMemcpyDeviceToDevice, s.stream),
                                "hipMemcpyAsync(src_dev)");
        else if (r.src_host) rc = ck(hipMemcpyAsync(dst, r.src_host, q->bytes, hipMemcpyHostToDevice, s.stream),
                                     "hipMemcpyAsync(src_host)");
        else {
            s.h_fill[nfill++] = FillJob{dst, q->plan->d_moduli, fill_words, r.seed, r.limbs, r.logN};
            fill_words += (uint64_t)r.limbs << r.logN;
        }
        if (rc) return rc;
    }
    if (nfill) {
        rc = ck(hipMemcpyAsync(s.d_fill, s.h_fill, nfill * sizeof(FillJob), hipMemcpyHostToDevice, s.stream),
                "hipMemcpyAsync(fills)");
        if (rc == NTT_OK) rc = launch_fill_jobs(s.d_fill, nfill, fill_words, s.stream);
        if (rc) return rc;
    }
    // 2. group by (logN, direction)
    // (logN, direction, lazy-capable, domain)
    std::map<std::tuple<int, int, int, int>, std::vector<size_t>> groups;
    for (size_t i = 0; i < reqs.size(); ++i)
        groups[std::make_tuple(reqs[i]->r.logN, reqs[i]->r.direction, reqs[i]->plan->lazy ? 1 : 0, reqs[i]->r.domain)]
            .push_back(i);
    size_t used = 0;                 // bytes of the slot's per-limb buffers, 16-byte aligned regions
    size_t nd = 0;
    struct FusedLaunch {
        int n;
        bool inv;
        size_t first, count;
        uint32_t items;
        FusedVariant v;
    };
    std::vector<FusedLaunch> fl;
    for (auto& kv : groups) {
        const int n = std::get<0>(kv.first);
        const bool inv = std::get<1>(kv.first) == NTT_INVERSE;
        size_t tot = 0;
        for (size_t i : kv.second) tot += reqs[i]->r.limbs;
        // the engine the autotuner measured faster for a launch of this size
        // (fused when the size is untuned)
        TunedEntry tuned;
        const bool tuned_two_pass = S->ctx->tuned_lookup(n, tot, &tuned) && tuned.engine_for(inv) == NTT_ENGINE_TWO_PASS;
        if (fused_plan_supported(n) && !tuned_two_pass) {
            const size_t first = nd;
            const FusedVariant v = fused_variant_for(S->ctx, n, tot, inv, std::get<2>(kv.first) != 0);
            uint64_t items = 0;
            for (size_t i : kv.second) {
                const Req& q = *reqs[i];
                FusedJob& j = s.h_desc[nd++];
                j = fused_job_plain(q.r.limbs);
                j.data = S->arena + q.off / sizeof(uint32_t);
                j.src = j.data;
                j.tw = inv ? q.plan->d_tw_inv : q.plan->d_tw_fwd;
                j.lc = inv ? q.plan->d_lc_inv : q.plan->d_lc_fwd;
                j.limbs = q.r.limbs;
                j.period = q.r.limbs;
                j.item_begin = (uint32_t)items;
                j.n = (uint16_t)n;
                j.inverse = inv ? 1 : 0;
                items += fused_items(n, q.r.limbs, v.logw);
            }
            fl.push_back(FusedLaunch{n, inv, first, kv.second.size(), (uint32_t)items, v});
        } else {
            size_t T = 0;
            for (size_t i : kv.second) T += reqs[i]->r.limbs;
            const size_t region = (T * SLOT_LIMB_BYTES + 15) & ~(size_t)15;
            if (used + region > s.limb_cap * SLOT_LIMB_BYTES) {
                // beyond the slot's per-limb capacity: one launch per request
                for (size_t i : kv.second) {
                    const Req& q = *reqs[i];
                    rc = launch_transform(plan_launch(q.plan, S->arena + q.off / sizeof(uint32_t), 1, inv), s.stream);
                    if (rc) return rc;
                    ++*two_pass;
                }
                continue;
            }
            // one two-pass transform over every limb of the group
            char* h = s.h_limb + used;
            char* d = s.d_limb + used;
            uint32_t** hA = reinterpret_cast<uint32_t**>(h);
            const uint32_t** hT = reinterpret_cast<const uint32_t**>(h + T * sizeof(void*));
            uint32_t* hq = reinterpret_cast<uint32_t*>(h + 2 * T * sizeof(void*));
            size_t k = 0;
            for (size_t i : kv.second) {
                const Req& q = *reqs[i];
                uint32_t* base = S->arena + q.off / sizeof(uint32_t);
                for (uint32_t l = 0; l < q.r.limbs; ++l, ++k) {
                    hA[k] = base + ((size_t)l << n);
                    hT[k] = inv ? q.plan->h_tw_inv[l] : q.plan->h_tw_fwd[l];
                    hq[k] = q.plan->h_q[l];
                    hq[T + k] = q.plan->h_q_inv[l];
                    hq[2 * T + k] = inv ? q.plan->h_scale_inv[l] : q.plan->h_scale_fwd[l];
                }
            }
            rc = ck(hipMemcpyAsync(d, h, T * SLOT_LIMB_BYTES, hipMemcpyHostToDevice, s.stream), "hipMemcpyAsync(limbs)");
            if (rc) return rc;
            used += region;
            TransformLaunch L{};
            L.A_ptrs = reinterpret_cast<uint32_t* const*>(d);
            L.tw_ptrs = reinterpret_cast<const uint32_t* const*>(d + T * sizeof(void*));
            L.moduli = reinterpret_cast<const uint32_t*>(d + 2 * T * sizeof(void*));
            L.q_inv = L.moduli + T;
            L.scale = L.moduli + 2 * T;
            L.scale_mode = (!inv && std::get<3>(kv.first) == NTT_DOMAIN_PLAIN) ? SCALE_NONE : SCALE_AS_IS;
            L.logN = n;
            L.limbs = T;
            L.count = 1;
            L.inverse = inv;
            L.lazy = std::get<2>(kv.first) != 0;
            L.xcd_remap = 1;
            TunedEntry t;
            if (S->ctx->tuned_lookup(n, T, &t)) L.cfg = t.cfg;
            else default_config(n, &L.cfg);
            if ((rc = launch_transform(L, s.stream))) return rc;
            ++*two_pass;
        }
    }
    if (nd) {
        rc = ck(hipMemcpyAsync(s.d_desc, s.h_desc, nd * sizeof(FusedJob), hipMemcpyHostToDevice, s.stream),
                "hipMemcpyAsync(jobs)");
        if (rc) return rc;
    }
    for (auto& g : fl) {
        FusedParams p{};
        p.jobs = s.d_desc + g.first;
        p.njobs = (uint32_t)g.count;
        p.total_items = g.items;
        // the slots' launches share the CUs: items are claimed dynamically; the
        // slot's launches run in order on its stream, and each kernel resets
        // the counter pair before the next one starts
        p.counter = s.d_counter;
        if ((rc = launch_fused(g.n, g.inv, g.v.eb, g.v.lazy, p, s.stream, 0, g.v.logw, g.v.ot))) return rc;
        ++*fused;
    }
    return ck(hipEventRecord(s.done, s.stream), "hipEventRecord(slot)");
}

void scheduler(ntt_server* S) {
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        bool progress = false;
        // 1. retire
        int busy = 0;
        for (int i = 0; i < S->nslots; ++i) {
            Slot& s = S->slots[i];
            if (!s.busy) continue;
            hipError_t e = hipEventQuery(s.done);
            if (e == hipErrorNotReady) {
                ++busy;
                continue;
            }
            const auto now = Clock::now();
            const int st = e == hipSuccess ? NTT_OK : record_hip_error(e, "launch slot");
            for (auto& q : s.reqs) {
                q->done = now;
                q->state = st ? NTT_REQ_FAILED : NTT_REQ_DONE;
                q->status = st;
                ++S->stats.completed;
                if (q->r.auto_release && q->placed) {
                    S->alloc.release(q->off, q->bytes);
                    q->placed = false;
                }
            }
            s.reqs.clear();
            s.busy = false;
            progress = true;
            S->cv_done.notify_all();
        }
        // shutting down: requests not yet admitted are failed, not waited for
        // (the arena may be held by results nobody will release)
        if (S->stop && !S->queued.empty()) {
            for (auto& q : S->queued) {
                q->state = NTT_REQ_FAILED;
                q->status = NTT_ERR_BUSY;
                q->done = Clock::now();
            }
            S->queued.clear();
            S->cv_done.notify_all();
        }
        // 2. admit in arrival order
        while (!S->queued.empty()) {
            auto& q = S->queued.front();
            size_t off = 0;
            if (!S->alloc.alloc(q->bytes, &off)) break;
            q->off = off;
            q->placed = true;
            q->state = NTT_REQ_ADMITTED;
            S->admitted.push_back(q);
            S->queued.pop_front();
            S->stats.arena_bytes = S->alloc.used();
            S->stats.arena_peak_bytes = std::max(S->stats.arena_peak_bytes, S->alloc.used());
            progress = true;
        }
        // 3. launch on free slots
        for (int i = 0; i < S->nslots && !S->admitted.empty(); ++i) {
            Slot& s = S->slots[i];
            if (s.busy) continue;
            std::vector<std::shared_ptr<Req>> take;
            while (!S->admitted.empty() && (int)take.size() < S->cfg.max_jobs_per_launch) {
                take.push_back(S->admitted.front());
                S->admitted.pop_front();
            }
            for (auto& q : take) q->state = NTT_REQ_RUNNING;
            s.busy = true;
            s.reqs = take;
            ++busy;
            S->stats.max_slots_busy = std::max(S->stats.max_slots_busy, busy);
            lk.unlock();
            uint64_t fused = 0, two = 0;
            const int rc = enqueue_launch(S, s, take, &fused, &two);
            lk.lock();
            S->stats.launches += 1;
            S->stats.fused_launches += fused;
            S->stats.two_pass_launches += two;
            S->jobs_launched += take.size();
            S->stats.mean_jobs_per_launch = (double)S->jobs_launched / (double)S->stats.launches;
            if (rc) {
                // nothing was recorded on the event for this launch: fail its requests now
                (void)hipStreamSynchronize(s.stream);
                for (auto& q : s.reqs) {
                    q->state = NTT_REQ_FAILED;
                    q->status = rc;
                    q->done = Clock::now();
                    ++S->stats.completed;
                    if (q->placed) {
                        S->alloc.release(q->off, q->bytes);
                        q->placed = false;
                    }
                }
                s.reqs.clear();
                s.busy = false;
                S->cv_done.notify_all();
            }
            progress = true;
        }
        bool any_busy = false;
        for (int i = 0; i < S->nslots; ++i) any_busy = any_busy || S->slots[i].busy;
        if (S->stop && !any_busy && S->admitted.empty() && S->queued.empty()) break;
        if (!progress) {
            // In-flight work: poll the slots' events again after a short
            // timed wait (a submit or release wakes it earlier); idle: sleep
            // until a submit.
            S->cv_work.wait_for(lk, any_busy ? std::chrono::microseconds(SERVER_POLL_US)
                                             : std::chrono::microseconds(5000));
        }
    }
}

}  // namespace

extern "C" {

int ntt_server_create(ntt_context* ctx, const ntt_server_config* cfg, ntt_server** out) {
    if (!ctx || !cfg || !out) return NTT_ERR_INVALID_ARGUMENT;
    if (cfg->slots < 1 || cfg->slots > SERVER_MAX_SLOTS || cfg->max_jobs_per_launch < 1 ||
        cfg->max_jobs_per_launch > 4096 || cfg->hbm_budget_bytes < ARENA_GRANULE)
        return NTT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int rc = ck(hipSetDevice(ctx->device), "hipSetDevice");
    if (rc) return rc;
    if ((rc = fused_prepare())) return rc;
    std::unique_ptr<ntt_server> S(new ntt_server());
    S->ctx = ctx;
    S->cfg = *cfg;
    if (hipMalloc(&S->arena, cfg->hbm_budget_bytes) != hipSuccess) {
        S->arena = nullptr;
        return NTT_ERR_OUT_OF_MEMORY;
    }
    S->alloc.reset(cfg->hbm_budget_bytes);
    S->nslots = cfg->slots;
    const size_t db = (size_t)cfg->max_jobs_per_launch * sizeof(FusedJob);
    for (int i = 0; i < S->nslots; ++i) {
        Slot& s = S->slots[i];
        if ((rc = ck(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
        if ((rc = ck(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate"))) return rc;
        if ((rc = ck(hipHostMalloc((void**)&s.h_desc, db, hipHostMallocDefault), "hipHostMalloc"))) return rc;
        if (hipMalloc(&s.d_desc, db) != hipSuccess) {
            s.d_desc = nullptr;
            return NTT_ERR_OUT_OF_MEMORY;
        }
        if (hipMalloc(&s.d_counter, 2 * sizeof(uint32_t)) != hipSuccess) {
            s.d_counter = nullptr;
            return NTT_ERR_OUT_OF_MEMORY;
        }
        if ((rc = ck(hipMemset(s.d_counter, 0, 2 * sizeof(uint32_t)), "hipMemset(counter)"))) return rc;
        if ((rc = ck(hipHostMalloc((void**)&s.h_fill, cfg->max_jobs_per_launch * sizeof(FillJob), hipHostMallocDefault),
                     "hipHostMalloc")))
            return rc;
        if (hipMalloc(&s.d_fill, cfg->max_jobs_per_launch * sizeof(FillJob)) != hipSuccess) {
            s.d_fill = nullptr;
            return NTT_ERR_OUT_OF_MEMORY;
        }
        s.limb_cap = (size_t)cfg->max_jobs_per_launch * SLOT_LIMBS_PER_JOB;
        if ((rc = ck(hipHostMalloc((void**)&s.h_limb, s.limb_cap * SLOT_LIMB_BYTES, hipHostMallocDefault),
                     "hipHostMalloc")))
            return rc;
        if (hipMalloc(&s.d_limb, s.limb_cap * SLOT_LIMB_BYTES) != hipSuccess) {
            s.d_limb = nullptr;
            return NTT_ERR_OUT_OF_MEMORY;
        }
    }
    ntt_server* raw = S.release();
    raw->worker = std::thread(scheduler, raw);
    *out = raw;
    return NTT_OK;
}

int ntt_server_destroy(ntt_server* server) {
    if (!server) return NTT_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> g(server->mu);
        server->stop = true;
    }
    server->cv_work.notify_all();
    if (server->worker.joinable()) server->worker.join();
    (void)hipDeviceSynchronize();
    delete server;
    return NTT_OK;
}

int ntt_server_submit(ntt_server* server, const ntt_request* req, uint64_t* id) {
    if (!server || !req || !id) return NTT_ERR_INVALID_ARGUMENT;
    const ntt_request& r = *req;
    if (r.limbs == 0 || r.logN < 1 || r.logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    if (r.direction != NTT_FORWARD && r.direction != NTT_INVERSE) return NTT_ERR_INVALID_ARGUMENT;
    if (r.domain != NTT_DOMAIN_PLAIN && r.domain != NTT_DOMAIN_MONTGOMERY) return NTT_ERR_INVALID_ARGUMENT;
    auto q = std::make_shared<Req>();
    q->r = r;
    q->r.moduli = nullptr;               // resolved into the plan below
    q->bytes = ((size_t)r.limbs << r.logN) * sizeof(uint32_t);
    q->arrival = Clock::now();
    if (q->bytes > server->alloc.size()) return NTT_ERR_BUDGET;
    int rc = plan_for(server, r, &q->plan);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(server->mu);
    q->id = server->next_id++;
    server->live[q->id] = q;
    server->queued.push_back(q);
    ++server->stats.submitted;
    *id = q->id;
    server->cv_work.notify_one();
    return NTT_OK;
}

int ntt_server_state(ntt_server* server, uint64_t id, int* state) {
    if (!server || !state) return NTT_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(server->mu);
    auto it = server->live.find(id);
    *state = it == server->live.end() ? NTT_REQ_UNKNOWN : it->second->state;
    return NTT_OK;
}

int ntt_server_wait(ntt_server* server, uint64_t id) {
    if (!server) return NTT_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(server->mu);
    auto it = server->live.find(id);
    if (it == server->live.end()) return NTT_ERR_NOT_FOUND;
    auto q = it->second;
    server->cv_done.wait(lk, [&] { return q->state == NTT_REQ_DONE || q->state == NTT_REQ_FAILED; });
    return q->status;
}

namespace {
// Frees a finished request's arena range and forgets it (caller holds mu).
void release_locked(ntt_server* S, const std::shared_ptr<Req>& q) {
    if (q->placed) {
        S->alloc.release(q->off, q->bytes);
        q->placed = false;
        S->stats.arena_bytes = S->alloc.used();
    }
    S->live.erase(q->id);
    S->cv_work.notify_one();
}
}  // namespace

int ntt_server_fetch(ntt_server* server, uint64_t id, uint32_t* dst, int dst_is_device, int release) {
    if (!server || !dst) return NTT_ERR_INVALID_ARGUMENT;
    std::shared_ptr<Req> q;
    {
        std::lock_guard<std::mutex> g(server->mu);
        auto it = server->live.find(id);
        if (it == server->live.end()) return NTT_ERR_NOT_FOUND;
        q = it->second;
        if (q->state != NTT_REQ_DONE) return q->state == NTT_REQ_FAILED ? q->status : NTT_ERR_BUSY;
        if (!q->placed || q->release_pending) return NTT_ERR_NOT_FOUND;   // auto-released or released
        ++q->pins;   // the arena range stays allocated until this copy is done
    }
    const int rc = ck(hipMemcpy(dst, server->arena + q->off / sizeof(uint32_t), q->bytes,
                                dst_is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost),
                      "hipMemcpy(fetch)");
    std::lock_guard<std::mutex> g(server->mu);
    --q->pins;
    if (release && rc == NTT_OK) q->release_pending = true;
    if (q->release_pending && q->pins == 0) release_locked(server, q);
    return rc;
}

int ntt_server_release(ntt_server* server, uint64_t id) {
    if (!server) return NTT_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(server->mu);
    auto it = server->live.find(id);
    if (it == server->live.end()) return NTT_ERR_NOT_FOUND;
    auto q = it->second;
    if (q->state != NTT_REQ_DONE && q->state != NTT_REQ_FAILED) return NTT_ERR_BUSY;
    if (q->release_pending) return NTT_ERR_NOT_FOUND;
    q->release_pending = true;
    // a fetch still copying the result frees it when it finishes
    if (q->pins == 0) release_locked(server, q);
    return NTT_OK;
}

int ntt_server_stats_get(ntt_server* server, ntt_server_stats* out) {
    if (!server || !out) return NTT_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(server->mu);
    *out = server->stats;
    out->arena_bytes = server->alloc.used();
    return NTT_OK;
}

int ntt_server_prepare(ntt_server* server, int logN, uint32_t limbs, int domain) {
    if (!server || limbs == 0 || logN < 1 || logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
    ntt_request r{};
    r.logN = logN;
    r.limbs = limbs;
    r.domain = domain;
    ntt_plan* P = nullptr;
    return plan_for(server, r, &P);
}

int ntt_server_run_trace(ntt_server* server, size_t n, const double* arrival_s, const int32_t* logN,
                         const uint32_t* limbs, const int32_t* direction, const uint64_t* seed, int domain,
                         double* latency_s, double* wall_s) {
    if (!server || (n && (!arrival_s || !logN || !limbs || !direction || !seed || !latency_s)))
        return NTT_ERR_INVALID_ARGUMENT;
    for (size_t i = 0; i < n; ++i) {
        int rc = ntt_server_prepare(server, logN[i], limbs[i], domain);
        if (rc) return rc;
    }
    std::vector<uint64_t> ids(n);
    const auto t0 = Clock::now();
    int rc = NTT_OK;
    for (size_t i = 0; i < n && !rc; ++i) {
        const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
        while (Clock::now() < due) {
            // coarse sleep while far from the arrival, spin for the last stretch
            if (due - Clock::now() > std::chrono::microseconds(200))
                std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
        ntt_request r{};
        r.logN = logN[i];
        r.limbs = limbs[i];
        r.direction = direction[i];
        r.domain = domain;
        r.seed = seed[i];
        r.auto_release = 1;
        rc = ntt_server_submit(server, &r, &ids[i]);
        if (rc == NTT_OK) {
            // arrival = the scheduled time, not the (slightly later) submit time
            std::lock_guard<std::mutex> g(server->mu);
            server->live[ids[i]]->arrival = due;
        }
    }
    if (rc) return rc;
    Clock::time_point last = t0;
    int status = NTT_OK;
    for (size_t i = 0; i < n; ++i) {
        int s = ntt_server_wait(server, ids[i]);
        if (s && !status) status = s;
        std::lock_guard<std::mutex> g(server->mu);
        auto q = server->live[ids[i]];
        latency_s[i] = std::chrono::duration<double>(q->done - q->arrival).count();
        last = std::max(last, q->done);
        server->live.erase(ids[i]);
    }
    if (wall_s) {
        const double first = n ? arrival_s[0] : 0.0;
        *wall_s = std::chrono::duration<double>(last - t0).count() - first;
    }
    return status;
}

}  // extern "C"
