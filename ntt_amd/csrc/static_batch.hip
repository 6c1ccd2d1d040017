// Allocation-free batch launch (ntt_batch_run_static in include/ntt_amd.h).
//
// ntt_batch_run launches uneven transforms through device-resident job
// tables that ntt_batch_create allocates: the fused kernel's FusedJob
// descriptors with per-limb table pointers and constants, and the per-limb
// operand / table / constant arrays of the two-pass groups.  This entry point
// runs the same kernels over jobs that bring their own device tables (the
// arrays the reference's launchers take, src/v04_montgomery.cu:603-658), and
// keeps those job tables in a fixed file-scope device array instead of an
// allocation, so it calls no hipMalloc/hipFree at all.
//
// Each call first launches a setup kernel (its job list travels in the kernel
// arguments) that writes the job tables into the array, then the transforms,
// all in order on the caller's stream.  Every call on one GPU uses that GPU's
// one array: calls that overlap on a device (other streams, other threads)
// race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/ntt_amd.h"
#include "arith.hpp"
#include "fused_types.hpp"
#include "host_math.hpp"
#include "launch.hpp"
#include "phase_kernels.hpp"
#include "trace.hpp"

namespace ntt {
int launch_fused(int n, bool inv, int eb, bool lazy, const FusedParams& p, hipStream_t stream, unsigned grid_limit,
                 int logw = 16, bool ot = false);
uint32_t fused_items(int logN, size_t limbs, int logw = 16);
bool fused_plan_supported(int logN);
}  // namespace ntt

using namespace ntt;

namespace {

constexpr size_t MAX_JOBS = NTT_STATIC_MAX_JOBS;
constexpr size_t MAX_LIMBS = NTT_STATIC_MAX_LIMBS;

// Byte layout of the array: the fused descriptors, then per-limb arrays.
constexpr size_t OFF_DESC = 0;
constexpr size_t OFF_TWP = OFF_DESC + MAX_JOBS * sizeof(FusedJob);      // const uint32_t* per limb
constexpr size_t OFF_LC = OFF_TWP + MAX_LIMBS * sizeof(void*);          // LimbConst per limb
constexpr size_t OFF_AP = OFF_LC + MAX_LIMBS * sizeof(LimbConst);       // uint32_t* per limb
constexpr size_t OFF_Q = OFF_AP + MAX_LIMBS * sizeof(void*);            // q, q_inv, scale per limb
constexpr size_t OFF_QI = OFF_Q + MAX_LIMBS * sizeof(uint32_t);
constexpr size_t OFF_SC = OFF_QI + MAX_LIMBS * sizeof(uint32_t);
constexpr size_t SCRATCH_BYTES = OFF_SC + MAX_LIMBS * sizeof(uint32_t);
static_assert(sizeof(FusedJob) % 16 == 0 && OFF_LC % 16 == 0, "aligned regions");

}  // namespace

// The job tables of the call in flight on this device (about 0.75 MiB).
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr uint32_t NO_DESC = 0xffffffffu;
constexpr int SETUP_JOBS = 16;    // jobs per setup launch (kernel arguments of about 1 KiB)

struct StaticJob {
    uint32_t* data;
    const uint32_t* tw;
    const uint32_t* q;
    const uint32_t* q_inv;
    const uint32_t* scale;
    uint32_t limbs;
    uint32_t limb_begin;   // first entry of this job in the per-limb arrays
    uint32_t desc;         // descriptor slot (fused job), or NO_DESC (two-pass job)
    uint32_t item_begin;   // fused job: first work item within its group
    int32_t logN;
    int32_t inverse;
};

struct SetupArgs {
    StaticJob job[SETUP_JOBS];
    unsigned char* base;   // device address of `scratch`
};

// One workgroup per job: its per-limb entries and, for a fused job, its
// descriptor (pointing at those entries).
__global__ __launch_bounds__(256) void static_setup_kernel(SetupArgs a) {
    const StaticJob& j = a.job[blockIdx.x];
    const uint32_t** twp = reinterpret_cast<const uint32_t**>(a.base + OFF_TWP);
    LimbConst* lc = reinterpret_cast<LimbConst*>(a.base + OFF_LC);
    uint32_t** ap = reinterpret_cast<uint32_t**>(a.base + OFF_AP);
    uint32_t* Q = reinterpret_cast<uint32_t*>(a.base + OFF_Q);
    uint32_t* QI = reinterpret_cast<uint32_t*>(a.base + OFF_QI);
    uint32_t* SC = reinterpret_cast<uint32_t*>(a.base + OFF_SC);
    for (uint32_t l = threadIdx.x; l < j.limbs; l += blockDim.x) {
        const uint32_t k = j.limb_begin + l;
        const uint32_t q = j.q[l], qi = j.q_inv[l], s = j.scale[l];
        twp[k] = j.tw + ((size_t)l << j.logN);
        ap[k] = j.data + ((size_t)l << j.logN);
        Q[k] = q;
        QI[k] = qi;
        SC[k] = s;
        // fused kernels multiply by the scale word with one mont_mul: R^2 on
        // the forward input (x R), N^-1 = from_mont(to_mont(N^-1)) on the
        // inverse output (NormalizeKernel, src/v04_montgomery.cu:567-580)
        lc[k] = LimbConst{q, qi, j.inverse ? mont_mul(s, 1u, q, qi) : s, 1u};
    }
    if (threadIdx.x == 0 && j.desc != NO_DESC) {
        FusedJob d{};
        d.src_period = j.limbs;
        d.sub_end = j.limbs;
        d.data = j.data;
        d.src = j.data;
        d.sub = nullptr;
        d.post = nullptr;
        d.tw = twp + j.limb_begin;
        d.lc = lc + j.limb_begin;
        d.limbs = j.limbs;
        d.period = j.limbs;
        d.item_begin = j.item_begin;
        d.n = (uint16_t)j.logN;
        d.inverse = (uint16_t)j.inverse;
        reinterpret_cast<FusedJob*>(a.base + OFF_DESC)[j.desc] = d;
    }
}

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;   // device -> address of its `scratch`

// The device address of this device's array, looked up once per device.
int scratch_for_current_device(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return record_hip_error(e, "hipGetDevice");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return record_hip_error(e, "hipGetSymbolAddress(scratch)");
        it = g_scratch_ptr.emplace(dev, static_cast<unsigned char*>(p)).first;
    }
    *out = it->second;
    return NTT_OK;
}

int ck(hipError_t e, const char* where) { return e == hipSuccess ? NTT_OK : record_hip_error(e, where); }

}  // namespace

extern "C" {

int ntt_batch_run_static(const ntt_table_job* jobs, size_t njobs, void* stream) {
    if (!jobs || njobs == 0) return NTT_ERR_INVALID_ARGUMENT;
    if (njobs > MAX_JOBS) return NTT_ERR_SCRATCH_LIMIT;
    size_t total_limbs = 0;
    for (size_t i = 0; i < njobs; ++i) {
        const ntt_table_job& j = jobs[i];
        if (!j.data || !j.twiddles || !j.moduli || !j.q_inv || !j.scale || j.limbs == 0) return NTT_ERR_INVALID_ARGUMENT;
        if (j.logN < 1 || j.logN > NTT_MAX_LOG_N) return NTT_ERR_INVALID_ARGUMENT;
        if (j.direction != NTT_FORWARD && j.direction != NTT_INVERSE) return NTT_ERR_INVALID_ARGUMENT;
        if (reinterpret_cast<uintptr_t>(j.data) & 15) return NTT_ERR_INVALID_ARGUMENT;   // 16 B vector access
        total_limbs += j.limbs;
        if (total_limbs > MAX_LIMBS) return NTT_ERR_SCRATCH_LIMIT;
    }
    TraceRange range("ntt_batch_run_static");
    unsigned char* base = nullptr;
    int rc = scratch_for_current_device(&base);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);

    // Group by (logN, direction): fused groups own a contiguous run of
    // descriptors, two-pass groups a contiguous run of per-limb entries.
    std::map<std::pair<int, int>, std::vector<size_t>> groups;
    for (size_t i = 0; i < njobs; ++i) groups[{jobs[i].logN, jobs[i].direction}].push_back(i);
    std::vector<StaticJob> sj(njobs);
    struct Launch {
        int n;
        bool inverse, fused;
        uint32_t begin, count, items;   // fused: descriptors; two-pass: limbs
    };
    std::vector<Launch> launches;
    uint32_t desc = 0, limb = 0;
    for (auto& kv : groups) {
        const int n = kv.first.first;
        const bool inv = kv.first.second == NTT_INVERSE;
        const bool fused = fused_plan_supported(n);   // logN 17: the split kernels
        Launch L{n, inv, fused, fused ? desc : limb, 0, 0};
        uint64_t items = 0;
        for (size_t i : kv.second) {
            const ntt_table_job& j = jobs[i];
            StaticJob& t = sj[i];
            t = StaticJob{j.data, j.twiddles, j.moduli, j.q_inv, j.scale, j.limbs, limb,
                          fused ? desc : NO_DESC, (uint32_t)items, n, inv ? 1 : 0};
            limb += j.limbs;
            if (fused) {
                ++desc;
                ++L.count;
                items += fused_items(n, j.limbs);
                if (items > 0xffffffffull) return NTT_ERR_INVALID_ARGUMENT;
            } else {
                L.count += j.limbs;
            }
        }
        L.items = (uint32_t)items;
        launches.push_back(L);
    }

    // 1. the job tables, SETUP_JOBS jobs per setup launch
    for (size_t first = 0; first < njobs; first += SETUP_JOBS) {
        SetupArgs a{};
        a.base = base;
        const size_t cnt = std::min<size_t>(SETUP_JOBS, njobs - first);
        for (size_t k = 0; k < cnt; ++k) a.job[k] = sj[first + k];
        hipLaunchKernelGGL(static_setup_kernel, dim3((unsigned)cnt), dim3(256), 0, s, a);
        if ((rc = ck(hipGetLastError(), "static setup kernel launch"))) return rc;
    }
    // 2. the transforms, in order on the same stream
    for (const Launch& L : launches) {
        if (L.fused) {
            FusedParams p{};
            p.jobs = reinterpret_cast<const FusedJob*>(base + OFF_DESC) + L.begin;
            p.njobs = L.count;
            p.total_items = L.items;
            p.counter = nullptr;
            p.stagger = 0;
            // canonical arithmetic: the moduli are only on the device, so the
            // lazy kernels' q < 2^30 condition cannot be checked here
            rc = launch_fused(L.n, L.inverse, FUSED_DEFAULT_EB_INV, false, p, s, 0);
        } else {
            TransformLaunch T{};
            T.A_ptrs = reinterpret_cast<uint32_t* const*>(base + OFF_AP) + L.begin;
            T.tw_ptrs = reinterpret_cast<const uint32_t* const*>(base + OFF_TWP) + L.begin;
            T.moduli = reinterpret_cast<const uint32_t*>(base + OFF_Q) + L.begin;
            T.q_inv = reinterpret_cast<const uint32_t*>(base + OFF_QI) + L.begin;
            T.scale = reinterpret_cast<const uint32_t*>(base + OFF_SC) + L.begin;
            // the scale words of the reference's launchers: R^2 forward,
            // to_mont(N^-1) inverse (v04:603-658)
            T.scale_mode = L.inverse ? SCALE_FROM_MONT : SCALE_AS_IS;
            T.logN = L.n;
            T.limbs = L.count;
            T.count = 1;
            T.inverse = L.inverse;
            T.lazy = false;
            T.xcd_remap = 1;
            default_config(L.n, &T.cfg);
            rc = launch_transform(T, s);
        }
        if (rc) return rc;
    }
    return NTT_OK;
}

}  // extern "C"
