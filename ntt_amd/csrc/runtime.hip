// Thin device-runtime helpers (include/ntt_amd_runtime.h) over the HIP
// runtime this library links.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "../../include/ntt_amd.h"
#include "../../include/ntt_amd_runtime.h"
#include "launch.hpp"

namespace {

inline int ck(hipError_t e, const char* where) {
    return e == hipSuccess ? NTT_OK : ntt::record_hip_error(e, where);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// One word per thread, grid-stride; value = floor(hash * q / 2^64).
__global__ void fill_uniform_kernel(uint32_t* dst, const uint32_t* moduli, uint32_t limbs, int logN,
                                    size_t total, uint64_t seed) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t limb = (uint32_t)((i >> logN) % limbs);
        const uint64_t h = splitmix64(seed ^ splitmix64(i));
        dst[i] = (uint32_t)__umul64hi(h, (uint64_t)moduli[limb]);
    }
}

// Many fills in one launch: workgroup b fills words [b * FILL_CHUNK, ...) of
// the concatenation of all jobs (job j starts at word_begin[j]); each word
// gets exactly the value fill_uniform_kernel gives it within its own job.
__global__ void fill_uniform_jobs_kernel(const ntt::FillJob* jobs, uint32_t njobs, uint64_t total) {
    const uint64_t begin = (uint64_t)blockIdx.x * ntt::FILL_CHUNK;
    if (begin >= total) return;
    const uint64_t end = begin + ntt::FILL_CHUNK < total ? begin + ntt::FILL_CHUNK : total;
    uint32_t lo = 0, hi = njobs;   // last job with word_begin <= begin
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].word_begin <= begin) lo = mid; else hi = mid;
    }
    for (uint64_t w = begin + threadIdx.x; w < end; w += blockDim.x) {
        uint32_t j = lo;
        while (j + 1 < njobs && jobs[j + 1].word_begin <= w) ++j;
        const ntt::FillJob& f = jobs[j];
        const uint64_t i = w - f.word_begin;
        const uint32_t limb = (uint32_t)((i >> f.logN) % f.limbs);
        const uint64_t h = splitmix64(f.seed ^ splitmix64(i));
        f.dst[i] = (uint32_t)__umul64hi(h, (uint64_t)f.moduli[limb]);
    }
}

}  // namespace

namespace ntt {
int fused_prepare();
int launch_fill_jobs(const FillJob* d_jobs, uint32_t njobs, uint64_t total_words, hipStream_t stream) {
    if (!njobs || !total_words) return NTT_OK;
    const uint64_t blocks = (total_words + FILL_CHUNK - 1) / FILL_CHUNK;
    if (blocks > 0x7fffffffull) return NTT_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(fill_uniform_jobs_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_jobs, njobs,
                       total_words);
    return ck(hipGetLastError(), "fill_uniform_jobs launch");
}
}  // namespace ntt

extern "C" {

int ntt_rt_device_count(int* count) {
    if (!count) return NTT_ERR_INVALID_ARGUMENT;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    return e == hipSuccess ? NTT_OK : ck(e, "hipGetDeviceCount");
}

int ntt_rt_set_device(int device) { return ck(hipSetDevice(device), "hipSetDevice"); }

int ntt_rt_device_info(int device, ntt_device_info* out) {
    if (!out) return NTT_ERR_INVALID_ARGUMENT;
    hipDeviceProp_t p;
    int rc = ck(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    if (rc) return rc;
    std::memset(out, 0, sizeof *out);
    std::snprintf(out->name, sizeof out->name, "%s", p.name);
    std::snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    out->compute_units = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    out->l2_bytes = p.l2CacheSize;
    out->lds_per_block = p.sharedMemPerBlock;
    out->total_global_mem = p.totalGlobalMem;
    return NTT_OK;
}

int ntt_rt_malloc(void** ptr, size_t bytes) {
    if (!ptr) return NTT_ERR_INVALID_ARGUMENT;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory) { (void)hipGetLastError(); return NTT_ERR_OUT_OF_MEMORY; }
    return ck(e, "hipMalloc");
}

int ntt_rt_free(void* ptr) { return ck(hipFree(ptr), "hipFree"); }

int ntt_rt_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return ck(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int ntt_rt_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return ck(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

int ntt_rt_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    return ck(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync D2D");
}

int ntt_rt_memset(void* ptr, int value, size_t bytes) { return ck(hipMemset(ptr, value, bytes), "hipMemset"); }

int ntt_rt_synchronize(void) { return ck(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int ntt_rt_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return NTT_ERR_INVALID_ARGUMENT;
    return ck(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int ntt_rt_stream_create(void** stream) {
    if (!stream) return NTT_ERR_INVALID_ARGUMENT;
    return ck(hipStreamCreateWithFlags(reinterpret_cast<hipStream_t*>(stream), hipStreamNonBlocking),
              "hipStreamCreate");
}

int ntt_rt_stream_destroy(void* stream) { return ck(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy"); }

int ntt_rt_stream_synchronize(void* stream) {
    return ck(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int ntt_rt_event_create(void** event) {
    if (!event) return NTT_ERR_INVALID_ARGUMENT;
    return ck(hipEventCreate(reinterpret_cast<hipEvent_t*>(event)), "hipEventCreate");
}

int ntt_rt_event_destroy(void* event) { return ck(hipEventDestroy((hipEvent_t)event), "hipEventDestroy"); }

int ntt_rt_event_record(void* event, void* stream) {
    return ck(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

int ntt_rt_event_synchronize(void* event) {
    return ck(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

int ntt_rt_event_elapsed_ms(float* ms, void* start, void* stop) {
    if (!ms) return NTT_ERR_INVALID_ARGUMENT;
    return ck(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

namespace {
struct CapturedGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};
}  // namespace

int ntt_rt_graph_begin(void* stream) {
    if (!stream) return NTT_ERR_INVALID_ARGUMENT;
    // kernel attributes are set once per process before any capture
    int rc = ntt::fused_prepare();
    if (rc) return rc;
    return ck(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
}

int ntt_rt_graph_end(void* stream, void** graph) {
    if (!stream || !graph) return NTT_ERR_INVALID_ARGUMENT;
    *graph = nullptr;
    CapturedGraph* g = new CapturedGraph();
    int rc = ck(hipStreamEndCapture((hipStream_t)stream, &g->graph), "hipStreamEndCapture");
    if (rc == NTT_OK) rc = ck(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    *graph = g;
    return NTT_OK;
}

int ntt_rt_graph_launch(void* graph, void* stream) {
    if (!graph) return NTT_ERR_INVALID_ARGUMENT;
    return ck(hipGraphLaunch(static_cast<CapturedGraph*>(graph)->exec, (hipStream_t)stream), "hipGraphLaunch");
}

int ntt_rt_graph_destroy(void* graph) {
    if (!graph) return NTT_ERR_INVALID_ARGUMENT;
    CapturedGraph* g = static_cast<CapturedGraph*>(graph);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return NTT_OK;
}

int ntt_rt_fill_uniform(uint32_t* dst, const uint32_t* moduli_dev, size_t limbs, int logN, size_t count,
                        uint64_t seed, void* stream) {
    if (!dst || !moduli_dev || limbs == 0 || count == 0 || logN < 1 || logN > 30) return NTT_ERR_INVALID_ARGUMENT;
    const size_t total = (limbs * count) << logN;
    const unsigned threads = 256;
    size_t blocks = (total + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3((unsigned)blocks), dim3(threads), 0, (hipStream_t)stream, dst,
                       moduli_dev, (uint32_t)limbs, logN, total, seed);
    return ck(hipGetLastError(), "fill_uniform launch");
}

}  // extern "C"
