// Roof measurements used to place the transform kernels (ntt_amd_runtime.h):
//   * the butterfly roof: how many butterflies per second the CUs sustain
//     with the transform's own butterfly code (arith.hpp) when operands sit
//     in registers, no memory or LDS traffic at all — the binding compute
//     roof of an NTT whose 32-bit modular multiplies run on the integer VALU;
//   * a plain streaming copy of known size, the calibration kernel for the
//     HBM byte counters (FETCH_SIZE / WRITE_SIZE) and the achievable-bandwidth
//     reference.
#include <hip/hip_runtime.h>

#include "../../include/ntt_amd.h"
#include "../../include/ntt_amd_runtime.h"
#include "arith.hpp"
#include "arith64.hpp"
#include "launch.hpp"

namespace {

constexpr int ROOF_CHAINS = 8;       // independent butterflies in flight per thread
constexpr int ROOF_ITERS = 2048;
constexpr unsigned ROOF_THREADS = 256;

inline int ck(hipError_t e, const char* where) {
    return e == hipSuccess ? NTT_OK : ntt::record_hip_error(e, where);
}

// Each thread runs ROOF_CHAINS dependent butterfly chains; every butterfly
// feeds the next, so nothing folds away, and the twiddle changes per step.
template <int VARIANT>
__global__ __launch_bounds__(ROOF_THREADS) void butterfly_roof_kernel(uint32_t* out, uint32_t q, uint32_t q_inv,
                                                                      uint32_t w0) {
    uint32_t X[ROOF_CHAINS], Y[ROOF_CHAINS];
    const uint32_t t = blockIdx.x * ROOF_THREADS + threadIdx.x;
#pragma unroll
    for (int i = 0; i < ROOF_CHAINS; ++i) {
        X[i] = (t * 7u + (uint32_t)i) % q;
        Y[i] = (t * 13u + 5u * (uint32_t)i) % q;
    }
    uint32_t w = w0;
    for (int it = 0; it < ROOF_ITERS; ++it) {
#pragma unroll
        for (int i = 0; i < ROOF_CHAINS; ++i) {
            if (VARIANT == 0) ntt::ct_butterfly(X[i], Y[i], w, q, q_inv);
            if (VARIANT == 1) ntt::ct_butterfly_lazy(X[i], Y[i], w, q, q_inv);
            if (VARIANT == 2) ntt::gs_butterfly(X[i], Y[i], w, q, q_inv);
            if (VARIANT == 3) ntt::gs_butterfly_lazy(X[i], Y[i], w, q, q_inv);
        }
        w = w + 1u < q ? w + 1u : 1u;
    }
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < ROOF_CHAINS; ++i) r ^= X[i] ^ Y[i];
    out[t] = r;
}

// Cooley-Tukey on x, y in [0, 4q) with a precomputed-quotient (Shoup /
// Harvey) product instead of Montgomery's: with w' = floor(w 2^32 / q),
// v = y w - floor(y w' / 2^32) q lies in [0, 2q) for every 32-bit y.  The
// product is a high multiply, a low multiply and a multiply-add (three VALU
// instructions, as the Montgomery product), and every twiddle needs its
// quotient word w' beside it.
__device__ __forceinline__ void ct_butterfly_shoup_lazy(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wq,
                                                        uint32_t q) {
    const uint32_t q2 = 2 * q;
    const uint32_t u = min(x, x - q2);
    const uint32_t v = y * w - __umulhi(y, wq) * q;
    x = u + v;
    y = u + q2 - v;
}

// Variants 6 (Shoup) and 7 (Montgomery) of the lazy Cooley-Tukey butterfly
// with the same loop: eight twiddles per thread (with their quotient words
// for variant 6) held in registers and used in turn, so that the two
// products are compared on equal terms.
template <int VARIANT>
__global__ __launch_bounds__(ROOF_THREADS) void butterfly_roof_pair_kernel(uint32_t* out, uint32_t q, uint32_t q_inv,
                                                                           uint32_t w0) {
    uint32_t X[ROOF_CHAINS], Y[ROOF_CHAINS], W[8], WQ[8];
    const uint32_t t = blockIdx.x * ROOF_THREADS + threadIdx.x;
#pragma unroll
    for (int i = 0; i < ROOF_CHAINS; ++i) {
        X[i] = (t * 7u + (uint32_t)i) % q;
        Y[i] = (t * 13u + 5u * (uint32_t)i) % q;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        W[k] = (w0 + t + 97u * (uint32_t)k) % q;
        WQ[k] = (uint32_t)(((uint64_t)W[k] << 32) / q);
    }
    for (int it = 0; it < ROOF_ITERS; it += 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
            for (int i = 0; i < ROOF_CHAINS; ++i) {
                if (VARIANT == 6) ct_butterfly_shoup_lazy(X[i], Y[i], W[k], WQ[k], q);
                if (VARIANT == 7) ntt::ct_butterfly_lazy(X[i], Y[i], W[k], q, q_inv);
            }
    }
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < ROOF_CHAINS; ++i) r ^= X[i] ^ Y[i];
    out[t] = r;
}

// The same for 64-bit words (variant 4: Cooley-Tukey, 5: Gentleman-Sande).
template <int VARIANT>
__global__ __launch_bounds__(ROOF_THREADS) void butterfly_roof64_kernel(uint32_t* out, uint64_t q, uint64_t q_inv,
                                                                        uint64_t w0) {
    uint64_t X[ROOF_CHAINS], Y[ROOF_CHAINS];
    const uint32_t t = blockIdx.x * ROOF_THREADS + threadIdx.x;
#pragma unroll
    for (int i = 0; i < ROOF_CHAINS; ++i) {
        X[i] = ((uint64_t)t * 7u + (uint64_t)i) % q;
        Y[i] = ((uint64_t)t * 13u + 5u * (uint64_t)i) % q;
    }
    uint64_t w = w0;
    for (int it = 0; it < ROOF_ITERS / 4; ++it) {
#pragma unroll
        for (int i = 0; i < ROOF_CHAINS; ++i) {
            if (VARIANT == 4) ntt::ct_butterfly64(X[i], Y[i], w, q, q_inv);
            if (VARIANT == 5) ntt::gs_butterfly64(X[i], Y[i], w, q, q_inv);
        }
        w = w + 1u < q ? w + 1u : 1u;
    }
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < ROOF_CHAINS; ++i) r ^= X[i] ^ Y[i];
    out[t] = (uint32_t)(r ^ (r >> 32));
}

__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                          size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" {

int ntt_rt_butterfly_roof(int variant, double* butterflies_per_s, void* stream) {
    if (!butterflies_per_s || variant < 0 || variant > 7) return NTT_ERR_INVALID_ARGUMENT;
    int dev = 0, cus = 0;
    int rc = ck(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if ((rc = ck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute")))
        return rc;
    // 8 workgroups of 4 waves per CU: 8 waves per SIMD, enough to hide the
    // multiply latency.
    const unsigned blocks = (unsigned)cus * 8u;
    uint32_t* out = nullptr;
    if ((rc = ck(hipMalloc(&out, (size_t)blocks * ROOF_THREADS * 4), "hipMalloc"))) return rc;
    void (*fn)(uint32_t*, uint32_t, uint32_t, uint32_t) =
        variant == 0 ? butterfly_roof_kernel<0>
        : variant == 1 ? butterfly_roof_kernel<1>
        : variant == 2 ? butterfly_roof_kernel<2>
        : variant == 3 ? butterfly_roof_kernel<3>
        : variant == 6 ? butterfly_roof_pair_kernel<6>
                       : butterfly_roof_pair_kernel<7>;
    const uint32_t q = 1073479681u;     // a prime < 2^30 with q = 1 mod 2^17
    uint32_t inv = 1;
    for (int i = 0; i < 5; ++i) inv *= 2u - q * inv;
    const uint32_t q_inv = 0u - inv;
    const uint64_t q64 = 4611686018425815041ull;   // the largest prime < 2^62 with q = 1 mod 2^18
    uint64_t inv64 = 1;
    for (int i = 0; i < 6; ++i) inv64 *= 2u - q64 * inv64;
    const uint64_t q_inv64 = 0u - inv64;
    const bool w64 = variant == 4 || variant == 5;
    void (*fn64)(uint32_t*, uint64_t, uint64_t, uint64_t) =
        variant == 4 ? butterfly_roof64_kernel<4> : butterfly_roof64_kernel<5>;
    auto launch = [&](uint32_t salt) {
        if (w64)
            hipLaunchKernelGGL(fn64, dim3(blocks), dim3(ROOF_THREADS), 0, (hipStream_t)stream, out, q64, q_inv64,
                               (uint64_t)salt * 0x9E3779B97F4A7C15ull % q64);
        else
            hipLaunchKernelGGL(fn, dim3(blocks), dim3(ROOF_THREADS), 0, (hipStream_t)stream, out, q, q_inv, salt);
    };
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = ck(hipEventCreate(&e0), "hipEventCreate");
    if (!rc) rc = ck(hipEventCreate(&e1), "hipEventCreate");
    const int reps = 5;
    if (!rc) {
        launch(12345u);   // warm-up
        rc = ck(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < reps && !rc; ++r) {
            launch(777u + (uint32_t)r);
            rc = ck(hipGetLastError(), "butterfly roof launch");
        }
        if (!rc) rc = ck(hipEventRecord(e1, s), "hipEventRecord");
        if (!rc) rc = ck(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        if (!rc) rc = ck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (!rc) {
            const double bfly = (double)reps * blocks * ROOF_THREADS * ROOF_CHAINS * (w64 ? ROOF_ITERS / 4 : ROOF_ITERS);
            *butterflies_per_s = ms > 0 ? bfly / (ms * 1e-3) : 0.0;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(out);
    return rc;
}

int ntt_rt_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return NTT_ERR_INVALID_ARGUMENT;
    int dev = 0, cus = 0;
    int rc = ck(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if ((rc = ck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute")))
        return rc;
    const size_t n16 = bytes / 16;
    size_t blocks = (n16 + 255) / 256;
    const size_t cap = (size_t)cus * 4;     // 4 workgroups per CU, grid-stride: measured best for large copies
    if (blocks > cap) blocks = cap;
    if (blocks == 0) return NTT_OK;
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)src, (uint4*)dst, n16);
    return ck(hipGetLastError(), "stream copy launch");
}

}  // extern "C"
