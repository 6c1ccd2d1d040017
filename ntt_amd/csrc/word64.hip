// 64-bit word transforms: the negacyclic NTT / INTT over primes q < 2^62,
// the word width of Cheddar's NTTHandler<uint64_t> (src/NTT_cheddar.cu:1313;
// the reference's own kernels are 32-bit only).
//
// Same data layout and orders as the 32-bit path (A[limb * N + i]; forward:
// natural-order input, bit-reversed output), normal domain in and out:
// twiddles are stored in Montgomery form (R = 2^64), so mont_mul(y, w) is the
// plain product y * psi^e mod q, and the inverse ends with mont_mul(x,
// to_mont(N^-1)).  Two passes like the 32-bit two-pass engine: a column pass
// over the top R1 index bits and a row pass over the low R2 bits; each pass
// stages a tile in LDS and runs its stages three at a time in registers
// (groups of 8 words), one barrier per three stages.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "trace.hpp"
#include "../../include/ntt_amd.h"
#include "arith64.hpp"
#include "context.hpp"
#include "host_math.hpp"
#include "launch.hpp"

namespace ntt {

typedef __attribute__((address_space(1))) uint64_t gu64;
typedef __attribute__((address_space(1))) const uint64_t gcu64;

struct Pass64Params {
    uint64_t* A;                       // operand k, limb l at A + (k * limbs + l) * N
    const uint64_t* const* tw;         // per-limb Montgomery-form table (this direction)
    const uint64_t* q;                 // per limb
    const uint64_t* q_inv;             // -q^-1 mod 2^64 per limb
    const uint64_t* scale;             // per-limb Montgomery multiplier on the last store, or null
    int logN, R, W;                    // pass stages; log2 tile columns (column pass) / rows (row pass)
    uint32_t tiles, limbs, count;
};

constexpr int PASS64_THREADS = 256;

namespace w64 {

// One round of K stages on pass bits [lb, lb + K) over a tile in LDS.  A
// thread takes groups of 2^K words whose pass indices differ only in those
// bits (group member m at x0 + (m << lb)), runs the K stages in registers and
// writes the group back.  For a word at pass index x, (word index) >> (b + 1)
// is x >> (b + 1) in a column pass and ((row << R) + x) >> (b + 1) in a row
// pass; for group members it is that of x0 plus m >> (kb + 1) (x0 is zero on
// the group bits, so nothing carries).
template <int K, bool INV, bool COL>
__device__ __forceinline__ void group_round(uint64_t* L64, gcu64* tw, uint64_t q, uint64_t q_inv, int n, int R, int W,
                                            uint32_t tile, int lb) {
    constexpr int G = 1 << K;
    const uint32_t words = 1u << (R + W), mask = (1u << W) - 1, lo_mask = (1u << lb) - 1;
    for (uint32_t gi = threadIdx.x; gi < words / G; gi += PASS64_THREADS) {
        const uint32_t lp = COL ? (gi & mask) : gi >> (R - K);
        const uint32_t xp = COL ? gi >> W : gi & ((1u << (R - K)) - 1);
        const uint32_t x0 = ((xp >> lb) << (lb + K)) | (xp & lo_mask);
        const uint32_t base = COL ? x0 : (((tile << W) + lp) << R) + x0;   // (word index) above, before shifting
        uint32_t e[G];
        uint64_t a[G];
#pragma unroll
        for (int m = 0; m < G; ++m) {
            const uint32_t x = x0 + ((uint32_t)m << lb);
            e[m] = COL ? (x << W) | lp : (lp << R) | x;
            a[m] = L64[e[m]];
        }
#pragma unroll
        for (int st = 0; st < K; ++st) {
            const int kb = INV ? st : K - 1 - st;          // group bit of this stage
            const int b = lb + kb;                         // pass bit
            const int bg = COL ? n - R + b : b;            // word-index bit
            const uint32_t tbase = (1u << (n - 1 - bg)) + (base >> (b + 1));
#pragma unroll
            for (int m = 0; m < G; ++m) {
                if (m & (1 << kb)) continue;
                const uint64_t w = tw[tbase + ((uint32_t)m >> (kb + 1))];
                if (INV)
                    gs_butterfly64(a[m], a[m | (1 << kb)], w, q, q_inv);
                else
                    ct_butterfly64(a[m], a[m | (1 << kb)], w, q, q_inv);
            }
        }
#pragma unroll
        for (int m = 0; m < G; ++m) L64[e[m]] = a[m];
    }
}

}  // namespace w64

// COL: the top R bits of the index, tiles of 2^W consecutive columns of the
// 2^R x 2^(n-R) matrix; row pass: the low R bits, tiles of 2^W rows.
template <bool INV, bool COL>
__global__ __launch_bounds__(PASS64_THREADS) void pass64_kernel(Pass64Params p) {
    extern __shared__ uint64_t L64[];
    const uint32_t b = blockIdx.x;
    const uint32_t tile = b % p.tiles, rest = b / p.tiles;
    const uint32_t op = rest % p.count, li = rest / p.count;
    const int n = p.logN, R = p.R, W = p.W;
    const size_t N = (size_t)1 << n;
    gu64* A = (gu64*)(p.A + ((size_t)op * p.limbs + li) * N);
    gcu64* tw = (gcu64*)p.tw[li];
    const uint64_t q = p.q[li], q_inv = p.q_inv[li];
    const uint32_t words = 1u << (R + W), mask = (1u << W) - 1;
    const uint32_t C = 1u << (n - R);   // COL: columns of the matrix
    // tile element e -> word index in the limb
    auto gidx = [&](uint32_t e) -> uint32_t {
        return COL ? (e >> W) * C + (tile << W) + (e & mask) : (tile << (R + W)) + e;
    };
    for (uint32_t e = threadIdx.x; e < words; e += PASS64_THREADS) L64[e] = A[gidx(e)];
    __syncthreads();
    // Stages K = 3 at a time (a group of 8 words over pass bits lb .. lb + 2:
    // one LDS round trip and one barrier per three stages, 12 butterflies and 7
    // twiddle loads per group); a last round of 1 or 2 stages when 3 does not
    // divide R.  Forward: bits from the top down; inverse: from the bottom up.
    for (int s = 0; s < R;) {
        const int K = R - s >= 3 ? 3 : R - s;
        const int lb = INV ? s : R - s - K;
        if (K == 3)
            w64::group_round<3, INV, COL>(L64, tw, q, q_inv, n, R, W, tile, lb);
        else if (K == 2)
            w64::group_round<2, INV, COL>(L64, tw, q, q_inv, n, R, W, tile, lb);
        else
            w64::group_round<1, INV, COL>(L64, tw, q, q_inv, n, R, W, tile, lb);
        s += K;
        __syncthreads();
    }
    const uint64_t sc = p.scale ? p.scale[li] : 0;
    for (uint32_t e = threadIdx.x; e < words; e += PASS64_THREADS) {
        uint64_t x = L64[e];
        if (p.scale) x = mont_mul64(x, sc, q, q_inv);
        A[gidx(e)] = x;
    }
}

// ---- host arithmetic (unsigned __int128) ------------------------------------
typedef unsigned __int128 u128;
static uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }
static uint64_t powmod64(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    for (; e; e >>= 1, b = mulmod64(b, b, q))
        if (e & 1) r = mulmod64(r, b, q);
    return r;
}
static uint64_t neg_inv64(uint64_t q) {   // -q^-1 mod 2^64, Newton: 1 -> 64 bits in 6 steps
    uint64_t inv = 1;
    for (int i = 0; i < 6; ++i) inv *= 2 - q * inv;
    return 0 - inv;
}
static uint64_t to_mont64(uint64_t x, uint64_t q) { return (uint64_t)(((u128)x << 64) % q); }
static bool is_prime64(uint64_t n) {   // deterministic Miller-Rabin for n < 3.3e24
    if (n < 2) return false;
    static const uint64_t P[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t p : P)
        if (n % p == 0) return n == p;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, ++s;
    for (uint64_t a : P) {
        uint64_t x = powmod64(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < s && comp; ++i) {
            x = mulmod64(x, x, n);
            if (x == n - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}
static uint32_t brv(uint32_t v, int w) {
    uint32_t r = 0;
    for (int i = 0; i < w; ++i) r |= ((v >> i) & 1u) << (w - 1 - i);
    return r;
}

}  // namespace ntt

struct ntt_plan64 {
    int logN = 0;
    size_t limbs = 0;
    int R1 = 0, R2 = 0;
    void* d_mem = nullptr;                 // tables + per-limb arrays, one allocation
    const uint64_t* const* d_tw_fwd = nullptr;
    const uint64_t* const* d_tw_inv = nullptr;
    uint64_t *d_q = nullptr, *d_q_inv = nullptr, *d_ninv = nullptr;
};

using namespace ntt;

static int launch_pass64(const ntt_plan64* P, uint64_t* A, size_t count, bool inv, bool col, hipStream_t stream) {
    const int R = col ? P->R1 : P->R2;
    if (R == 0) return NTT_OK;
    const int other = col ? P->R2 : P->R1;   // bits across tiles
    const int W = col ? (other < 4 ? other : 4) : (other < 2 ? other : 2);
    Pass64Params p{};
    p.A = A;
    p.tw = inv ? P->d_tw_inv : P->d_tw_fwd;
    p.q = P->d_q;
    p.q_inv = P->d_q_inv;
    // the inverse ends in the column pass (or the row pass when there is none)
    p.scale = (inv && (col || P->R1 == 0)) ? P->d_ninv : nullptr;
    p.logN = P->logN;
    p.R = R;
    p.W = W;
    p.tiles = 1u << (other - W);
    p.limbs = (uint32_t)P->limbs;
    p.count = (uint32_t)count;
    const size_t blocks = (size_t)p.tiles * P->limbs * count;
    if (blocks == 0 || blocks > 0x7fffffffu) return NTT_ERR_INVALID_ARGUMENT;
    const size_t lds = ((size_t)1 << (R + W)) * sizeof(uint64_t);
    void (*fn)(Pass64Params) = inv ? (col ? pass64_kernel<true, true> : pass64_kernel<true, false>)
                                   : (col ? pass64_kernel<false, true> : pass64_kernel<false, false>);
    hipError_t stale = hipGetLastError();
    if (stale != hipSuccess) record_hip_error(stale, "error pending before launch (cleared)");
    hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(PASS64_THREADS), lds, stream, p);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? NTT_OK : record_hip_error(e, "64-bit pass launch");
}

extern "C" {

int ntt_generate_moduli64(int logN, uint64_t bound, int count, ntt_modulus_config64* out) {
    if (logN < 1 || logN > NTT_MAX_LOG_N || count < 0 || (count && !out)) return -NTT_ERR_INVALID_ARGUMENT;
    if (bound > ((uint64_t)1 << 62)) bound = (uint64_t)1 << 62;
    const uint64_t step = (uint64_t)2 << logN;
    if (bound <= step + 1) return 0;
    uint64_t q = (bound - 2) / step * step + 1;   // largest q = 1 mod 2N below bound
    int found = 0;
    for (; found < count && q > step; q -= step) {
        if (!is_prime64(q)) continue;
        uint64_t g = 2;
        while (powmod64(g, (q - 1) / 2, q) != q - 1) ++g;   // a quadratic non-residue
        out[found].q = q;
        out[found].root = g;
        ++found;
    }
    return found;
}

int ntt_plan64_create(ntt_context* ctx, int logN, const ntt_modulus_config64* moduli, size_t limbs,
                      ntt_plan64** out) {
    if (!ctx || !moduli || !out || limbs == 0 || limbs > 0xffffffu || logN < 2 || logN > 17)
        return NTT_ERR_INVALID_ARGUMENT;
    const size_t N = (size_t)1 << logN;
    std::vector<uint64_t> tf(limbs * N), ti(limbs * N), q(limbs), qi(limbs), ninv(limbs);
    for (size_t l = 0; l < limbs; ++l) {
        const uint64_t Q = moduli[l].q;
        if (Q < 3 || !(Q & 1) || Q >= ((uint64_t)1 << 62)) return NTT_ERR_MODULUS_RANGE;
        if ((Q - 1) % ((uint64_t)2 << logN)) return NTT_ERR_MODULUS_CONGRUENCE;
        const uint64_t psi = powmod64(moduli[l].root, (Q - 1) / ((uint64_t)2 << logN), Q);
        if (powmod64(psi, N, Q) != Q - 1) return NTT_ERR_PRIMITIVE_ROOT;
        const uint64_t psi_inv = powmod64(psi, Q - 2, Q);
        // powers in natural order, then scattered to bit-reversed positions
        uint64_t pf = 1, pi = 1;
        for (size_t i = 0; i < N; ++i) {
            const uint32_t j = brv((uint32_t)i, logN);
            tf[l * N + j] = to_mont64(pf, Q);
            ti[l * N + j] = to_mont64(pi, Q);
            pf = mulmod64(pf, psi, Q);
            pi = mulmod64(pi, psi_inv, Q);
        }
        q[l] = Q;
        qi[l] = neg_inv64(Q);
        ninv[l] = to_mont64(powmod64(N % Q, Q - 2, Q), Q);
    }
    ntt_plan64* P = new ntt_plan64();
    P->logN = logN;
    P->limbs = limbs;
    P->R2 = logN < 9 ? logN : 9;
    P->R1 = logN - P->R2;
    const size_t tb = limbs * N * sizeof(uint64_t), pb = limbs * sizeof(void*), ab = limbs * sizeof(uint64_t);
    const size_t total = 2 * tb + 2 * pb + 3 * ab;
    if (hipMalloc(&P->d_mem, total) != hipSuccess) {
        delete P;
        return NTT_ERR_OUT_OF_MEMORY;
    }
    char* base = static_cast<char*>(P->d_mem);
    uint64_t* dtf = reinterpret_cast<uint64_t*>(base);
    uint64_t* dti = reinterpret_cast<uint64_t*>(base + tb);
    P->d_tw_fwd = reinterpret_cast<const uint64_t* const*>(base + 2 * tb);
    P->d_tw_inv = reinterpret_cast<const uint64_t* const*>(base + 2 * tb + pb);
    P->d_q = reinterpret_cast<uint64_t*>(base + 2 * tb + 2 * pb);
    P->d_q_inv = P->d_q + limbs;
    P->d_ninv = P->d_q_inv + limbs;
    std::vector<const uint64_t*> pf(limbs), pi(limbs);
    for (size_t l = 0; l < limbs; ++l) {
        pf[l] = dtf + l * N;
        pi[l] = dti + l * N;
    }
    std::vector<char> h(total);
    std::memcpy(h.data(), tf.data(), tb);
    std::memcpy(h.data() + tb, ti.data(), tb);
    std::memcpy(h.data() + 2 * tb, pf.data(), pb);
    std::memcpy(h.data() + 2 * tb + pb, pi.data(), pb);
    std::memcpy(h.data() + 2 * tb + 2 * pb, q.data(), ab);
    std::memcpy(h.data() + 2 * tb + 2 * pb + ab, qi.data(), ab);
    std::memcpy(h.data() + 2 * tb + 2 * pb + 2 * ab, ninv.data(), ab);
    if (hipMemcpy(P->d_mem, h.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(P->d_mem);
        delete P;
        return NTT_ERR_HIP;
    }
    *out = P;
    return NTT_OK;
}

int ntt_plan64_destroy(ntt_plan64* plan) {
    if (!plan) return NTT_ERR_INVALID_ARGUMENT;
    if (plan->d_mem) (void)hipFree(plan->d_mem);
    delete plan;
    return NTT_OK;
}

int ntt_plan64_forward(ntt_plan64* plan, uint64_t* A_dev, size_t count, void* stream) {
    TraceRange range("ntt64_forward");
    if (!plan || !A_dev || count == 0) return NTT_ERR_INVALID_ARGUMENT;
    int rc = launch_pass64(plan, A_dev, count, false, true, (hipStream_t)stream);
    if (rc == NTT_OK) rc = launch_pass64(plan, A_dev, count, false, false, (hipStream_t)stream);
    return rc;
}

int ntt_plan64_inverse(ntt_plan64* plan, uint64_t* A_dev, size_t count, void* stream) {
    TraceRange range("ntt64_inverse");
    if (!plan || !A_dev || count == 0) return NTT_ERR_INVALID_ARGUMENT;
    int rc = launch_pass64(plan, A_dev, count, true, false, (hipStream_t)stream);
    if (rc == NTT_OK) rc = launch_pass64(plan, A_dev, count, true, true, (hipStream_t)stream);
    return rc;
}

}  // extern "C"
