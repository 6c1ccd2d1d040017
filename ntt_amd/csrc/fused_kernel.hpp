// Fused limb-per-workgroup negacyclic NTT / INTT for gfx950.
//
// A workgroup holds 2^LOGW words in registers: 2^(LOGW-n) limbs of N = 2^n
// words.  LOGW = 16: one workgroup per CU (it takes 132 KiB of the CU's
// 160 KiB LDS); LOGW = 15: two independent workgroups per CU (66 KiB of LDS
// each), so that one can compute while the other loads or stores.  With
// window width EB, a thread holds 2^EB words and the workgroup has
// 2^(LOGW-EB) threads; per CU, EB = 6 is 1024 threads x 64 words (128
// registers per thread available), EB = 7 is 512 threads x 128 words (256
// registers per thread available).  Every stage of the transform
// runs from registers; between rounds the words move through LDS in two
// halves.  Each limb is read from and written to HBM exactly once, instead of
// once per pass as in the two-pass kernels (src/v04_montgomery.cu:196-565 is
// the reference's two-pass form).
//
// Schedule (all compile-time for a given n; checked by the host model in
// tests/fused_model.py):
//  * round k owns an EB-bit window [lo_k, lo_k + EB) of the word index;
//    thread u of a limb holds the 2^EB words
//        i = ((u >> lo) << (lo + EB)) | (x << lo) | (u & (2^lo - 1));
//  * forward windows go from the top bits down, inverse from the bottom up,
//    and consecutive windows share a bit sigma; the exchange between them
//    moves the words with sigma = 0, then those with sigma = 1, through a
//    32K-word LDS buffer (word i of a limb at compress(i without sigma) plus
//    one padding word per 32);
//  * round k stores window bit b in register bit perm_k(b), a transposition
//    chosen so that sigma occupies the same register bit on both sides of an
//    exchange: in each half a thread reads back exactly the registers it has
//    just written out;
//  * the twiddle of the stage on window bit j of word x is
//    tw[((2^(n-EB-lo) + (u >> lo)) << (EB-1-j)) + (x >> (j+1))].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "arith.hpp"
#include "fused_types.hpp"

namespace ntt {

namespace fused {


__host__ __device__ constexpr int nrounds(int n, int eb) { return n <= 2 * eb - 1 ? 2 : 3; }

// Window low bit of round k: forward from the top bits down, inverse from
// the bottom up; consecutive windows overlap in at least one bit.
__host__ __device__ constexpr int win_lo(int n, bool inv, int k, int eb) {
    return nrounds(n, eb) == 2 ? (inv ? (k == 0 ? 0 : n - eb) : (k == 0 ? n - eb : 0))
                               : (inv ? (k == 0 ? 0 : k == 1 ? eb - 1 : n - eb)
                                      : (k == 0 ? n - eb : k == 1 ? n - 2 * eb + 1 : 0));
}

// Window-relative range [jlo, jhi] of the stages round k performs.
__host__ __device__ constexpr int stage_jhi(int n, bool inv, int k, int eb) {
    return !inv ? (k == 0 ? eb - 1 : win_lo(n, inv, k - 1, eb) - 1 - win_lo(n, inv, k, eb)) : eb - 1;
}
__host__ __device__ constexpr int stage_jlo(int n, bool inv, int k, int eb) {
    return !inv ? 0 : (k == 0 ? 0 : win_lo(n, inv, k - 1, eb) + eb - win_lo(n, inv, k, eb));
}

// Split bit of the exchange after round k.
__host__ __device__ constexpr int sigma(int n, bool inv, int k, int eb) {
    return !inv ? win_lo(n, inv, k, eb) : win_lo(n, inv, k, eb) + eb - 1;
}

// Register bit of window bit b in round k.
__host__ __device__ constexpr int perm_bit(int n, bool inv, int k, int b, int eb) {
    int p[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    for (int r = 1; r <= k; ++r) {
        const int s = sigma(n, inv, r - 1, eb);
        const int ra = p[s - win_lo(n, inv, r - 1, eb)];
        const int jb = s - win_lo(n, inv, r, eb);
        int q[8] = {0, 1, 2, 3, 4, 5, 6, 7};
        q[jb] = ra;
        q[ra] = jb;
        for (int i = 0; i < eb; ++i) p[i] = q[i];
    }
    return p[b];
}

__host__ __device__ constexpr int reg_of(int n, bool inv, int k, int x, int eb) {
    int r = 0;
    for (int b = 0; b < eb; ++b)
        if ((x >> b) & 1) r |= 1 << perm_bit(n, inv, k, b, eb);
    return r;
}

// Register index as a constant expression.
template <int n, bool INV, int K, int X, int EB>
inline constexpr int REG = reg_of(n, INV, K, X, EB);

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1, so that every
// register index derived from I is a constant expression.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Word index of register x in a round with window LO, split into the part
// that depends on the thread (ubase) and the part that depends on x only
// (bit fields are disjoint, so the index is their sum): every access is a
// per-thread base plus a compile-time offset.
// A copy of v the compiler cannot see through: values derived from it are
// computed where they are used instead of once, early, and kept live (the
// kernel has no registers to spare while it holds 2^EB words per thread).
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <int LO, int EB>
__device__ __forceinline__ uint32_t ubase(uint32_t u) {
    return ((u >> LO) << (LO + EB)) + (u & ((1u << LO) - 1));
}

// LDS word of word i in an exchange split on bit S: i without bit S
// ("compress"), plus one padding word per 2^PAD_SHIFT.  Both steps distribute
// over sums of disjoint bit fields, so the position of a word is
// pad(compress(thread part)) + pad(compress(register part)).  The padding
// keeps the lanes of each exchange access on distinct banks: one word per 32
// for window width 6, one per 64 for width 7 (one per 32 puts the lanes of
// width-7 exchanges on 2-4 way conflicting banks).
__host__ __device__ constexpr int pad_shift(int eb) { return eb == 7 ? 6 : 5; }

template <int EB>
__host__ __device__ constexpr uint32_t lds_pad(uint32_t c) { return c + (c >> pad_shift(EB)); }

template <int S>
__host__ __device__ constexpr uint32_t compress(uint32_t i) {
    return ((i >> (S + 1)) << S) | (i & ((1u << S) - 1));
}

template <int S, int LO, int EB>
__device__ __forceinline__ uint32_t lds_ubase(uint32_t u) {
    return lds_pad<EB>(compress<S>(ubase<LO, EB>(u)));
}

template <int S, int LO, int X, int EB>
inline constexpr uint32_t LDS_OFF = lds_pad<EB>(compress<S>((uint32_t)X << LO));

// OT: in the round whose window is the low bits (LO == 0: every thread needs
// its own twiddles), each twiddle is formed on the fly from the table's first
// 64 words and its msb appendix (FusedJob tables, fused_types.hpp) instead of
// being read from the table's upper levels.
// TL: the stage twiddles (and the on-the-fly factors) come from the staged
// LDS copy twl (tw[0 .. FUSED_TWL_TW - 1], then the msb appendix) instead of
// the table in HBM.
// SPLIT: the transform is one half of a 2^(n+1)-word limb (after the forward's
// cross-half stage, before the inverse's), whose stage twiddles are those of
// the 2^(n+1)-word table at (tb0 m + g) for level m and group g, with
// tb0 = 2 + half (tb0 = 1: a whole limb); the lazy forward's input scale is
// then applied at the cross-half stage, not here.
template <int n, bool INV, bool LAZY, int K, int EB, bool OT = false, bool TL = false, bool SPLIT = false>
__device__ __forceinline__ void stages(uint32_t (&v)[1 << EB], gcu32* __restrict__ tw, const uint32_t* twl,
                                       uint32_t u, uint32_t q, uint32_t q_inv, uint32_t scale = 0,
                                       uint32_t tb0 = 1) {
    constexpr int E = 1 << EB;
    constexpr int LO = win_lo(n, INV, K, EB);
    constexpr int JLO = stage_jlo(n, INV, K, EB), JHI = stage_jhi(n, INV, K, EB);
    // In the forward's first round (window at the top bits, LO + EB == n)
    // every thread of a limb has u < 2^LO, so TB = 1 and the stage twiddles
    // are tw[1 .. 2^EB - 1] for the whole limb: compile-time offsets from the
    // (wave-uniform) table base, read with scalar loads into SGPRs.  (The
    // inverse's last round has the same property, but there the SGPR
    // twiddles raised the lazy kernels' spills and measured slower.)
    const uint32_t TB = (!INV && LO + EB == n) ? tb0 : (tb0 << (n - EB - LO)) + (opaque(u) >> LO);
    // The inverse's last stage (bit n - 1, twiddle tw[1] for every word) is
    // folded into the output scaling by run_item.
    constexpr bool FOLDED_TOP = INV && K == nrounds(n, EB) - 1;
    // The forward's last stage (bit 0, in the last round) leaves canonical
    // values when the arithmetic is lazy.
    constexpr bool LAST_CANON = !INV && K == nrounds(n, EB) - 1;
    // The lazy forward's first stage (bit n - 1, twiddle tw[1] for every
    // pair) applies the input scale s as well: x' = x s / R and
    // y' = y (s w) / R, both lazy in [0, 2q), then (x' + y', x' + 2q - y')
    // in [0, 4q): two Montgomery products per pair instead of three.
    constexpr bool FOLD_SCALE = LAZY && !INV && K == 0 && !SPLIT;
    static_for<EB>([&](auto JJ) __attribute__((always_inline)) {
        constexpr int j = INV ? decltype(JJ)::value : EB - 1 - decltype(JJ)::value;
        if constexpr (FOLD_SCALE && j == EB - 1) {
            static_assert(LO + EB == n, "the forward's first round holds the top bits");
            const uint32_t sw = mont_mul(scale, *(gcu32*)((gcu8*)tw + 4), q, q_inv);
            const uint32_t q2 = 2 * q;
            static_for<E / 2>([&](auto P) __attribute__((always_inline)) {
                constexpr int x = decltype(P)::value;
                constexpr int a = REG<n, INV, K, x, EB>, b = REG<n, INV, K, x | (1 << j), EB>;
                // opaque: otherwise the compiler widens every loaded word
                // into a zero-extended register pair for the 64-bit products
                // (twice the registers for the words in flight, and spills)
                const uint32_t xs = mont_mul_lazy(opaque(v[a]), scale, q, q_inv);
                const uint32_t ys = mont_mul_lazy(opaque(v[b]), sw, q, q_inv);
                v[a] = xs + ys;
                v[b] = xs + q2 - ys;
            });
        } else if constexpr (j >= JLO && j <= JHI && !(FOLDED_TOP && j == EB - 1)) {
            // on-the-fly twiddles of this stage: index (TB << k) + c with
            // c < 2^k, k <= 6, splits as lo = (TB mod 2^(6-k)) << k + c and
            // hi = TB >> (6 - k); msb[hi] is one word per thread per stage
            constexpr int k = EB - 1 - j;
            constexpr bool OTS = OT && LO == 0 && k <= OT_LSB_BITS;
            uint32_t ot_msb = 0, ot_lo = 0;
            if constexpr (OTS) {
                // the msb appendix follows the table (of 2^(n+1) words for a split limb)
                constexpr uint32_t MSB = 1u << (SPLIT ? n + 1 : n);
                // the staged msb words follow the staged table words
                constexpr uint32_t TWL_TW = SPLIT ? fused_split_twl_tw(INV) : FUSED_TWL_TW;
                ot_msb = TL ? twl[TWL_TW + (TB >> (OT_LSB_BITS - k))] : tw[MSB + (TB >> (OT_LSB_BITS - k))];
                ot_lo = (TB & ((1u << (OT_LSB_BITS - k)) - 1)) << k;
            }
            static_for<E / 2>([&](auto P) __attribute__((always_inline)) {
                // P enumerates the words with bit j clear: x = P with a 0 inserted at bit j
                constexpr int p = decltype(P)::value;
                constexpr int x = ((p >> j) << (j + 1)) | (p & ((1 << j) - 1));
                constexpr int a = REG<n, INV, K, x, EB>, b = REG<n, INV, K, x | (1 << j), EB>;
                uint32_t w;
                if constexpr (OTS)
                    w = mont_mul(TL ? twl[ot_lo + (uint32_t)(x >> (j + 1))] : tw[ot_lo + (uint32_t)(x >> (j + 1))],
                                 ot_msb, q, q_inv);
                else if constexpr (TL)
                    w = twl[(TB << (EB - 1 - j)) + (uint32_t)(x >> (j + 1))];
                else   // 32-bit byte offset: SGPR table base plus a VGPR offset per load
                    w = *(gcu32*)((gcu8*)tw + (((TB << (EB - 1 - j)) + (uint32_t)(x >> (j + 1))) << 2));
                if (LAZY) {
                    if (INV)
                        gs_butterfly_lazy(v[a], v[b], w, q, q_inv);
                    else if (LAST_CANON && j == 0)
                        ct_butterfly_lazy_last(v[a], v[b], w, q, q_inv);
                    else
                        ct_butterfly_lazy(v[a], v[b], w, q, q_inv);
                } else if (INV) {
                    gs_butterfly(v[a], v[b], w, q, q_inv);
                } else {
                    ct_butterfly(v[a], v[b], w, q, q_inv);
                }
            });
        }
    });
}

template <int n, bool INV, int K, int EB>
__device__ __forceinline__ void exchange(uint32_t (&v)[1 << EB], uint32_t* L, uint32_t u, bool valid) {
    constexpr int E = 1 << EB;
    constexpr int S = sigma(n, INV, K, EB);
    constexpr int LOA = win_lo(n, INV, K, EB), LOB = win_lo(n, INV, K + 1, EB);
    constexpr int JA = S - LOA, JB = S - LOB;
    // LDS accesses take a base register plus an immediate byte offset below
    // 64 KiB; words further out go through a second and third base register
    // (one per 16K words), made here and dead after the exchange, instead of
    // one address register per word kept live by the compiler.
    constexpr uint32_t WIN = 16384;
    const uint32_t ua = lds_ubase<S, LOA, EB>(opaque(u)), ub = lds_ubase<S, LOB, EB>(opaque(u));
    uint32_t* const LA[3] = {L + ua, L + opaque(ua + WIN), L + opaque(ua + 2 * WIN)};
    uint32_t* const LB[3] = {L + ub, L + opaque(ub + WIN), L + opaque(ub + 2 * WIN)};
    static_for<2>([&](auto H) __attribute__((always_inline)) {
        constexpr int h = decltype(H)::value;
        if (valid) {
            static_for<E>([&](auto X) __attribute__((always_inline)) {
                constexpr int x = decltype(X)::value;
                constexpr uint32_t o = LDS_OFF<S, LOA, x, EB>;
                if constexpr (((x >> JA) & 1) == h) LA[o / WIN][o % WIN] = v[REG<n, INV, K, x, EB>];
            });
        }
        __syncthreads();
        if (valid) {
            static_for<E>([&](auto X) __attribute__((always_inline)) {
                constexpr int x = decltype(X)::value;
                constexpr uint32_t o = LDS_OFF<S, LOB, x, EB>;
                if constexpr (((x >> JB) & 1) == h) v[REG<n, INV, K + 1, x, EB>] = LB[o / WIN][o % WIN];
            });
        }
        __syncthreads();
    });
}

// Every field of a job descriptor, and of the limb constants, is the same in
// all lanes (a work item is wave-uniform), but a descriptor read from memory,
// or merged with one from the kernel arguments, lands in VGPRs.  Moving the
// fields to SGPRs keeps pointers out of VGPRs (global accesses then take an
// SGPR base plus a 32-bit VGPR offset) and frees registers for the words.
// Orders the LDS accesses of one wavefront (they complete in order; this
// keeps the compiler from moving them across the point).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ FusedJob uniform_job(const FusedJob& j) {
    FusedJob u;
    u.data = uniform_ptr(j.data);
    u.src = uniform_ptr(j.src);
    u.sub = uniform_ptr(j.sub);
    u.post = uniform_ptr(j.post);
    u.tw = uniform_ptr(j.tw);
    u.lc = uniform_ptr(j.lc);
    u.subf = uniform_ptr(j.subf);
    u.limbs = __builtin_amdgcn_readfirstlane(j.limbs);
    u.period = __builtin_amdgcn_readfirstlane(j.period);
    u.item_begin = __builtin_amdgcn_readfirstlane(j.item_begin);
    u.skip_start = __builtin_amdgcn_readfirstlane(j.skip_start);
    u.skip_len = __builtin_amdgcn_readfirstlane(j.skip_len);
    u.src_period = __builtin_amdgcn_readfirstlane(j.src_period);
    u.src_first = __builtin_amdgcn_readfirstlane(j.src_first);
    u.sub_start = __builtin_amdgcn_readfirstlane(j.sub_start);
    u.sub_end = __builtin_amdgcn_readfirstlane(j.sub_end);
    u.n = j.n;
    u.inverse = j.inverse;
    return u;
}

// Word pairs per chunk of the split kernels' pipelined loads (two chunks in
// flight beside the 64 words a thread holds).
constexpr int SPLIT_LOAD_CHUNK = 8;

// The end of half `half` of a split inverse (run_item): the half's top
// stage (bit n - 1, twiddle tw[2 + half]).  Half 0 leaves its result T0 in
// half 0 of the output; half 1 reads T0 back at the positions this thread
// stored it (so no other thread's stores are involved) and computes the
// cross-half stage (bit n, twiddle tw[1]) folded with the output scale:
// ((a + b) s, (a - b) (s w) / R) / R, a = T0, b = T1, for both halves.
template <int n, bool LAZY, int EB>
__device__ __forceinline__ void split_inverse_store(uint32_t (&v)[1 << EB], const FusedJob& job, size_t dlimb,
                                                    uint32_t u, gcu32* tw, const LimbConst& c, uint32_t half) {
    constexpr int E = 1 << EB;
    constexpr int NR = nrounds(n, EB);
    constexpr int LOL = win_lo(n, true, NR - 1, EB);
    const uint32_t w = tw[2 + half];
    static_for<E / 2>([&](auto P) __attribute__((always_inline)) {
        constexpr int x = decltype(P)::value, y = x | (E / 2);
        constexpr int ra = REG<n, true, NR - 1, x, EB>, rb = REG<n, true, NR - 1, y, EB>;
        if (LAZY)
            gs_butterfly_lazy(v[ra], v[rb], w, c.q, c.q_inv);
        else
            gs_butterfly(v[ra], v[rb], w, c.q, c.q_inv);
    });
    gu32* const O = (gu32*)(job.data + (dlimb << (n + 1))) + ubase<LOL, EB>(opaque(u));
    if (half == 0) {
        static_for<E>([&](auto X) __attribute__((always_inline)) {
            constexpr int x = decltype(X)::value;
            O[x << LOL] = v[REG<n, true, NR - 1, x, EB>];
        });
        return;
    }
    const uint32_t s2 = mont_mul(c.scale, tw[1], c.q, c.q_inv);
    const uint32_t qq = LAZY ? 2 * c.q : c.q;
    // software-pipelined: the loads of chunk k + 1 are issued before chunk k
    // is combined and stored
    constexpr int NCH = E / SPLIT_LOAD_CHUNK;
    uint32_t to[2][SPLIT_LOAD_CHUNK];
    static_for<NCH + 1>([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value;
        if constexpr (k < NCH) {
            static_for<SPLIT_LOAD_CHUNK>([&](auto J) __attribute__((always_inline)) {
                constexpr int x = k * SPLIT_LOAD_CHUNK + decltype(J)::value;
                to[k & 1][decltype(J)::value] = O[x << LOL];
            });
        }
        asm volatile("" ::: "memory");
        if constexpr (k > 0) {
            static_for<SPLIT_LOAD_CHUNK>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value, x = (k - 1) * SPLIT_LOAD_CHUNK + j;
                const uint32_t a = to[(k - 1) & 1][j], b = v[REG<n, true, NR - 1, x, EB>];
                O[x << LOL] = mont_mul(a + b, c.scale, c.q, c.q_inv);
                O[(1u << n) + (x << LOL)] = mont_mul(a + qq - b, s2, c.q, c.q_inv);
            });
        }
    });
}

// TWL: one limb of 2^16 words per workgroup with on-the-fly twiddles; the
// twiddle words the rounds read are staged into LDS after the exchange
// buffer whenever the limb's table differs from the staged one (*staged).
// SPLIT: the item is a limb of 2^(n+1) words, transformed as its two halves
// one after the other by this workgroup (so no other workgroup reads or
// writes the limb).  Forward: the first pass reads both halves, computes the
// cross-half stage (bit n, twiddle tw[1]) with the input scale, keeps half 0
// in registers and stores half 1's stage output in half 1 of the output (the
// words each thread has just read there); then each half's 2^n-word
// sub-transform runs with the twiddles of the 2^(n+1)-word table (stages),
// half 1's reloading the words this thread stored.  Inverse: half 0's
// sub-transform leaves its result in half 0 of the output, and half 1's
// combines it with its own (split_inverse_store).
template <int n, bool INV, bool LAZY, int EB, bool MD, bool OT, bool TWL = false, bool SPLIT = false>
__device__ __forceinline__ void run_item(const FusedJob& job, uint32_t first_limb, uint32_t* lds,
                                         const uint32_t** staged) {
    static_assert(!SPLIT || (n == 16 && EB == 6 && !MD), "split limbs: 2 x 2^16 words, width 6");
    constexpr int NL = SPLIT ? n + 1 : n;   // log2 of the limb stride
    constexpr int E = 1 << EB;
    constexpr int G_BITS = n - EB;
    constexpr uint32_t STRIDE = (1u << (n - 1)) + (1u << (n - 1 - pad_shift(EB)));
    constexpr int NR = nrounds(n, EB);
    constexpr int LO0 = win_lo(n, INV, 0, EB), LOL = win_lo(n, INV, NR - 1, EB);
    // forward stores (the mod-down form's too) through an in-wave transpose
    // (one limb per wave, 64 words per thread, the last window at the low bits)
    constexpr bool XST = !INV && EB == 6 && G_BITS >= 6 && LOL == 0;
    const uint32_t t = threadIdx.x;
    uint32_t slot = t >> G_BITS;
    uint32_t u = t & ((1u << G_BITS) - 1);
    // Opaque per item: otherwise the persistent loop lets the compiler hoist
    // every address derived from u out of the loop and keep it live.
    asm volatile("" : "+v"(u));
    if (G_BITS >= 6) slot = __builtin_amdgcn_readfirstlane(slot);   // one limb per wave
    const uint32_t limb = first_limb + slot;
    const bool valid = limb < job.limbs;
    const uint32_t lsafe = valid ? limb : first_limb;
    // job limb -> (operand, plan limb), skipping [skip_start, skip_start + skip_len)
    const uint32_t vper = job.period - job.skip_len;
    const uint32_t op = lsafe / vper, pv = lsafe - op * vper;
    const uint32_t per = pv + (pv >= job.skip_start ? job.skip_len : 0);
    const size_t dlimb = (size_t)op * job.period + per;
    gcu32* const InL = (gcu32*)(job.src + (((size_t)op * job.src_period + job.src_first + per) << NL));
    gcu32* tw = (gcu32*)job.tw[per];
    LimbConst c = job.lc[per];
    if constexpr (G_BITS >= 6) {   // one limb per wave: the limb's table and constants are wave-uniform
        tw = uniform_ptr(tw);
        c.q = __builtin_amdgcn_readfirstlane(c.q);
        c.q_inv = __builtin_amdgcn_readfirstlane(c.q_inv);
        c.scale = __builtin_amdgcn_readfirstlane(c.scale);
        c.flags = __builtin_amdgcn_readfirstlane(c.flags);
    }
    uint32_t* L = lds + slot * STRIDE;
    uint32_t* const twl = lds + FUSED_LDS_WORDS;
    if constexpr (TWL) {
        static_assert(n == 16 && OT, "staged twiddles: one 2^16-word limb per workgroup, on-the-fly twiddles");
        if ((const uint32_t*)tw != *staged) {   // uniform: the table base is in SGPRs
            __syncthreads();   // the previous item's rounds have read the staged words
            // a split limb's table has 2^(n+1) words and an appendix of 2^(n-5);
            // its later rounds read tw[0 .. 2047] (inverse) or tw[0 .. 4095]
            constexpr uint32_t TWL_TW = SPLIT ? fused_split_twl_tw(INV) : FUSED_TWL_TW;
            constexpr uint32_t MSB_WORDS = SPLIT ? FUSED_SPLIT_MSB_WORDS : FUSED_TWL_WORDS - FUSED_TWL_TW;
            for (uint32_t i = threadIdx.x; i < TWL_TW; i += blockDim.x) twl[i] = tw[i];
            for (uint32_t i = threadIdx.x; i < MSB_WORDS; i += blockDim.x) twl[TWL_TW + i] = tw[(1u << NL) + i];
            __syncthreads();
            *staged = (const uint32_t*)tw;
        }
    }

#pragma clang loop unroll(disable)
    for (uint32_t half = 0; half < (SPLIT ? 2u : 1u); ++half) {
    gcu32* const In = InL + (SPLIT && INV ? half << n : 0u);
    uint32_t v[E];
    // ---- load (round 0 layout) ----
    if (SPLIT && !INV && half == 1) {
        // half 1's cross-half stage output, stored by this thread in pass 0
        gcu32* Ab = (gcu32*)(job.data + (dlimb << NL)) + (1u << n) + ubase<LO0, EB>(opaque(u));
        static_for<E>([&](auto X) __attribute__((always_inline)) {
            constexpr int x = decltype(X)::value;
            v[REG<n, INV, 0, x, EB>] = Ab[x << LO0];
        });
    } else if constexpr (SPLIT && !INV) {
        // the cross-half stage (bit n, twiddle tw[1]) with the input scale s:
        // x' = x s / R, y' = y (s w) / R, then x' + y' (half 0, kept) and
        // x' - y' (half 1, stored); lazy: both in [0, 2q), results in [0, 4q)
        const uint32_t w1 = tw[1];
        const bool sc = LAZY || (c.flags & 1);
        const uint32_t sw = mont_mul(sc ? c.scale : 1u, w1, c.q, c.q_inv);
        const uint32_t yf = sc ? sw : w1;   // canonical without scale: y w
        gcu32* A0 = In + ubase<LO0, EB>(opaque(u));
        gcu32* A1 = A0 + (1u << n);
        gu32* D1 = (gu32*)(job.data + (dlimb << NL)) + (1u << n) + ubase<LO0, EB>(opaque(u));
        const uint32_t q2 = 2 * c.q;
        // software-pipelined in chunks of SCH word pairs: the loads of chunk
        // k + 1 are issued before chunk k is combined (two chunks of loads in
        // flight beside the 64 words; all 128 loads at once would spill)
        constexpr int NCH = E / SPLIT_LOAD_CHUNK;
        uint32_t ta[2][SPLIT_LOAD_CHUNK], tb[2][SPLIT_LOAD_CHUNK];
        static_for<NCH + 1>([&](auto K) __attribute__((always_inline)) {
            constexpr int k = decltype(K)::value;
            if constexpr (k < NCH) {
                static_for<SPLIT_LOAD_CHUNK>([&](auto J) __attribute__((always_inline)) {
                    constexpr int x = k * SPLIT_LOAD_CHUNK + decltype(J)::value;
                    ta[k & 1][decltype(J)::value] = A0[x << LO0];
                    tb[k & 1][decltype(J)::value] = A1[x << LO0];
                });
            }
            asm volatile("" ::: "memory");
            if constexpr (k > 0) {
                static_for<SPLIT_LOAD_CHUNK>([&](auto J) __attribute__((always_inline)) {
                    constexpr int j = decltype(J)::value, x = (k - 1) * SPLIT_LOAD_CHUNK + j;
                    const uint32_t a = ta[(k - 1) & 1][j], b = tb[(k - 1) & 1][j];
                    if constexpr (LAZY) {
                        const uint32_t xs = mont_mul_lazy(opaque(a), c.scale, c.q, c.q_inv);
                        const uint32_t ys = mont_mul_lazy(opaque(b), sw, c.q, c.q_inv);
                        v[REG<n, INV, 0, x, EB>] = xs + ys;
                        D1[x << LO0] = xs + q2 - ys;
                    } else {
                        const uint32_t xs = sc ? mont_mul(a, c.scale, c.q, c.q_inv) : a;
                        const uint32_t ys = mont_mul(b, yf, c.q, c.q_inv);
                        v[REG<n, INV, 0, x, EB>] = add_mod(xs, ys, c.q);
                        D1[x << LO0] = sub_mod(xs, ys, c.q);
                    }
                });
            }
        });
    } else if (valid) {
        if (LO0 == 0) {
            gcu4* src = (gcu4*)(In + ubase<LO0, EB>(opaque(u)));
            static_for<E / 4>([&](auto X4) __attribute__((always_inline)) {
                constexpr int x4 = decltype(X4)::value;
                const u32x4 w = src[x4];
                v[REG<n, INV, 0, 4 * x4, EB>] = w.x;
                v[REG<n, INV, 0, 4 * x4 + 1, EB>] = w.y;
                v[REG<n, INV, 0, 4 * x4 + 2, EB>] = w.z;
                v[REG<n, INV, 0, 4 * x4 + 3, EB>] = w.w;
            });
        } else {
            gcu32* Ab = In + ubase<LO0, EB>(opaque(u));
            static_for<E>([&](auto X) __attribute__((always_inline)) {
                constexpr int x = decltype(X)::value;
                v[REG<n, INV, 0, x, EB>] = Ab[x << LO0];
            });
        }
        // the input scale (the lazy forward applies it in its first stage)
        if (!INV && !LAZY && !SPLIT && (c.flags & 1)) {
#pragma unroll
            for (int x = 0; x < E; ++x) v[x] = mont_mul(v[x], c.scale, c.q, c.q_inv);
        }
    } else {
#pragma unroll
        for (int x = 0; x < E; ++x) v[x] = 0;
    }

    // staged twiddles: every round of the inverse, the forward's rounds after
    // the first (whose twiddles tw[1 .. 63] are scalar loads)
    const uint32_t tb0 = SPLIT ? 2u + half : 1u;
    stages<n, INV, LAZY, 0, EB, OT, TWL && INV, SPLIT>(v, tw, twl, u, c.q, c.q_inv, c.scale, tb0);
    exchange<n, INV, 0, EB>(v, L, u, valid);
    stages<n, INV, LAZY, 1, EB, OT, TWL, SPLIT>(v, tw, twl, u, c.q, c.q_inv, 0, tb0);
    if constexpr (NR == 3) {
        exchange<n, INV, 1, EB>(v, L, u, valid);
        stages<n, INV, LAZY, 2, EB, OT, TWL, SPLIT>(v, tw, twl, u, c.q, c.q_inv, 0, tb0);
    }

    // ---- store (last round layout) ----
    // The words pass through an empty asm first: otherwise the compiler sinks
    // the whole last round into the conditional store block below, where its
    // schedule interleaves the round's products and spills.
#pragma unroll
    for (int x = 0; x < E; ++x) asm volatile("" : "+v"(v[x]));
    if constexpr (SPLIT && INV) {
        split_inverse_store<n, LAZY, EB>(v, job, dlimb, u, tw, c, half);
        continue;
    }
    if (valid) {
        if (INV) {
            // the last stage (bit n - 1, twiddle w = itw[1]) folded with the
            // output scaling: ((a + b) * scale, (a - b) * (scale w / R)) / R;
            // a + b and a - b + 2q (q for canonical a, b) stay below 2^32
            const uint32_t s2 = mont_mul(c.scale, tw[1], c.q, c.q_inv);
            const uint32_t qq = LAZY ? 2 * c.q : c.q;
            static_for<E / 2>([&](auto P) __attribute__((always_inline)) {
                constexpr int x = decltype(P)::value, y = x | (E / 2);
                constexpr int ra = REG<n, INV, NR - 1, x, EB>, rb = REG<n, INV, NR - 1, y, EB>;
                const uint32_t a = v[ra], b = v[rb];
                v[ra] = mont_mul(a + b, c.scale, c.q, c.q_inv);
                v[rb] = mont_mul(a + qq - b, s2, c.q, c.q_inv);
            });
        }   // forward values are canonical: the last stage's lazy butterflies leave [0, q)
        if constexpr (MD) {
            // mod-down: (sub - NTT) * post, sub read at the positions of the
            // words in registers (before the transposed store, if any); limbs
            // outside [sub_start, sub_end) subtract from 0
            uint32_t f = job.post[per];
            if constexpr (G_BITS >= 6) f = __builtin_amdgcn_readfirstlane(f);
            const bool has_sub = per >= job.sub_start && per < job.sub_end;
            if (has_sub) {
                const uint32_t sw = job.sub_end - job.sub_start, sl = per - job.sub_start;
                uint32_t g = job.subf ? job.subf[sl] : 0u;
                if constexpr (G_BITS >= 6) g = __builtin_amdgcn_readfirstlane(g);
                gcu32* Sb = (gcu32*)(job.sub + (((size_t)op * sw + sl) << n)) + ubase<LOL, EB>(opaque(u));
                auto one = [&](uint32_t s) __attribute__((always_inline)) {
                    return job.subf ? mont_mul(s, g, c.q, c.q_inv) : s;
                };
                if constexpr (LOL == 0) {
                    static_for<E / 4>([&](auto X4) __attribute__((always_inline)) {
                        constexpr int x4 = decltype(X4)::value;
                        const u32x4 s4 = ((gcu4*)Sb)[x4];
                        const uint32_t sv[4] = {s4.x, s4.y, s4.z, s4.w};
                        static_for<4>([&](auto K4) __attribute__((always_inline)) {
                            constexpr int r = REG<n, INV, NR - 1, 4 * x4 + decltype(K4)::value, EB>;
                            v[r] = mont_mul(sub_mod(one(sv[decltype(K4)::value]), v[r], c.q), f, c.q, c.q_inv);
                        });
                    });
                } else {
                    static_for<E>([&](auto X) __attribute__((always_inline)) {
                        constexpr int x = decltype(X)::value;
                        constexpr int r = REG<n, INV, NR - 1, x, EB>;
                        v[r] = mont_mul(sub_mod(one(Sb[x << LOL]), v[r], c.q), f, c.q, c.q_inv);
                    });
                }
            } else {
#pragma unroll
                for (int x = 0; x < E; ++x) v[x] = mont_mul(sub_mod(0u, v[x], c.q), f, c.q, c.q_inv);
            }
        }
        gu32* A = (gu32*)(job.data + (dlimb << NL)) + (SPLIT ? half << n : 0u);
        if constexpr (XST) {
            // The wave's 64 threads hold 4096 consecutive words, lane l the 64
            // words (16 chunks of 16 B) at 64 l.  Stored as they are, each
            // 16-byte store would put its lanes 256 B apart; instead the wave
            // transposes its block through its own 8 KiB of the (now idle)
            // exchange buffer, half of the chunks at a time: in phase h every
            // lane writes its chunks 8h .. 8h + 7 to a 128-byte row (slot s of
            // lane l at s ^ ((l >> 1) & 7), distinct banks for the row writes
            // and the column reads), then store k writes words 256 k + 4 l from
            // the lanes l whose chunk l & 15 is in this phase: 128-byte lines
            // written whole.  Only wave-local synchronisation is needed.
            // (opaque: the addresses below are computed per item, not hoisted
            // out of the persistent loop and kept live)
            const uint32_t lane = opaque(threadIdx.x & 63u);
            u32x4* const Wl = (u32x4*)(lds + ((threadIdx.x >> 6) << 11));
            // byte offset of store 0 from the (wave-uniform) limb base: an SGPR
            // base plus one 32-bit VGPR offset per store
            const uint32_t so = ((((opaque(u) >> 6) << 12) + 4 * lane) << 2);
            gu8* const Ab = (gu8*)A;
            const uint32_t wsw = (lane >> 1) & 7u;
            static_for<2>([&](auto H) __attribute__((always_inline)) {
                constexpr uint32_t h = decltype(H)::value;
                static_for<8>([&](auto S) __attribute__((always_inline)) {
                    constexpr uint32_t ss = decltype(S)::value, cc = 8 * h + ss;
                    Wl[(lane << 3) + (ss ^ wsw)] =
                        u32x4{v[REG<n, INV, NR - 1, 4 * cc, EB>], v[REG<n, INV, NR - 1, 4 * cc + 1, EB>],
                              v[REG<n, INV, NR - 1, 4 * cc + 2, EB>], v[REG<n, INV, NR - 1, 4 * cc + 3, EB>]};
                });
                wave_lds_sync();
                if (((lane >> 3) & 1u) == h) {   // lane & 15 in [8h, 8h + 8)
                    const uint32_t sl = lane & 7u;
                    static_for<16>([&](auto K) __attribute__((always_inline)) {
                        constexpr uint32_t kk = decltype(K)::value;
                        const uint32_t row = 4 * kk + (lane >> 4);   // the source lane
                        *(gu4*)(Ab + (so + 1024u * kk)) = Wl[(row << 3) + (sl ^ ((row >> 1) & 7u))];
                    });
                }
                wave_lds_sync();
            });
        } else if (LOL == 0) {
            gu4* dst = (gu4*)(A + ubase<LOL, EB>(opaque(u)));
            static_for<E / 4>([&](auto X4) __attribute__((always_inline)) {
                constexpr int x4 = decltype(X4)::value;
                dst[x4] = u32x4{v[REG<n, INV, NR - 1, 4 * x4, EB>], v[REG<n, INV, NR - 1, 4 * x4 + 1, EB>],
                                v[REG<n, INV, NR - 1, 4 * x4 + 2, EB>], v[REG<n, INV, NR - 1, 4 * x4 + 3, EB>]};
            });
        } else {
            gu32* Ab = A + ubase<LOL, EB>(opaque(u));
            static_for<E>([&](auto X) __attribute__((always_inline)) {
                constexpr int x = decltype(X)::value;
                Ab[x << LOL] = v[REG<n, INV, NR - 1, x, EB>];
            });
        }
    }
    // the next item's (or half's) first exchange writes the LDS the
    // transposes read
    if constexpr (XST) __syncthreads();
    }   // halves
}

}  // namespace fused

// Persistent loop over work items of jobs of one size 2^n and one direction;
// an item is 2^(LOGW-n) consecutive limbs of one job.  Every thread of the
// workgroup runs every barrier of every item.
// MD: forward mod-down form, storing (sub - NTT) * post (FusedJob).
// OT: on-the-fly twiddles in the low-bit round (stages).
// The register budget is that of 2^(16-EB) threads per CU (2^(8-EB) waves
// per SIMD) for either LOGW.
// SPLIT (n = 16): an item is one limb of 2^17 words (run_item).
template <int n, bool INV, bool LAZY, int EB, bool MD, bool OT, int LOGW, bool SPLIT>
__device__ __forceinline__ void fused_body(FusedParams p) {
    static_assert(n <= LOGW && LOGW >= 15 && LOGW <= 16, "a workgroup holds whole limbs of at most 2^LOGW words");
    static_assert(!SPLIT || (n == 16 && LOGW == 16), "split limbs: two 2^16-word halves");
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ uint32_t s_next;
    uint32_t iter = 0;
    constexpr bool TWL = OT && n == 16 && LOGW == 16 && EB == 6;
    const uint32_t* staged = nullptr;   // table whose words the LDS holds (TWL)
    for (uint32_t s = 0, ns = (blockIdx.x & 3u) * p.stagger; s < ns; ++s) __builtin_amdgcn_s_sleep(127);
    const uint32_t total = p.total_items;
    for (;;) {
        uint32_t item;
        if (p.counter) {
            if (threadIdx.x == 0) s_next = atomicAdd(p.counter, 1u);
            __syncthreads();
            item = s_next;
            __syncthreads();
        } else {
            item = blockIdx.x + iter * gridDim.x;
            ++iter;
        }
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= total) {
            // dynamic scheduling: every workgroup leaves after one claim past
            // the end, so the last one out (counter[1] reaching the grid size)
            // is the last to touch the counters and returns both to zero for
            // the next launch (graph replays included)
            if (p.counter && threadIdx.x == 0 && atomicAdd(p.counter + 1, 1u) == gridDim.x - 1) {
                atomicExch(p.counter, 0u);
                atomicExch(p.counter + 1, 0u);
            }
            break;
        }
        FusedJob job = p.job0;
        if (p.jobs) {
            // last job with item_begin <= item
            uint32_t lo = 0, hi = p.njobs;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (p.jobs[mid].item_begin <= item) lo = mid; else hi = mid;
            }
            job = p.jobs[lo];
        }
        job = fused::uniform_job(job);
        fused::run_item<n, INV, LAZY, EB, MD, OT, TWL, SPLIT>(job, (item - job.item_begin) << (LOGW - n), lds,
                                                                      &staged);
    }
}

template <int n, bool INV, bool LAZY, int EB, bool MD = false, bool OT = false, int LOGW = 16>
__global__ __launch_bounds__(1 << (LOGW - EB)) __attribute__((amdgpu_waves_per_eu(1 << (8 - EB))))
void fused_kernel(FusedParams p) {
    fused_body<n, INV, LAZY, EB, MD, OT, LOGW, false>(p);
}

// Limbs of 2^17 words as two 2^16-word halves (fused_body, SPLIT): width 6,
// table twiddles, one workgroup of 1024 threads per CU.
template <bool INV, bool LAZY, bool OT = false>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4)))
void fused_split_kernel(FusedParams p) {
    fused_body<16, INV, LAZY, 6, false, OT, 16, true>(p);
}

}  // namespace ntt
