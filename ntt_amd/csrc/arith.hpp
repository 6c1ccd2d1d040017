// Modular arithmetic for 32-bit RNS limbs on gfx950.
//
// Values are canonical residues in [0, q) with q odd and q < 2^31, the range
// the reference's single-subtraction Montgomery reduction supports
// (src/v04_montgomery.cu:110-118).  Conditional corrections use the unsigned
// min trick (one v_sub/v_add plus one v_min_u32) instead of compare+select.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ntt {

// Global-memory pointers: accesses through them use an SGPR base plus a
// 32-bit VGPR offset (global_load/global_store) instead of a 64-bit flat
// address per lane.
typedef __attribute__((address_space(1))) uint32_t gu32;
typedef __attribute__((address_space(1))) const uint32_t gcu32;
typedef __attribute__((address_space(1))) const unsigned char gcu8;
typedef __attribute__((address_space(1))) unsigned char gu8;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4 gu4;
typedef __attribute__((address_space(1))) const u32x4 gcu4;

// (a + b) mod q for a, b in [0, q): a + b < 2q < 2^32; if a + b < q the
// subtraction wraps to a value above 2^31 and min keeps a + b.
__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t q) {
    uint32_t s = a + b;
    return min(s, s - q);
}

// (a - b) mod q: if a < b the difference wraps above 2^31 while d + q wraps
// back into [0, q), so min picks the right representative.
__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t q) {
    uint32_t d = a - b;
    return min(d, d + q);
}

// Montgomery product a * b * 2^-32 mod q with q_inv = -q^-1 mod 2^32.
// Valid for a * b < 2^32 * q (so any 32-bit a when b < q): t + m q < 2^64 and
// the pre-correction result lies in [0, 2q).  The two 64-bit products map to
// v_mad_u64_u32, the low product m to v_mul_lo_u32.
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t q, uint32_t q_inv) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * q_inv;
    uint32_t u = (uint32_t)(((uint64_t)m * q + t) >> 32);
    return min(u, u - q);
}

// -q^-1 mod 2^32 by Newton iteration (q odd): each step doubles the number of
// correct low bits, 1 -> 2 -> 4 -> ... -> 32 bits after five steps.
__device__ __forceinline__ uint32_t neg_inv_mod_2_32(uint32_t q) {
    uint32_t inv = 1;
#pragma unroll
    for (int i = 0; i < 5; ++i) inv *= 2u - q * inv;
    return 0u - inv;
}

// 2^64 mod q (the Montgomery R^2 constant) for kernels that are handed only q.
__device__ __forceinline__ uint32_t r2_mod(uint32_t q) {
    uint64_t r = (UINT64_C(1) << 32) % q;
    return (uint32_t)((r * r) % q);
}

// Cooley-Tukey: (x, y) -> (x + w y, x - w y), w in Montgomery form.
__device__ __forceinline__ void ct_butterfly(uint32_t& x, uint32_t& y, uint32_t w, uint32_t q,
                                             uint32_t q_inv) {
    uint32_t v = mont_mul(y, w, q, q_inv);
    uint32_t u = x;
    x = add_mod(u, v, q);
    y = sub_mod(u, v, q);
}

// Gentleman-Sande: (x, y) -> (x + y, (x - y) w), w in Montgomery form.
__device__ __forceinline__ void gs_butterfly(uint32_t& x, uint32_t& y, uint32_t w, uint32_t q,
                                             uint32_t q_inv) {
    uint32_t u = x, v = y;
    x = add_mod(u, v, q);
    y = mont_mul(sub_mod(u, v, q), w, q, q_inv);
}

// ---- lazy reduction, q < 2^30 (so 4q < 2^32) ------------------------------
// Values are kept as representatives in [0, 4q) (forward) or [0, 2q)
// (inverse) and made canonical once, at the end of the transform; the
// arithmetic mod q, hence every canonical output, is unchanged.

// Montgomery product without the final correction: for a < 2^32 and b < q,
// t + m q < 2^64 and the result lies in [0, 2q).
__device__ __forceinline__ uint32_t mont_mul_lazy(uint32_t a, uint32_t b, uint32_t q, uint32_t q_inv) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * q_inv;
    return (uint32_t)(((uint64_t)m * q + t) >> 32);
}

// [0, 4q) -> [0, q)
__device__ __forceinline__ uint32_t canon_4q(uint32_t x, uint32_t q) {
    x = min(x, x - 2 * q);
    return min(x, x - q);
}

// Cooley-Tukey on x, y in [0, 4q): x is first brought into [0, 2q), then
// (x + wy, x - wy + 2q), both in [0, 4q).
__device__ __forceinline__ void ct_butterfly_lazy(uint32_t& x, uint32_t& y, uint32_t w, uint32_t q,
                                                  uint32_t q_inv) {
    const uint32_t q2 = 2 * q;
    const uint32_t u = min(x, x - q2);
    const uint32_t v = mont_mul_lazy(y, w, q, q_inv);
    x = u + v;
    y = u + q2 - v;
}

// The last Cooley-Tukey stage of a lazy transform, x, y in [0, 4q): the
// canonical (x + wy, x - wy) in [0, q), which saves canonicalising both
// outputs after a lazy butterfly (12 instructions instead of 8 + 8).
__device__ __forceinline__ void ct_butterfly_lazy_last(uint32_t& x, uint32_t& y, uint32_t w, uint32_t q,
                                                       uint32_t q_inv) {
    const uint32_t u = canon_4q(x, q);
    const uint32_t v = mont_mul(y, w, q, q_inv);
    x = add_mod(u, v, q);
    y = sub_mod(u, v, q);
}

// Gentleman-Sande on x, y in [0, 2q): (x + y mod 2q, (x - y + 2q) w), both in [0, 2q).
__device__ __forceinline__ void gs_butterfly_lazy(uint32_t& x, uint32_t& y, uint32_t w, uint32_t q,
                                                  uint32_t q_inv) {
    const uint32_t q2 = 2 * q;
    const uint32_t s = x + y;
    const uint32_t d = x + q2 - y;
    x = min(s, s - q2);
    y = mont_mul_lazy(d, w, q, q_inv);
}

}  // namespace ntt
