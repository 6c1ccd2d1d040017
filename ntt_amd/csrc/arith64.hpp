// Modular arithmetic for 64-bit words on gfx950: canonical residues in
// [0, q) with q odd and q < 2^62 (sums stay below 2^63); Montgomery form with
// R = 2^64.  Each 64 x 64 -> 128-bit product is a chain of 32-bit
// v_mad_u64_u32 on the integer VALU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ntt {

// a * b * 2^-64 mod q for a, b < q < 2^63, q_inv = -q^-1 mod 2^64:
// t + m q = 2^64 (hi(t) + hi(m q) + carry) with carry = (lo(t) != 0), since
// lo(t) + lo(m q) = 0 mod 2^64; the result lies in [0, 2q).
__device__ __forceinline__ uint64_t mont_mul64(uint64_t a, uint64_t b, uint64_t q, uint64_t q_inv) {
    const uint64_t tlo = a * b, thi = __umul64hi(a, b);
    const uint64_t m = tlo * q_inv;
    const uint64_t u = thi + __umul64hi(m, q) + (tlo != 0 ? 1u : 0u);
    return u >= q ? u - q : u;
}
__device__ __forceinline__ uint64_t add_mod64(uint64_t a, uint64_t b, uint64_t q) {
    const uint64_t s = a + b;   // < 2q < 2^63
    return s >= q ? s - q : s;
}
__device__ __forceinline__ uint64_t sub_mod64(uint64_t a, uint64_t b, uint64_t q) {
    return a >= b ? a - b : a + (q - b);
}
// Cooley-Tukey (x, y) -> (x + w y, x - w y); Gentleman-Sande (x, y) ->
// (x + y, (x - y) w); w in Montgomery form.
__device__ __forceinline__ void ct_butterfly64(uint64_t& x, uint64_t& y, uint64_t w, uint64_t q, uint64_t q_inv) {
    const uint64_t t = mont_mul64(y, w, q, q_inv);
    const uint64_t u = x;
    x = add_mod64(u, t, q);
    y = sub_mod64(u, t, q);
}
__device__ __forceinline__ void gs_butterfly64(uint64_t& x, uint64_t& y, uint64_t w, uint64_t q, uint64_t q_inv) {
    const uint64_t u = x, v = y;
    x = add_mod64(u, v, q);
    y = mont_mul64(sub_mod64(u, v, q), w, q, q_inv);
}

}  // namespace ntt
