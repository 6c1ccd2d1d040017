/*
 * CPU oracle for the RNS multi-limb negacyclic NTT/INTT.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
 * oracle/_build/libntt_oracle.so and loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product library (ntt_amd/) never links, loads or calls it.
 *
 * It restates, in plain C, the reference's CPU path (paths relative to the
 * reference repository root):
 *   add_mod / sub_mod / mul_mod ........ src/v04_montgomery.cu:93-108
 *   mont_reduce / mont_mul ............. src/v04_montgomery.cu:110-123
 *   mont_q_inv / to_mont / mont_r2 ..... src/v04_montgomery.cu:125-141
 *   from_mont .......................... src/v04_montgomery.cu:143-146
 *   pow_mod / bit_reverse .............. src/v04_montgomery.cu:660-679
 *   make_modulus_runtime ............... src/v04_montgomery.cu:681-718
 *   make_twiddle_table ................. src/v04_montgomery.cu:720-727
 *   to/from_mont_inplace ............... src/v04_montgomery.cu:729-740
 *   cpu_ntt_butterfly .................. src/v04_montgomery.cu:742-759
 *   cpu_intt_butterfly ................. src/v04_montgomery.cu:761-784
 *   fill_random ........................ src/v04_montgomery.cu:786-797
 *   plain-domain INTT normalisation .... src/v03_template_config.cu:516-527
 * plus the Montgomery-variant end-to-end semantics of launch_ntt/launch_intt
 * (src/v04_montgomery.cu:229-232, 567-580, 1082-1087): forward output is
 * to_mont(cpu_ntt_butterfly(x)), inverse output is
 * cpu_intt_butterfly(from_mont(x)).
 *
 * fill_random uses std::mt19937 and libstdc++'s
 * std::uniform_int_distribution<uint32_t>; both are restated here: the
 * MT19937 recurrence and tempering from the C++ standard [rand.eng.mers]
 * (Matsumoto & Nishimura 1998), and Lemire's nearly-divisionless bounded
 * draw (ACM TOMACS 29(1), 2019), which is the algorithm GCC 11's libstdc++
 * selects for a 32-bit generator (bits/uniform_int_dist.h, _S_nd).
 *
 * plus, at 64-bit word width (the last section of this file), the same
 * textbook transforms for primes q < 2^62 (Cheddar's NTTHandler<uint64_t>
 * word width, src/NTT_cheddar.cu:1313).
 *
 * Pinning (each function is checked in tests/test_oracle.py):
 *   pow_mod, bit_reverse      Python pow(), known answers
 *   modulus runtime           psi^2N = 1, psi^N = q-1, N*N^-1 = 1, q*q_inv = -1
 *   twiddle table             closed form psi^brv(i); tw[1]^2 = -1
 *   cpu_ntt_butterfly         brute-force evaluation at psi^(2brv(k)+1),
 *                             negacyclic convolution theorem, linearity
 *   cpu_intt_butterfly        INTT(NTT(x)) = x, brute-force inverse
 *   Montgomery helpers        Python big-int identities
 *   fill_random               KAT of mt19937 (10000th output 4123659995),
 *                             and a g++-built program using <random>
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ---- modular arithmetic: src/v04_montgomery.cu:93-108 ---- */
uint32_t oracle_add_mod(uint32_t a, uint32_t b, uint32_t q) {
    uint64_t s = (uint64_t)a + b;
    if (s >= q) s -= q;
    return (uint32_t)s;
}

uint32_t oracle_sub_mod(uint32_t a, uint32_t b, uint32_t q) {
    if (a >= b) return a - b;
    return (uint32_t)((uint64_t)a + q - b);
}

uint32_t oracle_mul_mod(uint32_t a, uint32_t b, uint32_t q) {
    return (uint32_t)(((uint64_t)a * b) % q);
}

/* ---- Montgomery, R = 2^32: src/v04_montgomery.cu:110-146 ---- */
uint32_t oracle_mont_reduce(uint64_t t, uint32_t q, uint32_t q_inv) {
    uint32_t m = (uint32_t)t * q_inv;
    uint64_t u = (t + (uint64_t)m * q) >> 32;
    if (u >= q) u -= q;
    return (uint32_t)u;
}

uint32_t oracle_mont_mul(uint32_t a, uint32_t b, uint32_t q, uint32_t q_inv) {
    return oracle_mont_reduce((uint64_t)a * b, q, q_inv);
}

/* -q^{-1} mod 2^32 by five Newton steps (q odd). */
uint32_t oracle_mont_q_inv(uint32_t q) {
    uint32_t inv = 1;
    for (int i = 0; i < 5; ++i) inv *= 2u - q * inv;
    return 0u - inv;
}

uint32_t oracle_to_mont(uint32_t x, uint32_t q) {
    uint64_t r = (UINT64_C(1) << 32) % q;
    return (uint32_t)(((uint64_t)x * r) % q);
}

uint32_t oracle_mont_r2(uint32_t q) {
    uint64_t r = (UINT64_C(1) << 32) % q;
    return (uint32_t)((r * r) % q);
}

uint32_t oracle_from_mont(uint32_t x, uint32_t q, uint32_t q_inv) {
    return oracle_mont_mul(x, 1, q, q_inv);
}

/* ---- src/v04_montgomery.cu:660-679 ---- */
uint32_t oracle_pow_mod(uint32_t base, uint64_t exp, uint32_t mod) {
    uint64_t result = 1, x = base;
    while (exp > 0) {
        if (exp & 1) result = (result * x) % mod;
        x = (x * x) % mod;
        exp >>= 1;
    }
    return (uint32_t)result;
}

uint32_t oracle_bit_reverse(uint32_t value, int width) {
    uint32_t r = 0;
    for (int i = 0; i < width; ++i) {
        r = (r << 1) | (value & 1u);
        value >>= 1;
    }
    return r;
}

/* Field order matches ModulusRuntime, src/v04_montgomery.cu:17-26. */
typedef struct {
    uint32_t qi, primitive_root, q_inv, r2, psi, psi_inv, omega, degree_inv;
} oracle_modulus_runtime;

/*
 * src/v04_montgomery.cu:681-718.  The reference calls std::exit on failure;
 * here the failure is a return code: 1 = q != 1 mod 2N (685-688),
 * 2 = root does not generate a primitive 2N-th root (702-706).
 */
int oracle_make_modulus_runtime(int logN, uint32_t qi, uint32_t g,
                                oracle_modulus_runtime* out) {
    uint64_t N = UINT64_C(1) << logN;
    uint64_t order = 2 * N;
    if (((uint64_t)qi - 1) % order != 0) return 1;
    uint32_t psi = oracle_pow_mod(g, ((uint64_t)qi - 1) / order, qi);
    uint32_t psi_inv = oracle_pow_mod(psi, qi - 2, qi);
    uint32_t omega = (uint32_t)(((uint64_t)psi * psi) % qi);
    uint32_t degree_inv = oracle_pow_mod((uint32_t)N, qi - 2, qi);
    if (oracle_pow_mod(psi, order, qi) != 1 || oracle_pow_mod(psi, N, qi) != qi - 1)
        return 2;
    out->qi = qi;
    out->primitive_root = g;
    out->q_inv = oracle_mont_q_inv(qi);
    out->r2 = oracle_mont_r2(qi);
    out->psi = psi;
    out->psi_inv = psi_inv;
    out->omega = omega;
    out->degree_inv = degree_inv;
    return 0;
}

/* src/v04_montgomery.cu:720-727: tw[i] = psi^brv(i, logN). */
void oracle_make_twiddle_table(int logN, uint32_t q, uint32_t psi, uint32_t* tw) {
    size_t N = (size_t)1 << logN;
    for (size_t i = 0; i < N; ++i)
        tw[i] = oracle_pow_mod(psi, oracle_bit_reverse((uint32_t)i, logN), q);
}

/* src/v04_montgomery.cu:729-740 */
void oracle_to_mont_inplace(uint32_t* v, size_t n, uint32_t q) {
    for (size_t i = 0; i < n; ++i) v[i] = oracle_to_mont(v[i], q);
}

void oracle_from_mont_inplace(uint32_t* v, size_t n, uint32_t q, uint32_t q_inv) {
    for (size_t i = 0; i < n; ++i) v[i] = oracle_from_mont(v[i], q, q_inv);
}

/*
 * src/v04_montgomery.cu:742-759.  In-place Cooley-Tukey over a
 * normal-domain bit-reversed psi table; natural-order input,
 * bit-reversed-order output.
 */
void oracle_cpu_ntt_butterfly(uint32_t* a, int logN, uint32_t q, const uint32_t* psi) {
    size_t N = (size_t)1 << logN;
    size_t t = N >> 1;
    for (size_t m = 1; m < N; m *= 2) {
        for (size_t j = 0; j < m; ++j) {
            uint32_t w = psi[m + j];
            for (size_t k = j * 2 * t; k < j * 2 * t + t; ++k) {
                uint32_t u = a[k];
                uint32_t v = (uint32_t)(((uint64_t)a[k + t] * w) % q);
                a[k] = oracle_add_mod(u, v, q);
                a[k + t] = oracle_sub_mod(u, v, q);
            }
        }
        t >>= 1;
    }
}

/*
 * src/v04_montgomery.cu:761-784 (and src/v03_template_config.cu:685-705,
 * 516-527 for the plain variant, which is the same computation).
 * Gentleman-Sande over the bit-reversed psi^-1 table, then x N^-1.
 */
void oracle_cpu_intt_butterfly(uint32_t* a, int logN, uint32_t q,
                               const uint32_t* inv_psi, uint32_t degree_inv) {
    size_t N = (size_t)1 << logN;
    size_t t = 1;
    for (size_t m = N >> 1; m >= 1; m >>= 1) {
        for (size_t j = 0; j < m; ++j) {
            uint32_t w = inv_psi[m + j];
            for (size_t k = j * 2 * t; k < j * 2 * t + t; ++k) {
                uint32_t u = a[k], v = a[k + t];
                a[k] = oracle_add_mod(u, v, q);
                a[k + t] = oracle_mul_mod(oracle_sub_mod(u, v, q), w, q);
            }
        }
        t <<= 1;
        if (m == 1) break;
    }
    for (size_t i = 0; i < N; ++i) a[i] = oracle_mul_mod(a[i], degree_inv, q);
}

/*
 * Whole-operand helpers over the A[limb * N + i] layout.  `domain` 0 is the
 * plain variant (src/v03_template_config.cu: normal-domain output), 1 the
 * Montgomery variant (src/v04_montgomery.cu: forward output in Montgomery
 * form, inverse input in Montgomery form, inverse output normal).
 * Returns 0, or the make_modulus_runtime code of the first failing limb.
 */
int oracle_forward(uint32_t* A, int logN, const uint32_t* qs, const uint32_t* roots,
                   size_t limbs, int domain) {
    size_t N = (size_t)1 << logN;
    uint32_t* tw = (uint32_t*)malloc(N * sizeof(uint32_t));
    if (!tw) return 3;
    for (size_t l = 0; l < limbs; ++l) {
        oracle_modulus_runtime rt;
        int rc = oracle_make_modulus_runtime(logN, qs[l], roots[l], &rt);
        if (rc) { free(tw); return rc; }
        oracle_make_twiddle_table(logN, rt.qi, rt.psi, tw);
        uint32_t* a = A + l * N;
        oracle_cpu_ntt_butterfly(a, logN, rt.qi, tw);
        if (domain == 1) oracle_to_mont_inplace(a, N, rt.qi);
    }
    free(tw);
    return 0;
}

int oracle_inverse(uint32_t* A, int logN, const uint32_t* qs, const uint32_t* roots,
                   size_t limbs, int domain) {
    size_t N = (size_t)1 << logN;
    uint32_t* tw = (uint32_t*)malloc(N * sizeof(uint32_t));
    if (!tw) return 3;
    for (size_t l = 0; l < limbs; ++l) {
        oracle_modulus_runtime rt;
        int rc = oracle_make_modulus_runtime(logN, qs[l], roots[l], &rt);
        if (rc) { free(tw); return rc; }
        oracle_make_twiddle_table(logN, rt.qi, rt.psi_inv, tw);
        uint32_t* a = A + l * N;
        if (domain == 1) oracle_from_mont_inplace(a, N, rt.qi, rt.q_inv);
        oracle_cpu_intt_butterfly(a, logN, rt.qi, tw, rt.degree_inv);
    }
    free(tw);
    return 0;
}

/* ---- fill_random: src/v04_montgomery.cu:786-797 ---- */
typedef struct {
    uint32_t mt[624];
    int idx;
} oracle_mt19937;

void oracle_mt_seed(oracle_mt19937* g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

uint32_t oracle_mt_next(oracle_mt19937* g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* Uniform draw in [0, range) (range >= 1) by Lemire's method. */
uint32_t oracle_uniform_below(oracle_mt19937* g, uint32_t range) {
    uint64_t product = (uint64_t)oracle_mt_next(g) * range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        uint32_t threshold = (0u - range) % range;
        while (low < threshold) {
            product = (uint64_t)oracle_mt_next(g) * range;
            low = (uint32_t)product;
        }
    }
    return (uint32_t)(product >> 32);
}

/* One limb: a fresh mt19937(seed), n draws uniform in [0, q-1]. Returns 1 if q <= 1. */
int oracle_fill_random(uint32_t* arr, size_t n, uint32_t q, uint32_t seed) {
    if (q <= 1) return 1;
    oracle_mt19937 g;
    oracle_mt_seed(&g, seed);
    for (size_t i = 0; i < n; ++i) arr[i] = oracle_uniform_below(&g, q);
    return 0;
}

/* The reference re-seeds every limb with the same seed (v04:932-934). */
int oracle_fill_random_limbs(uint32_t* A, int logN, const uint32_t* qs, size_t limbs,
                             uint32_t seed) {
    size_t N = (size_t)1 << logN;
    for (size_t l = 0; l < limbs; ++l) {
        int rc = oracle_fill_random(A + l * N, N, qs[l], seed);
        if (rc) return rc;
    }
    return 0;
}

/* ---- 64-bit words --------------------------------------------------------
 * The same textbook transforms (cpu_ntt_butterfly / cpu_intt_butterfly,
 * src/v04_montgomery.cu:742-784) over 64-bit words and primes q < 2^62, the
 * word width of Cheddar's NTTHandler<uint64_t> (src/NTT_cheddar.cu:1313),
 * which has no CPU path of its own.  Normal domain throughout: the forward
 * maps natural-order coefficients to the bit-reversed evaluations at
 * psi^(2 brv(k) + 1); the inverse undoes it including the N^-1 scaling.
 * psi = root^((q-1) / 2N) must be a primitive 2N-th root (psi^N = q - 1).
 * Pinned in tests/test_oracle.py: equals the 32-bit oracle for q < 2^31,
 * brute-force evaluation with Python integers, roundtrip. */
typedef unsigned __int128 u128;

uint64_t oracle_mul_mod64(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }

uint64_t oracle_pow_mod64(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = oracle_mul_mod64(r, b, q);
        b = oracle_mul_mod64(b, b, q);
        e >>= 1;
    }
    return r;
}

/* psi for (q, root) at logN; 0 if q != 1 mod 2N or psi is not primitive. */
uint64_t oracle_psi64(int logN, uint64_t q, uint64_t root) {
    const uint64_t two_n = (uint64_t)2 << logN;
    if (q < 3 || (q - 1) % two_n) return 0;
    const uint64_t psi = oracle_pow_mod64(root, (q - 1) / two_n, q);
    if (oracle_pow_mod64(psi, two_n >> 1, q) != q - 1) return 0;
    return psi;
}

static void twiddles64(int logN, uint64_t q, uint64_t psi, uint64_t* tw) {
    const uint32_t N = 1u << logN;
    for (uint32_t i = 0; i < N; ++i) tw[i] = oracle_pow_mod64(psi, oracle_bit_reverse(i, logN), q);
}

int oracle_forward64(uint64_t* A, int logN, const uint64_t* qs, const uint64_t* roots, size_t limbs) {
    const uint32_t N = 1u << logN;
    uint64_t* tw = (uint64_t*)malloc(sizeof(uint64_t) * N);
    if (!tw) return 2;
    for (size_t l = 0; l < limbs; ++l) {
        const uint64_t q = qs[l], psi = oracle_psi64(logN, q, roots[l]);
        if (!psi) { free(tw); return 1; }
        twiddles64(logN, q, psi, tw);
        uint64_t* a = A + l * N;
        for (uint32_t m = 1, t = N >> 1; m < N; m <<= 1, t >>= 1)
            for (uint32_t i = 0; i < m; ++i) {
                const uint64_t w = tw[m + i];
                for (uint32_t j = 2 * i * t; j < 2 * i * t + t; ++j) {
                    const uint64_t u = a[j], v = oracle_mul_mod64(a[j + t], w, q);
                    a[j] = (uint64_t)(((u128)u + v) % q);
                    a[j + t] = (uint64_t)(((u128)u + q - v) % q);
                }
            }
    }
    free(tw);
    return 0;
}

int oracle_inverse64(uint64_t* A, int logN, const uint64_t* qs, const uint64_t* roots, size_t limbs) {
    const uint32_t N = 1u << logN;
    uint64_t* tw = (uint64_t*)malloc(sizeof(uint64_t) * N);
    if (!tw) return 2;
    for (size_t l = 0; l < limbs; ++l) {
        const uint64_t q = qs[l], psi = oracle_psi64(logN, q, roots[l]);
        if (!psi) { free(tw); return 1; }
        twiddles64(logN, q, oracle_pow_mod64(psi, q - 2, q), tw);
        uint64_t* a = A + l * N;
        for (uint32_t m = N >> 1, t = 1; m >= 1; m >>= 1, t <<= 1)
            for (uint32_t i = 0; i < m; ++i) {
                const uint64_t w = tw[m + i];
                for (uint32_t j = 2 * i * t; j < 2 * i * t + t; ++j) {
                    const uint64_t u = a[j], v = a[j + t];
                    a[j] = (uint64_t)(((u128)u + v) % q);
                    a[j + t] = oracle_mul_mod64((uint64_t)(((u128)u + q - v) % q), w, q);
                }
            }
        const uint64_t ninv = oracle_pow_mod64(N, q - 2, q);
        for (uint32_t j = 0; j < N; ++j) a[j] = oracle_mul_mod64(a[j], ninv, q);
    }
    free(tw);
    return 0;
}
