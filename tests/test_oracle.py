"""Pins the CPU oracle (oracle/ntt_oracle.c) against mathematics that does not
depend on it: Python big-int arithmetic, closed-form evaluation, the
negacyclic convolution theorem, textbook MT19937 known answers and libstdc++'s
own <random>.  Each check is chosen so that a plausible slip in the oracle (a
wrong twiddle index, a dropped normalisation, a sign error) fails it; the
planted-mistake tests at the end show that they do.
"""
import os
import random
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from tests import mathref

R32 = 1 << 32

# default_moduli(), src/ntt_config.h:75-94 (q, smallest primitive root)
DEFAULT_MODULI = [
    (974258177, 3), (1081212929, 6), (1196556289, 7), (993263617, 5),
    (989986817, 3), (1074266113, 5), (1168900097, 5), (1010565121, 7),
    (957349889, 6), (1073872897, 7), (1209139201, 31), (994705409, 3),
    (998244353, 3), (1073479681, 11), (1158676481, 3), (1016463361, 38),
]
SMALL_PRIMES = [(17, 3), (97, 5), (7681, 17), (12289, 11), (998244353, 3), (1073479681, 11)]


def test_pow_mod_matches_python(oracle_mod):
    rng = random.Random(1)
    for _ in range(2000):
        m = rng.randrange(2, 1 << 31)
        b = rng.randrange(0, 1 << 32)
        e = rng.randrange(0, 1 << 40)
        assert oracle_mod.pow_mod(b, e, m) == pow(b % m if m > 1 else 0, e, m) or m == 1


def test_bit_reverse(oracle_mod):
    assert oracle_mod.bit_reverse(1, 16) == 32768
    assert oracle_mod.bit_reverse(0b0011, 4) == 0b1100
    for w in (1, 5, 12, 17):
        for v in range(0, 1 << w, max(1, (1 << w) // 97)):
            assert oracle_mod.bit_reverse(v, w) == int(format(v, f"0{w}b")[::-1], 2)


@pytest.mark.parametrize("q", [m[0] for m in DEFAULT_MODULI] + [17, 7681, 2147483629])
def test_montgomery_identities(oracle_mod, q):
    q_inv = oracle_mod.mont_q_inv(q)
    assert (q * q_inv) % R32 == R32 - 1                   # q * q_inv = -1 mod 2^32
    assert oracle_mod.mont_r2(q) == pow(2, 64, q)
    rinv = pow(R32, -1, q)
    rng = random.Random(q)
    for _ in range(300):
        a, b = rng.randrange(q), rng.randrange(q)
        assert oracle_mod.to_mont(a, q) == a * R32 % q
        assert oracle_mod.mont_mul(a, b, q, q_inv) == a * b * rinv % q
        assert oracle_mod.from_mont(a, q, q_inv) == a * rinv % q
        assert oracle_mod.add_mod(a, b, q) == (a + b) % q
        assert oracle_mod.sub_mod(a, b, q) == (a - b) % q
        assert oracle_mod.mul_mod(a, b, q) == a * b % q
        # R^2 conversion accepts any 32-bit input (src/v04_montgomery.cu:229-232)
        x = rng.randrange(R32)
        assert oracle_mod.mont_mul(x, oracle_mod.mont_r2(q), q, q_inv) == x * R32 % q


@pytest.mark.parametrize("logN", [3, 10, 16])
def test_modulus_runtime_invariants(oracle_mod, logN):
    N = 1 << logN
    for q, g in DEFAULT_MODULI + [(17, 3), (97, 5), (7681, 17)]:
        if (q - 1) % (2 * N):
            assert oracle_mod.make_modulus_runtime(logN, q, g)[0] == 1
            continue
        rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
        assert rc == 0
        assert pow(rt["psi"], 2 * N, q) == 1 and pow(rt["psi"], N, q) == q - 1
        assert rt["psi"] * rt["psi_inv"] % q == 1
        assert rt["omega"] == rt["psi"] ** 2 % q
        assert N * rt["degree_inv"] % q == 1
        assert (q * rt["q_inv"]) % R32 == R32 - 1
        assert rt["r2"] == pow(2, 64, q)
        assert rt["psi"] == pow(g, (q - 1) // (2 * N), q)


def test_modulus_runtime_errors(oracle_mod):
    assert oracle_mod.make_modulus_runtime(17, 974258177, 3)[0] == 1    # v2(q-1) = 17 < 18
    assert oracle_mod.make_modulus_runtime(10, 974258177, 1)[0] == 2    # root 1 is not primitive
    # a quadratic residue is never a primitive root
    assert oracle_mod.make_modulus_runtime(10, 998244353, 4)[0] == 2


@pytest.mark.parametrize("logN", [1, 4, 9, 13])
def test_twiddle_table_closed_form(oracle_mod, logN):
    q, g = 998244353, 3
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    tw = oracle_mod.twiddle_table(logN, q, rt["psi"])
    assert tw[0] == 1
    if logN >= 1:
        assert int(tw[1]) ** 2 % q == q - 1                 # psi^(N/2) = sqrt(-1)
    for i in range(0, 1 << logN, max(1, (1 << logN) // 257)):
        assert tw[i] == pow(rt["psi"], int(format(i, f"0{logN}b")[::-1], 2), q)


@pytest.mark.parametrize("q,g", SMALL_PRIMES + DEFAULT_MODULI[:3])
@pytest.mark.parametrize("logN", [1, 2, 3, 5, 8])
def test_ntt_matches_brute_force(oracle_mod, q, g, logN):
    N = 1 << logN
    if (q - 1) % (2 * N):
        pytest.skip("modulus does not support this N")
    rng = random.Random(q * 31 + logN)
    a = [rng.randrange(q) for _ in range(N)]
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    assert rc == 0
    tw = oracle_mod.twiddle_table(logN, q, rt["psi"])
    got = oracle_mod.cpu_ntt_butterfly(np.array(a, dtype=np.uint32), logN, q, tw)
    assert got.tolist() == mathref.negacyclic_eval_bitrev(a, q, rt["psi"])
    itw = oracle_mod.twiddle_table(logN, q, rt["psi_inv"])
    back = oracle_mod.cpu_intt_butterfly(got, logN, q, itw, rt["degree_inv"])
    assert back.tolist() == a
    # inverse closed form on an arbitrary (not NTT-produced) input
    b = [rng.randrange(q) for _ in range(N)]
    assert oracle_mod.cpu_intt_butterfly(np.array(b, dtype=np.uint32), logN, q, itw, rt["degree_inv"]).tolist() == \
        mathref.negacyclic_interp_bitrev(b, q, rt["psi"])


@pytest.mark.parametrize("logN", [4, 6, 8])
def test_negacyclic_convolution_theorem(oracle_mod, logN):
    q, g = 1073479681, 11
    N = 1 << logN
    rng = np.random.default_rng(logN)
    a = rng.integers(0, q, N, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, q, N, dtype=np.uint64).astype(np.uint32)
    fa = oracle_mod.forward(a[None, :], logN, [(q, g)], oracle_mod.PLAIN)[0]
    fb = oracle_mod.forward(b[None, :], logN, [(q, g)], oracle_mod.PLAIN)[0]
    prod = (fa.astype(object) * fb.astype(object) % q).astype(np.uint32)
    c = oracle_mod.inverse(prod[None, :], logN, [(q, g)], oracle_mod.PLAIN)[0]
    assert c.tolist() == mathref.negacyclic_mul(a.tolist(), b.tolist(), q)


def test_linearity_and_roundtrip_multilimb(oracle_mod):
    logN = 12
    moduli = DEFAULT_MODULI
    rng = np.random.default_rng(7)
    qs = np.array([m[0] for m in moduli], dtype=np.uint64)[:, None]
    x = (rng.integers(0, 1 << 62, (16, 1 << logN), dtype=np.uint64) % qs).astype(np.uint32)
    y = (rng.integers(0, 1 << 62, (16, 1 << logN), dtype=np.uint64) % qs).astype(np.uint32)
    s = ((x.astype(np.uint64) + y) % qs).astype(np.uint32)
    fx = oracle_mod.forward(x, logN, moduli, oracle_mod.PLAIN)
    fy = oracle_mod.forward(y, logN, moduli, oracle_mod.PLAIN)
    fs = oracle_mod.forward(s, logN, moduli, oracle_mod.PLAIN)
    assert np.array_equal(fs, ((fx.astype(np.uint64) + fy) % qs).astype(np.uint32))
    assert np.array_equal(oracle_mod.inverse(fx, logN, moduli, oracle_mod.PLAIN), x)


def test_montgomery_variant_semantics(oracle_mod):
    """Forward output is Montgomery form, inverse takes Montgomery form and
    returns normal domain (src/v04_montgomery.cu:229-232, 567-580, 1082-1087)."""
    logN = 10
    moduli = DEFAULT_MODULI[:4]
    x = oracle_mod.fill_random_limbs(logN, moduli, 0)
    plain = oracle_mod.forward(x, logN, moduli, oracle_mod.PLAIN)
    mont = oracle_mod.forward(x, logN, moduli, oracle_mod.MONTGOMERY)
    for l, (q, _) in enumerate(moduli):
        assert np.array_equal(mont[l], (plain[l].astype(np.uint64) * (R32 % q) % q).astype(np.uint32))
    assert np.array_equal(oracle_mod.inverse(mont, logN, moduli, oracle_mod.MONTGOMERY), x)


def test_mt19937_known_answer(oracle_mod):
    import ctypes
    L = oracle_mod.lib()

    class MT(ctypes.Structure):
        _fields_ = [("mt", ctypes.c_uint32 * 624), ("idx", ctypes.c_int)]

    g = MT()
    L.oracle_mt_seed.argtypes = [ctypes.POINTER(MT), ctypes.c_uint32]
    L.oracle_mt_next.argtypes = [ctypes.POINTER(MT)]
    L.oracle_mt_next.restype = ctypes.c_uint32
    L.oracle_mt_seed(ctypes.byref(g), 5489)
    vals = [L.oracle_mt_next(ctypes.byref(g)) for _ in range(10000)]
    assert vals[0] == 3499211612          # first output of default-seeded MT19937
    assert vals[-1] == 4123659995         # C++ [rand.predef]: 10000th output


CPP_PROG = r"""
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <random>
int main(int argc, char** argv) {
    uint32_t q = (uint32_t)strtoul(argv[1], 0, 10);
    uint32_t seed = (uint32_t)strtoul(argv[2], 0, 10);
    int n = atoi(argv[3]);
    std::mt19937 gen(seed);
    std::uniform_int_distribution<uint32_t> dist(0, q - 1);
    for (int i = 0; i < n; ++i) printf("%u\n", dist(gen));
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with libstdc++")
@pytest.mark.parametrize("q,seed", [(1073479681, 0), (974258177, 1234), (3, 7), (2147483647, 0), (1 << 31, 5)])
def test_fill_random_matches_libstdcxx(oracle_mod, q, seed):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "r.cpp")
        exe = os.path.join(d, "r")
        with open(src, "w") as f:
            f.write(CPP_PROG)
        subprocess.run(["g++", "-O1", "-o", exe, src], check=True)
        n = 5000
        out = subprocess.run([exe, str(q), str(seed), str(n)], check=True, capture_output=True, text=True).stdout
        expect = [int(v) for v in out.split()]
    assert oracle_mod.fill_random(n, q, seed).tolist() == expect


# ---- planted mistakes: the checks above must reject a wrong oracle --------
def test_planted_mistake_natural_order_twiddles_is_caught(oracle_mod):
    """A twiddle table in natural instead of bit-reversed order (a plausible
    indexing slip) must fail the brute-force check."""
    logN, q, g = 5, 7681, 17
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    N = 1 << logN
    wrong = np.array([pow(rt["psi"], i, q) for i in range(N)], dtype=np.uint32)
    a = list(range(1, N + 1))
    got = oracle_mod.cpu_ntt_butterfly(np.array(a, dtype=np.uint32), logN, q, wrong)
    assert got.tolist() != mathref.negacyclic_eval_bitrev(a, q, rt["psi"])


def test_planted_mistake_cyclic_root_is_caught(oracle_mod):
    """Using omega (an N-th root: cyclic NTT) instead of psi must fail."""
    logN, q, g = 5, 7681, 17
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    tw = oracle_mod.twiddle_table(logN, q, rt["omega"])
    a = list(range(3, (1 << logN) + 3))
    got = oracle_mod.cpu_ntt_butterfly(np.array(a, dtype=np.uint32), logN, q, tw)
    assert got.tolist() != mathref.negacyclic_eval_bitrev(a, q, rt["psi"])


def test_planted_mistake_missing_normalisation_is_caught(oracle_mod):
    logN, q, g = 4, 97, 5
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    itw = oracle_mod.twiddle_table(logN, q, rt["psi_inv"])
    b = list(range(1 << logN))
    got = oracle_mod.cpu_intt_butterfly(np.array(b, dtype=np.uint32), logN, q, itw, 1)
    assert got.tolist() != mathref.negacyclic_interp_bitrev(b, q, rt["psi"])


def _mod_down_case(oracle_mod, logN=10):
    """RNS exact division, the identity ModDown rests on: with P coprime to
    every q and f = P^-1 mod q, (NTT(x + P y) - NTT(x)) * f = NTT(y)."""
    moduli = DEFAULT_MODULI[:6]
    P = 1000003 * 999983                       # two primes outside the moduli
    rng = np.random.default_rng(11)
    qs = [q for q, _ in moduli]
    x = np.array([rng.integers(0, q, 1 << logN) for q in qs], dtype=np.uint32)
    y = np.array([rng.integers(0, q, 1 << logN) for q in qs], dtype=np.uint32)
    xpy = np.array([(x[l].astype(object) + P * y[l].astype(object)) % q for l, q in enumerate(qs)],
                   dtype=np.uint32)
    f = [pow(P, -1, q) for q in qs]
    mont = oracle_mod.MONTGOMERY
    return (oracle_mod.forward(x, logN, moduli, mont), oracle_mod.forward(xpy, logN, moduli, mont), f, moduli,
            oracle_mod.forward(y, logN, moduli, mont))


def test_mod_down_exact_division(oracle_mod):
    ntt_x, sub, f, moduli, ntt_y = _mod_down_case(oracle_mod)
    assert np.array_equal(oracle_mod.mod_down(ntt_x, sub, f, moduli), ntt_y)
    # f = 1 is the plain difference, checked against Python integers
    d = oracle_mod.mod_down(ntt_x, sub, [1] * len(moduli), moduli)
    l, i, q = 3, 77, moduli[3][0]
    assert int(d[l, i]) == (int(sub[l, i]) - int(ntt_x[l, i])) % q


def test_mod_down_sub_range_and_padding(oracle_mod):
    """src2 limb range and padding (NTT_cheddar.cu:566-590), pinned by the same
    exact division: with padding g_l and sub limb S_l = NTT(x + P y)_l * g_l^-1
    for l in the range, the result is NTT(y)_l; outside the range it is
    (0 - NTT(x)_l) * f_l = NTT(-x P^-1)_l by linearity of the transform."""
    ntt_x, ntt_xpy, f, moduli, ntt_y = _mod_down_case(oracle_mod)
    L = len(moduli)
    qs = [q for q, _ in moduli]
    lo, hi = 1, 4
    g = [7919 + 13 * k for k in range(hi - lo)]
    sub = np.array([ntt_xpy[l].astype(object) * pow(g[l - lo], -1, qs[l]) % qs[l] for l in range(lo, hi)],
                   dtype=np.uint32)
    got = oracle_mod.mod_down(ntt_x, sub, f, moduli, sub_start=lo, sub_end=hi, sub_factors=g)
    x = oracle_mod.inverse(ntt_x, 10, moduli, oracle_mod.MONTGOMERY)     # back to x
    mx = np.array([(-(x[l].astype(object)) * f[l]) % qs[l] for l in range(L)], dtype=np.uint32)
    ntt_mx = oracle_mod.forward(mx, 10, moduli, oracle_mod.MONTGOMERY)
    for l in range(L):
        assert np.array_equal(got[l], ntt_y[l] if lo <= l < hi else ntt_mx[l]), l
    # a padding applied to the wrong sub limb (off by one) fails
    bad = oracle_mod.mod_down(ntt_x, sub, f, moduli, sub_start=lo, sub_end=hi, sub_factors=g[1:] + g[:1])
    assert not np.array_equal(bad[lo], ntt_y[lo])


def test_planted_mistake_mod_down_operand_order_is_caught(oracle_mod):
    """NTT - sub instead of sub - NTT (a sign slip) gives NTT(-y), not NTT(y)."""
    ntt_x, sub, f, moduli, ntt_y = _mod_down_case(oracle_mod)
    assert not np.array_equal(oracle_mod.mod_down(sub, ntt_x, f, moduli), ntt_y)


def test_twiddle_split_identity(oracle_mod):
    """The identity the on-the-fly twiddles rest on: with tw[i] = psi^brv(i),
    tw[i] = tw[i mod 64] * tw[64 (i div 64)] mod q for every i (brv splits over
    the disjoint bit fields of i), for the forward and the inverse table."""
    for logN, (q, g) in ((10, (12289, 11)), (13, (998244353, 3)), (16, (1073479681, 11))):
        rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
        assert rc == 0
        for root in (rt["psi"], rt["psi_inv"]):
            tw = oracle_mod.twiddle_table(logN, q, root).astype(np.uint64)
            i = np.arange(1 << logN)
            assert np.array_equal(tw, tw[i & 63] * tw[(i >> 6) << 6] % q)


def test_planted_mistake_twiddle_split_is_caught(oracle_mod):
    """An msb table indexed tw[i div 64] instead of tw[64 (i div 64)] fails."""
    logN, q, g = 11, 12289, 11
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    assert rc == 0
    tw = oracle_mod.twiddle_table(logN, q, rt["psi"]).astype(np.uint64)
    i = np.arange(1 << logN)
    assert np.array_equal(tw, tw[i & 63] * tw[(i >> 6) << 6] % q)
    assert not np.array_equal(tw, tw[i & 63] * tw[i >> 6] % q)


# (q, a quadratic non-residue) with q < 2^62, q = 1 mod 2^18, found with
# Python integers (Miller-Rabin over fixed bases, Euler's criterion)
def _primes64(count, logN=17, bound=1 << 62):
    def is_prime(n):
        if n < 2:
            return False
        for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            if n % p == 0:
                return n == p
        d, s = n - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
            x = pow(a, d, n)
            if x in (1, n - 1):
                continue
            for _ in range(s - 1):
                x = x * x % n
                if x == n - 1:
                    break
            else:
                return False
        return True
    step = 1 << (logN + 1)
    q = (bound - 1) // step * step + 1
    out = []
    while len(out) < count:
        if q < bound and is_prime(q):
            g = 2
            while pow(g, (q - 1) // 2, q) != q - 1:
                g += 1
            out.append((q, g))
        q -= step
    return out


def test_oracle64_equals_32bit_oracle_below_2_31(oracle_mod):
    """For q < 2^31 the 64-bit transforms equal the 32-bit oracle's plain variant."""
    logN = 11
    moduli = [(998244353, 3), (1073479681, 11), (12289, 11)]
    x = oracle_mod.fill_random_limbs(logN, moduli, 5)
    f32 = oracle_mod.forward(x, logN, moduli, oracle_mod.PLAIN)
    f64 = oracle_mod.forward64(x.astype(np.uint64), logN, moduli)
    assert np.array_equal(f64, f32.astype(np.uint64))
    assert np.array_equal(oracle_mod.inverse64(f64, logN, moduli), x.astype(np.uint64))


def test_oracle64_brute_force_and_roundtrip(oracle_mod):
    """62-bit primes: evaluations at psi^(2 brv(k) + 1) with Python integers
    (N = 2^6), and forward/inverse roundtrip at N = 2^12."""
    moduli = _primes64(2)
    logN = 6
    N = 1 << logN
    rng = random.Random(3)
    for q, g in moduli:
        psi = oracle_mod.psi64(logN, q, g)
        assert psi and pow(psi, N, q) == q - 1
        a = [rng.randrange(q) for _ in range(N)]
        got = oracle_mod.forward64(np.array([a], dtype=np.uint64), logN, [(q, g)])[0]
        for k in range(N):
            e = 2 * oracle_mod.bit_reverse(k, logN) + 1
            assert int(got[k]) == sum(a[j] * pow(psi, e * j, q) for j in range(N)) % q
    logN = 12
    x = np.array([[rng.randrange(q) for _ in range(1 << logN)] for q, _ in moduli], dtype=np.uint64)
    assert np.array_equal(oracle_mod.inverse64(oracle_mod.forward64(x, logN, moduli), logN, moduli), x)


def test_planted_mistake_oracle64_non_primitive_root_is_refused(oracle_mod):
    """A root whose psi is not primitive (a quadratic residue: 4) is refused."""
    q, _ = _primes64(1)[0]
    assert oracle_mod.psi64(10, q, 4) == 0
