"""CPU check of the split decomposition the logN-17 fused kernels use
(fused_kernel.hpp, SPLIT): a transform of 2^(n+1) words is the cross-half
stage (twiddle tw[1]) followed by two independent 2^n-word transforms whose
stage twiddles are those of the 2^(n+1)-word table at (2 + h) m + g for level
m and group g; the inverse is the two sub-inverses (same indexing on the
inverse table), then the cross-half stage folded with N^-1:
((T0 + T1) N^-1, (T0 - T1) itw[1] N^-1).  Plain integers on the oracle's own
tables, compared with the oracle's transforms at several sizes (including the
kernels' 2^16-word halves at logN 17), so a wrong table offset, a swapped half
or a wrong sign in the combine fails."""
import numpy as np
import pytest

import ntt_amd


def _sub_forward(b, tw, h, q):
    # CT stages of the reference loop (src/v04_montgomery.cu:742-759) over
    # len(b) words, twiddle tw[(2 + h) m + g]
    b = list(b)
    M = len(b)
    t, m = M, 1
    while m < M:
        t //= 2
        for g in range(m):
            w = tw[(2 + h) * m + g]
            for j in range(2 * g * t, 2 * g * t + t):
                u, v = b[j], b[j + t] * w % q
                b[j], b[j + t] = (u + v) % q, (u - v) % q
        m *= 2
    return b


def _sub_inverse(b, itw, h, q):
    # GS stages of the reference loop (src/v04_montgomery.cu:761-784), no
    # normalisation, twiddle itw[(2 + h) m + g]
    b = list(b)
    M = len(b)
    t, m = 1, M // 2
    while m >= 1:
        for g in range(m):
            w = itw[(2 + h) * m + g]
            for j in range(2 * g * t, 2 * g * t + t):
                u, v = b[j], b[j + t]
                b[j], b[j + t] = (u + v) % q, (u - v) * w % q
        t *= 2
        m //= 2
    return b


@pytest.mark.parametrize("logN", [5, 9, 12, 17])
def test_split_decomposition_matches_oracle(oracle_mod, logN):
    q, g = [tuple(m) for m in ntt_amd.generate_moduli(logN, 1)][0]
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    assert rc == 0
    tw = [int(v) for v in oracle_mod.twiddle_table(logN, q, rt["psi"])]
    itw = [int(v) for v in oracle_mod.twiddle_table(logN, q, rt["psi_inv"])]
    N, M = 1 << logN, 1 << (logN - 1)
    x = [int(v) for v in np.random.default_rng(logN).integers(0, q, N)]
    ref = oracle_mod.forward(np.array([x], np.uint32), logN, [(q, g)], 0)[0].tolist()
    w1 = tw[1]
    y0 = [(x[j] + w1 * x[j + M]) % q for j in range(M)]
    y1 = [(x[j] - w1 * x[j + M]) % q for j in range(M)]
    assert _sub_forward(y0, tw, 0, q) + _sub_forward(y1, tw, 1, q) == ref
    # inverse of the oracle's forward output, back to x
    t0 = _sub_inverse(ref[:M], itw, 0, q)
    t1 = _sub_inverse(ref[M:], itw, 1, q)
    ninv = rt["degree_inv"]
    back = [(a + b) * ninv % q for a, b in zip(t0, t1)] + [(a - b) * itw[1] * ninv % q for a, b in zip(t0, t1)]
    assert back == x
    # a planted slip (the halves' table offsets swapped) is caught
    assert _sub_forward(y0, tw, 1, q) + _sub_forward(y1, tw, 0, q) != ref
