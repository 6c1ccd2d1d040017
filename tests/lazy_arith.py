"""Exact 32-bit models of the kernels' modular arithmetic (ntt_amd/csrc/arith.hpp).

Each operation is computed with Python integers and reduced modulo 2^32 the
way the GPU does, with the range claims of the kernels asserted: a value that
leaves its stated range, or a product that would overflow 64 bits, fails.
"""

M32 = (1 << 32) - 1


def umin(a, b):
    return a if a <= b else b


def mont_lazy(a, b, q, qinv):
    assert 0 <= a <= M32 and 0 <= b < q
    t = a * b
    assert t < (1 << 32) * q
    m = ((t & M32) * qinv) & M32
    s = t + m * q
    assert s < (1 << 64)
    u = s >> 32
    assert u < 2 * q
    return u


def mont(a, b, q, qinv):
    u = mont_lazy(a, b, q, qinv)
    return umin(u, (u - q) & M32)


def canon_4q(x, q):
    assert 0 <= x < 4 * q
    x = umin(x, (x - 2 * q) & M32)
    return umin(x, (x - q) & M32)


def ct_lazy(x, y, w, q, qinv):
    assert 0 <= x < 4 * q and 0 <= y < 4 * q and 4 * q <= M32 + 1
    q2 = 2 * q
    u = umin(x, (x - q2) & M32)
    v = mont_lazy(y, w, q, qinv)
    x2, y2 = (u + v) & M32, (u + q2 - v) & M32
    assert x2 < 4 * q and y2 < 4 * q
    return x2, y2


def gs_lazy(x, y, w, q, qinv):
    assert 0 <= x < 2 * q and 0 <= y < 2 * q
    q2 = 2 * q
    s = (x + y) & M32
    d = (x + q2 - y) & M32
    x2 = umin(s, (s - q2) & M32)
    y2 = mont_lazy(d, w, q, qinv)
    assert x2 < 2 * q
    return x2, y2


def qinv_of(q):
    return (-pow(q, -1, 1 << 32)) & M32


def to_mont(x, q):
    return (x << 32) % q
