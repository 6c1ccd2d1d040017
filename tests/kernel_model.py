"""Host model of the phase kernels' schedule (ntt_amd/csrc/phase_kernels.hpp).

Replays, thread by thread and round by round, the index arithmetic the HIP
kernels use: workgroup -> tile, thread -> (lane, col/row), the S-bit window
of each round, the LDS word of every value and the twiddle index of every
butterfly.  It checks on the CPU that every global and LDS access is in
bounds, that each round's LDS reads and writes touch only the thread's own
words (the one-barrier-per-round scheme has no cross-thread hazard), that a
pass loads and stores every word exactly once, and that the arithmetic it
drives reproduces the transform.  Values are kept in the normal domain; the
kernels' Montgomery representation does not change any index.
"""
from __future__ import annotations


def lds_pad(i: int) -> int:
    return i + (i >> 5)


def lds_words(R: int, W: int) -> int:
    w = 1 << (R + W)
    return w + (w >> 5) + 1


def run_pass(A, tw, q, n, R, S, W, inv, col, mult=None, lazy=False, final_pass=False):
    """One pass over limb A (list of N ints, modified in place).

    lazy: the kernels' 32-bit lazy arithmetic (tests/lazy_arith.py) with
    Montgomery-form twiddles; mult is then the Montgomery multiplier, and
    `final_pass` marks the pass that must leave canonical values."""
    from tests import lazy_arith as lz
    qinv = lz.qinv_of(q)
    N = 1 << n
    E = 1 << S
    lane_bits = R - S
    tiles = 1 << (n - R - W)
    threads = 1 << (R - S + W)
    assert threads <= 1024 and tiles >= 1
    K = (R + S - 1) // S
    rem = R - S * (K - 1)
    C = 1 << (n - R)
    loaded, stored = [0] * N, [0] * N
    nlds = lds_words(R, W)
    for tile in range(tiles):
        lds = [None] * nlds
        regs = [[0] * E for _ in range(threads)]
        ctx = []
        for t in range(threads):
            if col:
                c, lane = t & ((1 << W) - 1), t >> W
                rib = g = 0
            else:
                lane, rib = t & ((1 << lane_bits) - 1), t >> lane_bits
                g = (tile << W) + rib
                c = 0
            ctx.append((lane, c, rib, g))

        def vidx(lane, lo, x):
            return ((lane >> lo) << (lo + S)) | (x << lo) | (lane & ((1 << lo) - 1))

        def gaddr(lane, c, g, v):
            a = v * C + (tile << W) + c if col else (g << R) + v
            assert 0 <= a < N
            return a

        def laddr(c, rib, v):
            a = lds_pad((v << W) + c if col else (rib << R) + v)
            assert 0 <= a < nlds
            return a

        def twbase(lane, g, lo):
            base = (n - R) if col else 0
            ext = lane >> lo
            if not col:
                ext |= g << (R - lo - S)
            return (1 << (n - S - base - lo)) + ext

        for k in range(K):
            last = k == K - 1
            lo = (R - S if last else S * k) if inv else (0 if last else R - S * (k + 1))
            owners = {}
            for t in range(threads):
                lane, c, rib, g = ctx[t]
                v = regs[t]
                addrs = []
                for x in range(E):
                    vi = vidx(lane, lo, x)
                    assert 0 <= vi < (1 << R)
                    if k == 0:
                        a = gaddr(lane, c, g, vi)
                        loaded[a] += 1
                        v[x] = A[a]
                        if not inv and mult is not None:
                            v[x] = lz.mont_lazy(v[x], mult, q, qinv) if lazy else v[x] * mult % q
                    else:
                        la = laddr(c, rib, vi)
                        assert lds[la] is not None, "read of an LDS word nobody wrote"
                        v[x] = lds[la]
                        addrs.append(la)
                # stages
                TB = twbase(lane, g, lo)
                js = range(S) if inv else range(S - 1, -1, -1)
                jlo, jhi = ((S - rem if last else 0), S - 1) if inv else (0, (rem - 1 if last else S - 1))
                for j in js:
                    if j < jlo or j > jhi:
                        continue
                    for x in range(E):
                        if x & (1 << j):
                            continue
                        ti = (TB << (S - 1 - j)) + (x >> (j + 1))
                        assert 0 < ti < N
                        w = tw[ti]
                        y = x + (1 << j)
                        if lazy:
                            v[x], v[y] = (lz.gs_lazy if inv else lz.ct_lazy)(v[x], v[y], lz.to_mont(w, q), q, qinv)
                        elif inv:
                            u, vv = v[x], v[y]
                            v[x], v[y] = (u + vv) % q, (u - vv) * w % q
                        else:
                            u, vv = v[x], v[y] * w % q
                            v[x], v[y] = (u + vv) % q, (u - vv) % q
                if not last:
                    for x in range(E):
                        la = laddr(c, rib, vidx(lane, lo, x))
                        if k > 0:
                            assert la in addrs, "write outside the words this thread read"
                        assert owners.setdefault(la, t) == t, "two threads own one LDS word"
                        lds[la] = v[x]
                else:
                    for x in range(E):
                        if lazy:
                            if inv and mult is not None:
                                v[x] = lz.mont(v[x], mult, q, qinv)
                            elif final_pass and not inv:
                                v[x] = lz.canon_4q(v[x], q)
                        elif inv and mult is not None:
                            v[x] = v[x] * mult % q
                        a = gaddr(lane, c, g, vidx(lane, lo, x))
                        stored[a] += 1
                        A[a] = v[x]
    assert all(c == 1 for c in loaded) and all(c == 1 for c in stored)


def forward(a, q, psi_tw, n, cfg, lazy=False, mult=None):
    """cfg = ((R1,S1,W1),(R2,S2,W2),...) as TransformConfig.astuple()."""
    A = list(a)
    (r1, s1, w1), (r2, s2, w2) = cfg[0], cfg[1]
    if r1:
        run_pass(A, psi_tw, q, n, r1, s1, w1, inv=False, col=True, mult=mult, lazy=lazy, final_pass=not r2)
    if r2:
        run_pass(A, psi_tw, q, n, r2, s2, w2, inv=False, col=False, mult=None if r1 else mult, lazy=lazy,
                 final_pass=True)
    return A


def inverse(a, q, ipsi_tw, n, cfg, ninv, lazy=False):
    """ninv: N^-1 (plain multiplier), or with lazy the Montgomery multiplier."""
    A = list(a)
    (r1, s1, w1), (r2, s2, w2) = cfg[2], cfg[3]
    if r1:
        run_pass(A, ipsi_tw, q, n, r1, s1, w1, inv=True, col=False, mult=None if r2 else ninv, lazy=lazy,
                 final_pass=not r2)
    if r2:
        run_pass(A, ipsi_tw, q, n, r2, s2, w2, inv=True, col=True, mult=ninv, lazy=lazy, final_pass=True)
    return A
