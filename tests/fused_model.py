"""Host model of the fused limb kernel's schedule (ntt_amd/csrc/fused_kernel.hpp).

A 1024-thread workgroup holds 64K words in registers: 2^(16-n) limbs of
N = 2^n words, 2^(n-6) threads x 64 words per limb.  A round performs the
pending stages of a 6-bit window of the word index in registers; between
rounds the words move through LDS in two halves (LDS holds 32K words), split
by a bit sigma that is a register bit of both windows.  Each round maps its
window bits to register bits through a transposition chosen so sigma sits in
the same register bit on both sides of an exchange: the registers a thread
reads in half h are exactly those it wrote out in half h (no register is
overwritten before it is saved).

This model replays that schedule with the same formulas and checks bounds,
the hazard rule above, LDS ownership, and the transform against the oracle.
"""
from __future__ import annotations

E_BITS = 6          # window width; the kernel is built for 6 (1024 threads) and 7 (512 threads)
WG_WORDS = 1 << 16  # words held by one workgroup


def nrounds(n: int, eb: int = None) -> int:
    eb = E_BITS if eb is None else eb
    return 2 if n <= 2 * eb - 1 else 3


def windows(n: int, inv: bool, eb: int = None):
    """Window low bits of the rounds, in execution order."""
    eb = E_BITS if eb is None else eb
    if nrounds(n, eb) == 2:
        return [0, n - eb] if inv else [n - eb, 0]
    return [0, eb - 1, n - eb] if inv else [n - eb, n - 2 * eb + 1, 0]


def stage_bits(n: int, inv: bool, eb: int = None):
    """Per round, the window-relative bits whose stages run, in order."""
    eb = E_BITS if eb is None else eb
    ws = windows(n, inv, eb)
    out = []
    for k, lo in enumerate(ws):
        if not inv:
            top = lo + eb - 1 if k == 0 else ws[k - 1] - 1
            out.append([b - lo for b in range(top, lo - 1, -1)])
        else:
            bot = lo if k == 0 else ws[k - 1] + eb
            out.append([b - lo for b in range(bot, lo + eb)])
    return out


def sigmas(n: int, inv: bool, eb: int = None):
    eb = E_BITS if eb is None else eb
    ws = windows(n, inv, eb)
    return [(ws[k] if not inv else ws[k] + eb - 1) for k in range(len(ws) - 1)]


def perms(n: int, inv: bool, eb: int = None):
    """perm[k][window bit] = register bit, per round."""
    eb = E_BITS if eb is None else eb
    ws = windows(n, inv, eb)
    sg = sigmas(n, inv, eb)
    out = [list(range(eb))]
    for k in range(1, len(ws)):
        s = sg[k - 1]
        ra = out[k - 1][s - ws[k - 1]]
        jb = s - ws[k]
        assert 0 <= jb < eb, "split bit must be a register bit of both windows"
        p = list(out[0])
        # transposition jb <-> ra
        p[jb], p[ra] = ra, jb
        out.append(p)
    return out


def reg_of(p, xw: int) -> int:
    r = 0
    for b in range(len(p)):
        if (xw >> b) & 1:
            r |= 1 << p[b]
    return r


def vidx(lo: int, u: int, x: int, eb: int = None) -> int:
    eb = E_BITS if eb is None else eb
    return ((u >> lo) << (lo + eb)) | (x << lo) | (u & ((1 << lo) - 1))


def compress(i: int, s: int) -> int:
    return ((i >> (s + 1)) << s) | (i & ((1 << s) - 1))


def pad_shift(eb: int) -> int:
    """One LDS padding word per 2^pad_shift words (fused_kernel.hpp)."""
    return 6 if eb == 7 else 5


def lds_slot_stride(n: int, eb: int = None) -> int:
    eb = E_BITS if eb is None else eb
    half = 1 << (n - 1)
    return half + (half >> pad_shift(eb))


def run(limbs_data, tws, qs, n: int, inv: bool, scale=None, lazy=False, eb: int = None):
    """limbs_data: list of 2^(16-n) limbs (lists), modified copies returned.

    lazy=False: canonical arithmetic on normal-domain values (scale is a plain
    multiplier).  lazy=True: the kernels' 32-bit lazy arithmetic
    (tests/lazy_arith.py) with Montgomery-form twiddles; scale is the
    Montgomery multiplier applied on load (forward) or store (inverse)."""
    from tests import lazy_arith as la
    eb = E_BITS if eb is None else eb
    E = 1 << eb
    THREADS = WG_WORDS >> eb
    N = 1 << n
    g = n - eb
    lpw = 1 << (16 - n)
    assert len(limbs_data) <= lpw
    A = [list(a) for a in limbs_data]
    ws, sb, pm, sg = windows(n, inv, eb), stage_bits(n, inv, eb), perms(n, inv, eb), sigmas(n, inv, eb)
    assert sum(len(s) for s in sb) == n
    stride = lds_slot_stride(n, eb)
    ps = pad_shift(eb)
    lds_words = lpw * stride
    assert lds_words <= 33 * 1024 + 64
    regs = [[None] * E for _ in range(THREADS)]
    valid = [(t >> g) < len(A) for t in range(THREADS)]
    for k, lo in enumerate(ws):
        p = pm[k]
        if k == 0:
            for t in range(THREADS):
                if not valid[t]:
                    continue
                slot, u = t >> g, t & ((1 << g) - 1)
                for xw in range(E):
                    i = vidx(lo, u, xw, eb)
                    assert 0 <= i < N
                    v = A[slot][i]
                    if not inv and scale is not None:
                        v = la.mont_lazy(v, scale[slot], qs[slot], la.qinv_of(qs[slot])) if lazy \
                            else v * scale[slot] % qs[slot]
                    regs[t][reg_of(p, xw)] = v
        # stages
        for t in range(THREADS):
            if not valid[t]:
                continue
            slot, u = t >> g, t & ((1 << g) - 1)
            q = qs[slot]
            qinv = la.qinv_of(q)
            tb = (1 << (n - eb - lo)) + (u >> lo)
            for j in sb[k]:
                for xw in range(E):
                    if (xw >> j) & 1:
                        continue
                    ti = (tb << (eb - 1 - j)) + (xw >> (j + 1))
                    assert 0 < ti < N
                    w = tws[slot][ti]
                    a, b = reg_of(p, xw), reg_of(p, xw | (1 << j))
                    x, y = regs[t][a], regs[t][b]
                    if lazy:
                        wm = la.to_mont(w, q)
                        regs[t][a], regs[t][b] = (la.gs_lazy if inv else la.ct_lazy)(x, y, wm, q, qinv)
                    elif not inv:
                        y = y * w % q
                        regs[t][a], regs[t][b] = (x + y) % q, (x - y) % q
                    else:
                        regs[t][a], regs[t][b] = (x + y) % q, (x - y) * w % q
        if k == len(ws) - 1:
            break
        # exchange to round k+1 in two halves
        s = sg[k]
        lo_b, pb = ws[k + 1], pm[k + 1]
        ja, jb = s - lo, s - lo_b
        ra = p[ja]
        assert pb[jb] == ra
        for h in (0, 1):
            lds = [None] * lds_words
            written = [set() for _ in range(THREADS)]
            for t in range(THREADS):
                if not valid[t]:
                    continue
                slot, u = t >> g, t & ((1 << g) - 1)
                for xw in range(E):
                    if ((xw >> ja) & 1) != h:
                        continue
                    i = vidx(lo, u, xw, eb)
                    assert ((i >> s) & 1) == h
                    c = compress(i, s)
                    a = slot * stride + c + (c >> ps)
                    assert lds[a] is None, "two writers of one LDS word"
                    lds[a] = regs[t][reg_of(p, xw)]
                    written[t].add(reg_of(p, xw))
            for t in range(THREADS):
                if not valid[t]:
                    continue
                slot, u = t >> g, t & ((1 << g) - 1)
                for xw in range(E):
                    if ((xw >> jb) & 1) != h:
                        continue
                    r = reg_of(pb, xw)
                    assert r in written[t], "register read before its value was saved"
                    i = vidx(lo_b, u, xw, eb)
                    c = compress(i, s)
                    a = slot * stride + c + (c >> ps)
                    assert lds[a] is not None, "read of an unwritten LDS word"
                    regs[t][r] = lds[a]
    lo = ws[-1]
    p = pm[-1]
    stored = [[0] * N for _ in A]
    for t in range(THREADS):
        if not valid[t]:
            continue
        slot, u = t >> g, t & ((1 << g) - 1)
        for xw in range(E):
            i = vidx(lo, u, xw, eb)
            v = regs[t][reg_of(p, xw)]
            if lazy:
                v = la.mont(v, scale[slot], qs[slot], la.qinv_of(qs[slot])) if inv else la.canon_4q(v, qs[slot])
            elif inv and scale is not None:
                v = v * scale[slot] % qs[slot]
            A[slot][i] = v
            stored[slot][i] += 1
    assert all(c == 1 for s_ in stored for c in s_)
    return A
