"""CPU check of the fused limb kernel's schedule (tests/fused_model.py) for
every supported size: bounds, exchange hazards, LDS ownership and the
transform against the oracle."""
import numpy as np
import pytest

import ntt_amd
from tests import fused_model as fm


@pytest.mark.parametrize("eb", [6, 7])
@pytest.mark.parametrize("n", [10, 11, 12, 13, 14, 15, 16])
def test_fused_schedule(oracle_mod, n, eb):
    lpw = 1 << (16 - n)
    nl = min(lpw, 3)
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(n, nl)]
    rng = np.random.default_rng(n)
    xs, tws, itws, qs, dinv = [], [], [], [], []
    for q, g in moduli:
        rc, rt = oracle_mod.make_modulus_runtime(n, q, g)
        tws.append([int(v) for v in oracle_mod.twiddle_table(n, q, rt["psi"])])
        itws.append([int(v) for v in oracle_mod.twiddle_table(n, q, rt["psi_inv"])])
        qs.append(q)
        dinv.append(rt["degree_inv"])
        xs.append([int(v) for v in rng.integers(0, q, 1 << n)])
    fwd = fm.run(xs, tws, qs, n, inv=False, eb=eb)
    ref = oracle_mod.forward(np.array(xs, np.uint32), n, moduli, 0)
    assert np.array_equal(np.array(fwd, np.uint32), ref)
    back = fm.run(fwd, itws, qs, n, inv=True, scale=dinv, eb=eb)
    assert back == xs
    # the lazy 32-bit arithmetic: plain-domain forward, inverse with the
    # Montgomery multiplier to_mont(N^-1)
    lfwd = fm.run(xs, tws, qs, n, inv=False, lazy=True, eb=eb)
    assert np.array_equal(np.array(lfwd, np.uint32), ref)
    mdinv = [(d << 32) % q for d, q in zip(dinv, qs)]
    assert fm.run(lfwd, itws, qs, n, inv=True, scale=mdinv, lazy=True, eb=eb) == xs
    # Montgomery-domain forward: x R^2 on load gives to_mont(NTT(x))
    r2 = [pow(2, 64, q) for q in qs]
    mfwd = fm.run(xs, tws, qs, n, inv=False, scale=r2, lazy=True, eb=eb)
    assert mfwd == [[(v << 32) % q for v in row] for row, q in zip(ref.tolist(), qs)]


@pytest.mark.parametrize("eb", [6, 7])
def test_fused_windows_cover_all_bits(eb):
    for n in range(10, 17):
        for inv in (False, True):
            bits = [b for k, r in enumerate(fm.stage_bits(n, inv, eb)) for b in (x + fm.windows(n, inv, eb)[k] for x in r)]
            assert sorted(bits) == list(range(n))
            assert bits == (sorted(bits) if inv else sorted(bits, reverse=True))
