"""CPU check of the phase kernels' index schedule over every valid config at
small sizes (see tests/kernel_model.py): bounds, LDS ownership and the
transform it computes against the oracle."""
import itertools

import numpy as np
import pytest

import ntt_amd
from tests import kernel_model


def valid_configs(logN):
    out = []
    for r1 in range(0, logN + 1):
        r2 = logN - r1
        p1s = [(0, 0, 0)] if r1 == 0 else [(r1, s, w) for s in range(1, min(r1, 5) + 1) for w in range(0, r2 + 1)]
        p2s = [(0, 0, 0)] if r2 == 0 else [(r2, s, w) for s in range(1, min(r2, 5) + 1) for w in range(0, r1 + 1)]
        for p1, p2 in itertools.product(p1s, p2s):
            cfg = ntt_amd.TransformConfig.make(p1, p2)
            if ntt_amd.validate_config(logN, cfg) == 0:
                out.append(cfg.astuple())
    return out


@pytest.mark.parametrize("logN", [1, 2, 3, 4, 5, 6, 7])
def test_schedule_all_configs(oracle_mod, logN):
    q, g = [tuple(m) for m in ntt_amd.generate_moduli(logN, 1)][0]
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    tw = [int(v) for v in oracle_mod.twiddle_table(logN, q, rt["psi"])]
    itw = [int(v) for v in oracle_mod.twiddle_table(logN, q, rt["psi_inv"])]
    rng = np.random.default_rng(logN)
    x = [int(v) for v in rng.integers(0, q, 1 << logN)]
    ref = oracle_mod.forward(np.array([x], np.uint32), logN, [(q, g)], 0)[0].tolist()
    cfgs = valid_configs(logN)
    assert cfgs
    mdinv = (rt["degree_inv"] << 32) % q
    for cfg in cfgs:
        fwd = kernel_model.forward(x, q, tw, logN, cfg)
        assert fwd == ref, cfg
        assert kernel_model.inverse(fwd, q, itw, logN, cfg, rt["degree_inv"]) == x, cfg
        lf = kernel_model.forward(x, q, tw, logN, cfg, lazy=True)
        assert lf == ref, cfg
        assert kernel_model.inverse(lf, q, itw, logN, cfg, mdinv, lazy=True) == x, cfg


@pytest.mark.parametrize("logN,cfg", [
    (16, ((7, 4, 6), (9, 3, 0), (9, 3, 0), (7, 4, 6))),
    (16, ((7, 4, 4), (9, 3, 0), (9, 3, 0), (7, 4, 4))),
    (12, ((3, 3, 6), (9, 3, 0), (9, 3, 0), (3, 3, 6))),
    (10, ((0, 0, 0), (10, 3, 0), (10, 3, 0), (0, 0, 0))),
])
def test_schedule_default_shapes(oracle_mod, logN, cfg):
    q, g = [tuple(m) for m in ntt_amd.generate_moduli(logN, 1)][0]
    rc, rt = oracle_mod.make_modulus_runtime(logN, q, g)
    tw = [int(v) for v in oracle_mod.twiddle_table(logN, q, rt["psi"])]
    itw = [int(v) for v in oracle_mod.twiddle_table(logN, q, rt["psi_inv"])]
    x = [int(v) for v in np.random.default_rng(1).integers(0, q, 1 << logN)]
    ref = oracle_mod.forward(np.array([x], np.uint32), logN, [(q, g)], 0)[0].tolist()
    fwd = kernel_model.forward(x, q, tw, logN, cfg)
    assert fwd == ref
    assert kernel_model.inverse(fwd, q, itw, logN, cfg, rt["degree_inv"]) == x


def test_default_configs_are_modelled_for_every_size():
    for logN in range(1, 21):
        cfg = ntt_amd.default_launch_config(logN).astuple()
        (r1, s1, w1), (r2, s2, w2) = cfg[0], cfg[1]
        assert r1 + r2 == logN
