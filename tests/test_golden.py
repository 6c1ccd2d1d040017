"""Golden digests of the C1-C3 and reference workloads (tests/golden/digests.json,
made by tests/golden/make_digests.py): the oracle must keep reproducing them
(CPU), and the GPU path must match them without consulting the oracle (GPU)."""
import hashlib
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "digests.json")))
CASES = [k for k in GOLD if not k.startswith("_")]


def _digests(a: np.ndarray):
    return [hashlib.sha256(np.ascontiguousarray(r, dtype="<u4").tobytes()).hexdigest() for r in a]


def _check(a: np.ndarray, rec):
    assert _digests(a) == [r["sha256"] for r in rec]
    for row, r in zip(a, rec):
        assert [int(v) for v in row[:16]] == r["head"] and [int(v) for v in row[-16:]] == r["tail"]


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_golden(oracle_mod, case):
    g = GOLD[case]
    moduli = [tuple(m) for m in g["moduli"]]
    x = oracle_mod.fill_random_limbs(g["logN"], moduli, g["seed"])
    _check(x, g["input"])
    _check(oracle_mod.forward(x, g["logN"], moduli, g["domain"]), g["forward"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_matches_golden(oracle_mod, case):
    """Inputs come from the oracle's fill_random (checked against the fixture
    first); the forward output is compared with the fixture digests only."""
    import ntt_amd
    from tests import gpu_util as gu
    if not gu.gpu_available():
        pytest.skip("no GPU")
    g = GOLD[case]
    moduli = [tuple(m) for m in g["moduli"]]
    x = oracle_mod.fill_random_limbs(g["logN"], moduli, g["seed"])
    _check(x, g["input"])
    print("case golden", case, flush=True)
    for engine in (ntt_amd.ENGINE_FUSED, ntt_amd.ENGINE_TWO_PASS):
        if engine == ntt_amd.ENGINE_FUSED and g["logN"] > 16:
            continue
        ctx = ntt_amd.Context(0)
        plan = ctx.plan(g["logN"], moduli, g["domain"])
        plan.set_engine(engine)
        A = gu.to_dev(x)
        plan.forward(A)
        gu.sync()
        _check(gu.to_host(A).reshape(x.shape), g["forward"])
        plan.inverse(A)
        gu.sync()
        _check(gu.to_host(A).reshape(x.shape), g["input"])
