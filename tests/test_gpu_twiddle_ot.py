"""GPU parity of the fused engine with on-the-fly twiddles
(ntt_plan_set_twiddle_mode(NTT_TWIDDLE_OT)): every fused size, both
directions and domains, lazy-capable and canonical-only moduli, ragged limb
counts and several operands per launch, bit-exact against the oracle
(integer arithmetic: tolerance 0) and against the table mode."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(oracle_mod):
    import ntt_amd
    from tests import gpu_util
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    return ntt_amd, gpu_util, oracle_mod


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("logN,limbs", [(10, 67), (11, 5), (12, 9), (13, 3), (14, 6), (15, 3), (16, 4)])
@pytest.mark.parametrize("domain", [0, 1])
@pytest.mark.parametrize("arith", [1, 0])
def test_ot_matches_oracle(env, logN, limbs, domain, arith):
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    count = 2
    x = np.stack([gu.random_operand(logN, moduli, 300 + k) for k in range(count)])
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, domain)
    plan.set_engine(ntt_amd.ENGINE_FUSED)
    plan.set_twiddle_mode(ntt_amd.TWIDDLE_OT)
    plan.set_fused_arith(arith)
    for direction in (0, 1):
        assert plan.fused_variant_full(count, direction) == (6, arith == 1, 16, True)
    A = gu.to_dev(x)
    _print("case ot", logN, limbs, domain, arith)
    plan.forward(A, count)
    gu.sync()
    fwd = gu.to_host(A).reshape(x.shape)
    for k in range(count):
        assert np.array_equal(fwd[k], orc.forward(x[k], logN, moduli, domain)), k
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(x.shape), x)
    # the inverse alone, on the oracle's forward output
    B = gu.to_dev(fwd)
    plan.inverse(B, count)
    gu.sync()
    for k in range(count):
        assert np.array_equal(gu.to_host(B).reshape(x.shape)[k], orc.inverse(fwd[k], logN, moduli, domain)), k


def test_ot_canonical_moduli_and_table_mode_agree(env):
    """default_moduli() has q >= 2^30 (canonical forward OT kernel); the OT
    and table modes give identical words."""
    ntt_amd, gu, orc = env
    logN = 16
    moduli = [tuple(m) for m in ntt_amd.default_moduli()]
    x = gu.random_operand(logN, moduli, 17)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, 1)
    plan.set_engine(ntt_amd.ENGINE_FUSED)
    a, b = gu.to_dev(x), gu.to_dev(x)
    plan.set_twiddle_mode(ntt_amd.TWIDDLE_TABLE)
    assert not plan.fused_variant_full(1, 0)[3]
    plan.forward(a)
    plan.set_twiddle_mode(ntt_amd.TWIDDLE_OT)
    _print("case ot canonical")
    plan.forward(b)
    gu.sync()
    got = gu.to_host(b).reshape(x.shape)
    assert np.array_equal(got, gu.to_host(a).reshape(x.shape))
    assert np.array_equal(got, orc.forward(x, logN, moduli, 1))
    plan.inverse(b)
    gu.sync()
    assert np.array_equal(gu.to_host(b).reshape(x.shape), x)
    with pytest.raises(ntt_amd.NTTError):
        plan.set_twiddle_mode(2)
