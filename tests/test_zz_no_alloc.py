"""GPU test of the allocation-free batch launch (ntt_batch_run_static).

It runs the kernels of ntt_batch_run with the job tables in the library's
fixed device array instead of an allocation.  Its outputs must equal those of
ntt_batch_run (the existing entry point) bit for bit on the same jobs, over
two calls in a row, the second with fewer jobs and limbs than the first (the
array keeps the first call's tables; nothing of them may leak into the
second).  Jobs cover fused sizes (10..17, 17 on the split kernels) and the
two-pass size 9, both directions.  Sizes beyond the fixed tables are refused with
NTT_ERR_SCRATCH_LIMIT before anything is launched."""
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


def _jobs(ntt_amd, gu, spec, seed):
    """spec: (logN, limbs, direction) per job; per job its operand, the
    reference-style device tables of its moduli and two device copies."""
    from ntt_amd import workloads
    out = []
    for k, (logN, limbs, direction) in enumerate(spec):
        m = workloads.moduli_for(logN, limbs)
        t = ntt_amd.make_tables(logN, m, ntt_amd.MONTGOMERY)
        x = gu.random_operand(logN, m, seed + k)
        out.append(dict(
            logN=logN, limbs=limbs, direction=direction, moduli=m, x=x,
            a=gu.to_dev(x), b=gu.to_dev(x),
            tw=gu.to_dev(t.inv_twiddles if direction else t.twiddles),
            q=gu.to_dev(t.moduli), qi=gu.to_dev(t.q_inv),
            sc=gu.to_dev(t.degree_inv if direction else t.r2)))
    return out


def _run_both(env, spec, seed):
    ntt_amd, gu, orc = env
    jobs = _jobs(ntt_amd, gu, spec, seed)
    ctx = ntt_amd.Context(0)
    batch = ctx.batch([(j["a"], j["logN"], j["moduli"], j["direction"], ntt_amd.MONTGOMERY) for j in jobs])
    batch.run()
    ntt_amd.batch_run_static([(j["b"], j["tw"], j["q"], j["qi"], j["sc"], j["limbs"], j["logN"], j["direction"])
                              for j in jobs])
    gu.sync()
    for k, j in enumerate(jobs):
        ref = gu.to_host(j["a"]).reshape(j["x"].shape)
        got = gu.to_host(j["b"]).reshape(j["x"].shape)
        assert np.array_equal(got, ref), f"job {k}: logN {j['logN']} limbs {j['limbs']} dir {j['direction']}"
    # and one job of each direction against the oracle
    for d in (0, 1):
        j = next(j for j in jobs if j["direction"] == d)
        exp = orc.forward(j["x"], j["logN"], j["moduli"], 1) if d == 0 else orc.inverse(j["x"], j["logN"], j["moduli"], 1)
        assert np.array_equal(gu.to_host(j["b"]).reshape(j["x"].shape), exp)
    batch.close()


def test_static_batch_matches_batch_two_calls(env):
    first = [(16, 16, 0), (16, 5, 1), (14, 9, 0), (12, 33, 1), (10, 3, 0), (13, 2, 1),
             (17, 6, 0), (17, 4, 1), (9, 3, 0), (15, 7, 1), (16, 2, 0), (11, 40, 0)]
    second = [(12, 3, 0), (16, 1, 1), (17, 2, 1)]
    _print("case static batch: first call", len(first), "jobs")
    _run_both(env, first, 1000)
    _print("case static batch: second, smaller call", len(second), "jobs")
    _run_both(env, second, 2000)


def test_static_batch_limits(env):
    ntt_amd, gu, orc = env
    j = _jobs(ntt_amd, gu, [(10, 1, 0)], 7)[0]
    one = (j["b"], j["tw"], j["q"], j["qi"], j["sc"], 1, 10, 0)
    with pytest.raises(ntt_amd.NTTError) as e:
        ntt_amd.batch_run_static([one] * 1025)
    assert e.value.status == 12
    big = (j["b"], j["tw"], j["q"], j["qi"], j["sc"], 16385, 10, 0)
    with pytest.raises(ntt_amd.NTTError) as e:
        ntt_amd.batch_run_static([big])
    assert e.value.status == 12
    # nothing was launched: the operand is unchanged
    gu.sync()
    assert np.array_equal(gu.to_host(j["b"]).reshape(j["x"].shape), j["x"])
