"""Two primitive roots of one modulus in one context.

The reference accepts any primitive_root g whose psi = g^((q-1)/2N) has order
exactly 2N (src/v04_montgomery.cu:681-718), and its twiddle tables are powers
of that psi (v04:720-727).  For q = 1073479681 at N = 2^16, g = 11 and g = 17
both pass and give psi = 1070907127 and psi = 223030499.  A twiddle cache
keyed by q alone would hand the second plan the first plan's tables; every
result here is compared with the oracle for its own root, on both engines,
through a batch and through the server, all sharing one context."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Q = 1073479681
ROOTS = (11, 17)
LOGN = 16


@pytest.fixture(scope="module")
def env(oracle_mod):
    import ntt_amd
    from tests import gpu_util
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    return ntt_amd, gpu_util, oracle_mod


def _print(*a):
    print(*a, flush=True)


def _moduli(root, limbs=2):
    # the same (q, root) on every limb: limb l uses table l of the plan
    return [(Q, root)] * limbs


def test_psi_differs_between_roots(oracle_mod):
    for logN in (LOGN, 17):
        rts = [oracle_mod.make_modulus_runtime(logN, Q, g) for g in ROOTS]
        assert all(rc == 0 for rc, _ in rts)
        assert len({rt["psi"] for _, rt in rts}) == 2


@pytest.mark.parametrize("engine_name", ["ENGINE_FUSED", "ENGINE_TWO_PASS"])
@pytest.mark.parametrize("domain", [1, 0])
def test_plans_of_two_roots_share_context(env, engine_name, domain):
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    plans, xs = [], []
    for g in ROOTS:
        m = _moduli(g)
        p = ctx.plan(LOGN, m, domain)
        p.set_engine(getattr(ntt_amd, engine_name))
        plans.append((p, m))
        xs.append(gu.random_operand(LOGN, m, 100 + g))
    # one table per (q, root, direction)
    assert ctx.twiddle_entries == 2 * len(ROOTS)
    for (p, m), x in zip(plans, xs):
        _print("case roots plan", engine_name, domain, m[0])
        A = gu.to_dev(x)
        p.forward(A)
        gu.sync()
        fwd = gu.to_host(A).reshape(x.shape)
        assert np.array_equal(fwd, orc.forward(x, LOGN, m, domain)), f"forward root {m[0][1]}"
        p.inverse(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), x), f"roundtrip root {m[0][1]}"
        # the inverse alone against the oracle's inverse of the same input
        B = gu.to_dev(fwd)
        p.inverse(B)
        gu.sync()
        assert np.array_equal(gu.to_host(B).reshape(x.shape), orc.inverse(fwd, LOGN, m, domain))


def test_batch_with_two_roots(env):
    """logN 16 jobs run on one fused launch per direction, logN 17 jobs
    (q = 1 mod 2^18 holds for this q) on the two-pass kernels."""
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    jobs, ref = [], []
    for logN in (LOGN, 17):
        for k, g in enumerate(ROOTS):
            for direction in (0, 1):
                m = _moduli(g, 3)
                x = gu.random_operand(logN, m, 200 + 10 * k + direction + logN)
                buf = gu.to_dev(x)
                jobs.append((buf, logN, m, direction, 1))
                ref.append(orc.forward(x, logN, m, 1) if direction == 0 else orc.inverse(x, logN, m, 1))
    _print("case roots batch")
    b = ctx.batch(jobs)
    b.run()
    gu.sync()
    for k, (job, exp) in enumerate(zip(jobs, ref)):
        got = gu.to_host(job[0]).reshape(exp.shape)
        assert np.array_equal(got, exp), f"batch job {k} logN {job[1]} root {job[2][0][1]} direction {job[3]}"
    assert ctx.twiddle_entries == 2 * 2 * len(ROOTS)


def test_server_with_two_roots(env):
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    srv = ctx.server(hbm_budget_bytes=256 << 20, slots=2, max_jobs_per_launch=16)
    reqs = []
    for k, g in enumerate(ROOTS):
        for direction in (0, 1):
            m = _moduli(g, 2)
            x = gu.random_operand(LOGN, m, 300 + 10 * k + direction)
            rid = srv.submit(LOGN, len(m), direction, 1, moduli=m, src_host=x)
            reqs.append((rid, m, direction, x))
    _print("case roots server")
    for rid, m, direction, x in reqs:
        srv.wait(rid)
        got = srv.fetch(rid, LOGN, len(m))
        exp = orc.forward(x, LOGN, m, 1) if direction == 0 else orc.inverse(x, LOGN, m, 1)
        assert np.array_equal(got, exp), f"server root {m[0][1]} direction {direction}"
    srv.close()
    assert ctx.twiddle_entries == 2 * len(ROOTS)
