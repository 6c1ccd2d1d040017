"""GPU parity of batches of uneven transforms (ntt_batch_*): every job
bit-exact against the oracle, eager and replayed from a captured hipGraph,
plus the full-size uneven workload (256 jobs, logN 12..17, 1..48 limbs,
about 1 GB) through the forward/inverse roundtrip and oracle spot checks."""
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


def _expected(orc, x, job_logN, moduli, direction, domain):
    if direction == 0:
        return orc.forward(x, job_logN, moduli, domain)
    return orc.inverse(x, job_logN, moduli, domain)


def _mixed_jobs(ntt_amd, gu, seed, sizes, max_limbs):
    from ntt_amd import workloads
    rng = np.random.default_rng(seed)
    jobs = []
    for k, logN in enumerate(sizes):
        limbs = int(rng.integers(1, max_limbs + 1))
        direction = int(rng.integers(0, 2))
        domain = int(rng.integers(0, 2))
        moduli = workloads.moduli_for(logN, limbs)
        x = gu.random_operand(logN, moduli, seed * 1000 + k)
        jobs.append(dict(logN=logN, moduli=moduli, direction=direction, domain=domain, x=x, buf=gu.to_dev(x)))
    return jobs


def _check_all(env, jobs):
    ntt_amd, gu, orc = env
    for k, j in enumerate(jobs):
        got = gu.to_host(j["buf"]).reshape(j["x"].shape)
        exp = _expected(orc, j["x"], j["logN"], j["moduli"], j["direction"], j["domain"])
        assert np.array_equal(got, exp), f"job {k}: logN {j['logN']} limbs {len(j['moduli'])} dir {j['direction']}"


def _reset(jobs):
    for j in jobs:
        j["buf"].upload(j["x"])


def test_batch_eager_mixed(env):
    """Sizes 8..17 (fused groups for 10..16, split fused groups for 17,
    two-pass for 8 and 9), ragged limb counts, both directions and domains,
    several jobs per group."""
    ntt_amd, gu, orc = env
    sizes = [10, 11, 12, 13, 14, 15, 16, 17, 8, 9, 12, 12, 16, 16, 13, 17, 10, 15]
    jobs = _mixed_jobs(ntt_amd, gu, 5, sizes, 6)
    ctx = ntt_amd.Context(0)
    b = ctx.batch([(j["buf"], j["logN"], j["moduli"], j["direction"], j["domain"]) for j in jobs])
    info = b.info()
    _print("case batch-eager", info)
    assert info["two_pass_jobs"] == 2
    s = ntt_amd.Stream()
    b.run(s)
    s.synchronize()
    _check_all(env, jobs)


def test_batch_graph_replays(env):
    """Capture once, replay three times with eager work and synchronizations
    between replays; inputs are restored before each replay and every replay
    is checked against the oracle."""
    ntt_amd, gu, orc = env
    sizes = [12, 13, 16, 17, 12, 14, 16, 10, 11, 17]
    jobs = _mixed_jobs(ntt_amd, gu, 7, sizes, 5)
    ctx = ntt_amd.Context(0)
    b = ctx.batch([(j["buf"], j["logN"], j["moduli"], j["direction"], j["domain"]) for j in jobs])
    s = ntt_amd.Stream()
    b.capture(s)
    other_m = ntt_amd.workloads.moduli_for(12, 3)
    other = gu.random_operand(12, other_m, 99)
    obuf = gu.to_dev(other)
    oplan = ctx.plan(12, other_m, 1)
    for r in range(3):
        _reset(jobs)
        gu.sync()
        _print("case batch-replay", r)
        b.replay(s)
        s.synchronize()
        _check_all(env, jobs)
        oplan.forward(obuf, 1, s)     # unrelated eager work between replays
        oplan.inverse(obuf, 1, s)
        s.synchronize()
    assert np.array_equal(gu.to_host(obuf).reshape(other.shape), other)


def test_batch_set_data(env):
    """Re-pointing fused-group jobs (a 2^13 job and a 2^17 job on the split
    kernels) and a two-pass job (2^9) keeps the captured graph (all read their
    operand addresses from device memory); replays give the oracle's result on
    the new operands."""
    ntt_amd, gu, orc = env
    sizes = [13, 17, 9, 13]
    jobs = _mixed_jobs(ntt_amd, gu, 11, sizes, 4)
    ctx = ntt_amd.Context(0)
    b = ctx.batch([(j["buf"], j["logN"], j["moduli"], j["direction"], j["domain"]) for j in jobs])
    s = ntt_amd.Stream()
    b.replay(s)
    s.synchronize()
    _check_all(env, jobs)
    assert b.info()["two_pass_jobs"] == 1
    for k in (0, 1, 2):
        j = jobs[k]
        j["x"] = gu.random_operand(j["logN"], j["moduli"], 500 + k)
        j["buf"] = gu.to_dev(j["x"])
        b.set_data(k, j["buf"])
    _reset(jobs)
    gu.sync()
    _print("case batch-set-data")
    b.replay(s)
    s.synchronize()
    _check_all(env, jobs)


def test_batch_full_uneven_workload(env):
    """256 jobs: logN ~ U{12..17}, L ~ U{1..48}, direction ~ Bernoulli(1/2),
    Montgomery domain, seed 0 (about 1 GB of operands).  Forward/inverse
    roundtrip of every job through two batches (the second with every
    direction flipped), plus oracle checks of one job per size."""
    ntt_amd, gu, orc = env
    from ntt_amd import workloads
    wl = workloads.batched_uneven(0)
    ctx = ntt_amd.Context(0)
    qs = {}
    bufs, xs = [], []
    for k, job in enumerate(wl):
        m = workloads.moduli_for(job.logN, job.limbs)
        buf = ntt_amd.DeviceBuffer(4 * job.words)
        qd = qs.get((job.logN, job.limbs))
        if qd is None:
            qd = qs[(job.logN, job.limbs)] = gu.to_dev(np.array([q for q, _ in m], np.uint32))
        ntt_amd.device.fill_uniform(buf, qd, job.limbs, job.logN, 1, seed=k)
        bufs.append(buf)
    gu.sync()
    for buf in bufs:
        xs.append(gu.to_host(buf))
    spec = [(bufs[k], j.logN, workloads.moduli_for(j.logN, j.limbs), j.direction, 1) for k, j in enumerate(wl)]
    flip = [(b_, n, m, 1 - d, dom) for (b_, n, m, d, dom) in spec]
    fwd, back = ctx.batch(spec), ctx.batch(flip)
    s = ntt_amd.Stream()
    _print("case batch-full", fwd.info())
    fwd.replay(s)
    s.synchronize()
    seen = set()
    for k, j in enumerate(wl):
        if j.logN in seen:
            continue
        seen.add(j.logN)
        m = workloads.moduli_for(j.logN, j.limbs)
        x = xs[k].reshape(j.limbs, 1 << j.logN)
        got = gu.to_host(bufs[k]).reshape(x.shape)
        assert np.array_equal(got, _expected(orc, x, j.logN, m, j.direction, 1)), f"job {k}"
    back.replay(s)
    s.synchronize()
    for k in range(len(wl)):
        assert np.array_equal(gu.to_host(bufs[k]), xs[k]), f"roundtrip job {k}"


def test_batch_many_small_jobs(env):
    """600 one-limb jobs of sizes 10..13 in both directions (a few fused
    groups with hundreds of job descriptors each), oracle-checked."""
    ntt_amd, gu, orc = env
    from ntt_amd import workloads
    rng = np.random.default_rng(21)
    jobs = []
    for k in range(600):
        logN = int(rng.integers(10, 14))
        m = workloads.moduli_for(logN, 1)
        x = gu.random_operand(logN, m, 4000 + k)
        jobs.append(dict(logN=logN, moduli=m, direction=int(rng.integers(0, 2)), domain=1, x=x, buf=gu.to_dev(x)))
    ctx = ntt_amd.Context(0)
    b = ctx.batch([(j["buf"], j["logN"], j["moduli"], j["direction"], j["domain"]) for j in jobs])
    _print("case batch-many", b.info())
    s = ntt_amd.Stream()
    b.replay(s)
    s.synchronize()
    _check_all(env, jobs)
