"""GPU tests of the serving loop (ntt_server_*): each result is bit-identical
alone and in a concurrent mix (tolerance 0, integer arithmetic) and equal to
the oracle's; admission respects the HBM budget; a Poisson trace completes
with batching and overlapping launches."""
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


def _seeded_input(ntt_amd, gu, logN, limbs, seed):
    """The operand the server generates for (shape, seed), read back."""
    from ntt_amd import workloads
    m = workloads.moduli_for(logN, limbs)
    qd = gu.to_dev(np.array([q for q, _ in m], np.uint32))
    buf = ntt_amd.DeviceBuffer(4 * (limbs << logN))
    ntt_amd.device.fill_uniform(buf, qd, limbs, logN, 1, seed=seed)
    gu.sync()
    return gu.to_host(buf).reshape(limbs, 1 << logN), m


def _oracle(orc, x, logN, m, direction, domain=1):
    return orc.forward(x, logN, m, domain) if direction == 0 else orc.inverse(x, logN, m, domain)


def test_server_alone_vs_mix(env):
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    shape = dict(logN=14, limbs=5, direction=0, seed=42)
    srv = ctx.server(hbm_budget_bytes=512 << 20, slots=4, max_jobs_per_launch=32)
    _print("case server-alone")
    rid = srv.submit(shape["logN"], shape["limbs"], shape["direction"], seed=shape["seed"])
    srv.wait(rid)
    alone = srv.fetch(rid, shape["logN"], shape["limbs"])
    x, m = _seeded_input(ntt_amd, gu, shape["logN"], shape["limbs"], shape["seed"])
    assert np.array_equal(alone, _oracle(orc, x, shape["logN"], m, 0))

    _print("case server-mix")
    rng = np.random.default_rng(3)
    mix = []
    for k in range(80):
        if k == 37:
            mix.append(dict(shape))
        else:
            mix.append(dict(logN=int(rng.integers(10, 18)), limbs=int(rng.integers(1, 9)),
                            direction=int(rng.integers(0, 2)), seed=1000 + k))
    ids = [srv.submit(j["logN"], j["limbs"], j["direction"], seed=j["seed"]) for j in mix]
    for rid in ids:
        srv.wait(rid)
    st = srv.stats()
    _print("server stats", st)
    in_mix = srv.fetch(ids[37], shape["logN"], shape["limbs"])
    assert np.array_equal(in_mix, alone)
    # oracle checks across sizes and directions, including two-pass sizes
    checked = set()
    for j, rid in zip(mix, ids):
        key = (j["logN"], j["direction"])
        if key in checked or rid == ids[37]:
            continue
        checked.add(key)
        got = srv.fetch(rid, j["logN"], j["limbs"], release=False)
        xj, mj = _seeded_input(ntt_amd, gu, j["logN"], j["limbs"], j["seed"])
        assert np.array_equal(got, _oracle(orc, xj, j["logN"], mj, j["direction"])), j
    assert st["launches"] < len(mix) + 1          # requests were batched
    srv.close()


def test_server_host_and_device_sources(env):
    ntt_amd, gu, orc = env
    from ntt_amd import workloads
    ctx = ntt_amd.Context(0)
    srv = ctx.server(hbm_budget_bytes=64 << 20, slots=2, max_jobs_per_launch=8)
    logN, limbs = 12, 3
    m = workloads.moduli_for(logN, limbs)
    x = gu.random_operand(logN, m, 8)
    d = gu.to_dev(x)
    _print("case server-sources")
    a = srv.submit(logN, limbs, 0, src_host=x)
    b = srv.submit(logN, limbs, 0, src_dev=d, moduli=m)
    c = srv.submit(logN, limbs, 1, domain=0, src_host=x)
    for rid in (a, b, c):
        srv.wait(rid)
    ra, rb = srv.fetch(a, logN, limbs), srv.fetch(b, logN, limbs)
    assert np.array_equal(ra, orc.forward(x, logN, m, 1))
    assert np.array_equal(rb, ra)
    assert np.array_equal(srv.fetch(c, logN, limbs), orc.inverse(x, logN, m, 0))
    srv.close()


def test_server_budget_admission(env):
    """Budget of 4 operands: the fifth waits in the queue until a result is
    released; an operand larger than the budget is refused."""
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    logN, limbs = 16, 4                     # 1 MiB per operand
    op = 4 * (limbs << logN)
    srv = ctx.server(hbm_budget_bytes=4 * op, slots=2, max_jobs_per_launch=8)
    _print("case server-budget")
    ids = [srv.submit(logN, limbs, 0, seed=k) for k in range(5)]
    for rid in ids[:4]:
        srv.wait(rid)
    assert srv.state(ids[4]) == ntt_amd.REQ_QUEUED
    st = srv.stats()
    assert st["arena_peak_bytes"] <= 4 * op
    srv.release(ids[0])
    srv.wait(ids[4])
    x, m = _seeded_input(ntt_amd, gu, logN, limbs, 4)
    assert np.array_equal(srv.fetch(ids[4], logN, limbs), orc.forward(x, logN, m, 1))
    with pytest.raises(ntt_amd.NTTError):
        srv.submit(17, 40, 0, seed=1)       # 20 MiB > 4 MiB budget
    srv.close()


def test_server_poisson_trace(env):
    """A short served stream (logN 13..17, L 4..64 log-uniform) at a high
    rate: all requests complete, launches batch several requests and several
    launches are in flight at once."""
    ntt_amd, gu, orc = env
    from ntt_amd import workloads
    ctx = ntt_amd.Context(0)
    srv = ctx.server(hbm_budget_bytes=2 << 30, slots=4, max_jobs_per_launch=64)
    reqs = workloads.serving_stream(5, 400, rate_per_s=200000.0)
    _print("case server-trace")
    lat, wall = srv.run_trace(reqs)
    st = srv.stats()
    _print("trace", st, "wall", wall, "p50", np.percentile(lat, 50), "p99", np.percentile(lat, 99))
    assert st["completed"] == 400 and np.all(lat > 0)
    assert st["mean_jobs_per_launch"] > 1.0
    assert st["max_slots_busy"] >= 2
    srv.close()


def test_server_fifo_admission_order(env):
    """Budget for one operand at a time: requests are admitted and finished
    strictly in arrival order, and every result still matches the oracle."""
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    logN, limbs = 13, 2
    op = 4 * (limbs << logN)
    srv = ctx.server(hbm_budget_bytes=op, slots=3, max_jobs_per_launch=4)
    _print("case server-fifo")
    ids = [srv.submit(logN, limbs, k % 2, seed=300 + k) for k in range(6)]
    for k, rid in enumerate(ids):
        srv.wait(rid)
        # the next request cannot have been admitted while this result is held
        if k + 1 < len(ids):
            assert srv.state(ids[k + 1]) == ntt_amd.REQ_QUEUED
        x, m = _seeded_input(ntt_amd, gu, logN, limbs, 300 + k)
        assert np.array_equal(srv.fetch(rid, logN, limbs), _oracle(orc, x, logN, m, k % 2))
    assert srv.stats()["arena_peak_bytes"] <= op
    srv.close()


def test_server_argument_errors(env):
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    with pytest.raises(ntt_amd.NTTError):
        ctx.server(hbm_budget_bytes=1 << 20, slots=0)
    srv = ctx.server(hbm_budget_bytes=1 << 20, slots=1, max_jobs_per_launch=2)
    for bad in (dict(logN=0, limbs=1), dict(logN=12, limbs=0), dict(logN=25, limbs=1)):
        with pytest.raises(ntt_amd.NTTError):
            srv.submit(bad["logN"], bad["limbs"])
    with pytest.raises(ntt_amd.NTTError):
        srv.submit(12, 1, direction=2)
    with pytest.raises(ntt_amd.NTTError):
        srv.wait(12345)                       # unknown id
    srv.close()


def test_server_serving_stream_results(env):
    """The served-stream workload itself (logN U{13..17}, L log-uniform on
    [4, 64], direction Bernoulli(1/2), Poisson arrivals, seeded operands):
    a 2000-request prefix submitted at its arrival times with results kept
    (about 8 GB resident).  Every result is checked by the opposite transform
    on the device against its regenerated seeded operand, and the first
    request of every (logN, direction) against the oracle."""
    import time
    ntt_amd, gu, orc = env
    from ntt_amd import workloads
    reqs = workloads.serving_stream(0, 2000, rate_per_s=50000.0)
    total = sum(4 * r.job.words for r in reqs)
    ctx = ntt_amd.Context(0)
    srv = ctx.server(hbm_budget_bytes=total + (256 << 20), slots=4, max_jobs_per_launch=64)
    _print("case server-c5 submit", len(reqs), "requests", total >> 20, "MiB")
    ids = []
    t0 = time.perf_counter()
    for r in reqs:
        while time.perf_counter() - t0 < r.arrival_s:
            pass
        ids.append(srv.submit(r.job.logN, r.job.limbs, r.job.direction, seed=r.seed))
    for rid in ids:
        srv.wait(rid)
    st = srv.stats()
    _print("case server-c5 done", st)
    assert st["completed"] == len(reqs)
    assert st["mean_jobs_per_launch"] > 1.0            # requests were batched
    plans, qdev = {}, {}
    oracle_done = set()
    for k, (r, rid) in enumerate(zip(reqs, ids)):
        logN, limbs, direction = r.job.logN, r.job.limbs, r.job.direction
        shape = (logN, limbs)
        m = workloads.moduli_for(logN, limbs)
        if shape not in plans:
            plans[shape] = ctx.plan(logN, m, ntt_amd.MONTGOMERY)
            qdev[shape] = gu.to_dev(np.array([q for q, _ in m], np.uint32))
        words = limbs << logN
        res = ntt_amd.DeviceBuffer(4 * words)
        srv.fetch_device(rid, res, release=False)
        x = ntt_amd.DeviceBuffer(4 * words)
        ntt_amd.device.fill_uniform(x, qdev[shape], limbs, logN, 1, seed=r.seed)
        gu.sync()
        xh = gu.to_host(x).reshape(limbs, 1 << logN)
        if (logN, direction) not in oracle_done:
            oracle_done.add((logN, direction))
            exp = _oracle(orc, xh, logN, m, direction)
            assert np.array_equal(gu.to_host(res).reshape(xh.shape), exp), f"request {k} vs oracle"
        # the opposite transform brings the result back to the operand
        if direction == ntt_amd.FORWARD:
            plans[shape].inverse(res)
        else:
            plans[shape].forward(res)
        gu.sync()
        assert np.array_equal(gu.to_host(res).reshape(xh.shape), xh), \
            f"request {k}: logN {logN} limbs {limbs} direction {direction}"
        srv.release(rid)
    assert len(oracle_done) == 10
    srv.close()
