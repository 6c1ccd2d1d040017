"""GPU parity: the HIP kernels against the CPU oracle, bit-exact.

Inputs are seeded uniform words in [0, q); every comparison is exact equality
of all L*N words (the reference's `validate`, src/v04_montgomery.cu:799-806),
plus the NTT -> INTT roundtrip (v04:1109-1132).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "moduli.json")))
M16 = [tuple(x) for x in FIX["M16"]]
M48 = [tuple(x) for x in FIX["M48"]]


@pytest.fixture(scope="module")
def env(oracle_mod):
    import ntt_amd
    from tests import gpu_util
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    return ntt_amd, gpu_util, oracle_mod


def _print(*a):
    print(*a, flush=True)


def run_dropin(env, logN, moduli, domain, x, cfg=None):
    ntt_amd, gu, orc = env
    d = gu.DropIn(logN, moduli, domain)
    A = gu.to_dev(x)
    d.forward(A, cfg)
    gu.sync()
    fwd = gu.to_host(A).reshape(x.shape)
    d.inverse(A, cfg)
    gu.sync()
    back = gu.to_host(A).reshape(x.shape)
    return fwd, back


@pytest.mark.parametrize("domain", [1, 0])
def test_reference_workload_default_moduli(env, domain):
    """The reference's own run: DefaultNTTLaunchConfig<16>, default_moduli(),
    fill_random with SEED 1234 (v04:28, 872-934)."""
    ntt_amd, gu, orc = env
    logN = 16
    moduli = [tuple(m) for m in ntt_amd.default_moduli()]
    x = orc.fill_random_limbs(logN, moduli, 1234)
    ref = orc.forward(x, logN, moduli, domain)
    for cfg in (ntt_amd.reference_launch_config(16), ntt_amd.default_launch_config(16)):
        _print("case reference-workload", domain, cfg)
        fwd, back = run_dropin(env, logN, moduli, domain, x, cfg)
        assert np.array_equal(fwd, ref)
        assert np.array_equal(back, x)


@pytest.mark.parametrize("domain", [1, 0])
def test_baseline_c1_c2_m16_seed0(env, domain):
    """N=2^16, 16 limbs, M16 moduli, seed 0: plain (C1) and Montgomery (C2)."""
    ntt_amd, gu, orc = env
    x = orc.fill_random_limbs(16, M16, 0)
    _print("case C1/C2 domain", domain)
    fwd, back = run_dropin(env, 16, M16, domain, x)
    assert np.array_equal(fwd, orc.forward(x, 16, M16, domain))
    assert np.array_equal(back, x)


def test_baseline_c3_large(env):
    """N=2^17, 48 limbs (M48), seed 0, Montgomery."""
    ntt_amd, gu, orc = env
    x = orc.fill_random_limbs(17, M48, 0)
    _print("case C3")
    fwd, back = run_dropin(env, 17, M48, 1, x)
    assert np.array_equal(fwd, orc.forward(x, 17, M48, 1))
    assert np.array_equal(back, x)


@pytest.mark.parametrize("logN", list(range(1, 16)))
def test_all_sizes_default_config(env, logN):
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, 3)]
    x = gu.random_operand(logN, moduli, logN)
    for domain in (1, 0):
        _print("case size", logN, domain)
        fwd, back = run_dropin(env, logN, moduli, domain, x)
        assert np.array_equal(fwd, orc.forward(x, logN, moduli, domain)), (logN, domain)
        assert np.array_equal(back, x)


@pytest.mark.parametrize("logN", [18, 19, 20])
def test_largest_sizes(env, logN):
    """The sizes above the baseline workloads up to the largest accepted
    (ntt_max_log_n): drop-in launchers (default split) and plans, both
    domains, against the oracle; one size past it is refused."""
    ntt_amd, gu, orc = env
    assert ntt_amd.lib().ntt_max_log_n() == 20
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, 2)]
    x = gu.random_operand(logN, moduli, logN)
    ctx = ntt_amd.Context(0)
    for domain in (1, 0):
        _print("case large size", logN, domain)
        exp = orc.forward(x, logN, moduli, domain)
        fwd, back = run_dropin(env, logN, moduli, domain, x)
        assert np.array_equal(fwd, exp), (logN, domain)
        assert np.array_equal(back, x)
        plan = ctx.plan(logN, moduli, domain)
        A = gu.to_dev(x)
        plan.forward(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), exp)
        plan.inverse(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), x)
    if logN == 20:
        with pytest.raises(ntt_amd.NTTError):
            ctx.plan(21, [tuple(m) for m in ntt_amd.generate_moduli(21, 1)])


def _configs(logN):
    """A spread of valid phase configs for logN (both pass orders and shapes)."""
    out = []
    for r1 in range(0, logN + 1):
        r2 = logN - r1
        for s1 in range(1, 6):
            for s2 in range(1, 6):
                if (r1 and s1 > r1) or (r2 and s2 > r2) or (not r1 and s1 != 1) or (not r2 and s2 != 1):
                    continue
                for w1 in (0, 2, 4, 6):
                    for w2 in (0, 1, 2):
                        p1 = (r1, s1, w1) if r1 else (0, 0, 0)
                        p2 = (r2, s2, w2) if r2 else (0, 0, 0)
                        out.append((p1, p2))
    return out


@pytest.mark.parametrize("logN", [6, 10, 12, 14])
def test_config_sweep(env, logN):
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, 2)]
    x = gu.random_operand(logN, moduli, 100 + logN)
    ref = {d: orc.forward(x, logN, moduli, d) for d in (0, 1)}
    rng = np.random.default_rng(logN)
    cands = [c for c in _configs(logN)
             if ntt_amd.validate_config(logN, ntt_amd.TransformConfig.make(*c)) == 0]
    pick = [cands[i] for i in rng.choice(len(cands), size=min(40, len(cands)), replace=False)]
    for p1, p2 in pick:
        cfg = ntt_amd.TransformConfig.make(p1, p2)
        d = int(rng.integers(0, 2))
        _print("case sweep", logN, p1, p2, d)
        fwd, back = run_dropin(env, logN, moduli, d, x, cfg)
        assert np.array_equal(fwd, ref[d]), (p1, p2, d)
        assert np.array_equal(back, x), (p1, p2, d)


@pytest.mark.parametrize("logN,limbs", [(12, 5), (13, 7), (15, 3), (16, 16), (17, 9)])
@pytest.mark.parametrize("domain", [1, 0])
def test_plans_match_oracle(env, logN, limbs, domain):
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    plan = ctx.plan(logN, moduli, domain)
    x = gu.random_operand(logN, moduli, limbs)
    A = gu.to_dev(x)
    _print("case plan", logN, limbs, domain)
    plan.forward(A)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(x.shape), orc.forward(x, logN, moduli, domain))
    plan.inverse(A)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(x.shape), x)
    # the cache holds one table per (q, logN, direction)
    assert ctx.twiddle_entries == 2 * limbs


@pytest.mark.parametrize("col", [(8, 4, 5), (8, 4, 6), (8, 5, 5), (8, 5, 6), (9, 4, 5), (9, 5, 5), (9, 5, 6)])
@pytest.mark.parametrize("domain", [1, 0])
def test_logN17_column_shapes(env, col, domain):
    """The column passes compiled for one geometry (logN 17, R1 = 8, 9, W = 5,
    6; picked whenever a two-pass launch at logN 17 uses that split), forward
    and inverse, over several operands of several limbs."""
    ntt_amd, gu, orc = env
    logN, limbs, count = 17, 3, 2
    ctx = ntt_amd.Context(0)
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    row = (17 - col[0], 4, 1)
    cfg = ntt_amd.TransformConfig.make(col, row, row, col)
    plan = ctx.plan(logN, moduli, domain, cfg)
    plan.set_engine(ntt_amd.ENGINE_TWO_PASS)
    xs = [gu.random_operand(logN, moduli, 170 + 7 * k + domain) for k in range(count)]
    A = gu.to_dev(np.stack(xs))
    _print("case col17", col, domain)
    plan.forward(A, count)
    gu.sync()
    got = gu.to_host(A).reshape(count, limbs, -1)
    for k in range(count):
        assert np.array_equal(got[k], orc.forward(xs[k], logN, moduli, domain)), (col, k)
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(count, limbs, -1), np.stack(xs))


def test_plan_many_operands(env):
    ntt_amd, gu, orc = env
    logN, limbs, count = 14, 6, 5
    ctx = ntt_amd.Context(0)
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    plan = ctx.plan(logN, moduli, 1)
    xs = [gu.random_operand(logN, moduli, 7 + k) for k in range(count)]
    A = gu.to_dev(np.stack(xs))
    _print("case plan-many")
    plan.forward(A, count)
    gu.sync()
    got = gu.to_host(A).reshape(count, limbs, -1)
    for k in range(count):
        assert np.array_equal(got[k], orc.forward(xs[k], logN, moduli, 1))
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(count, limbs, -1), np.stack(xs))


def test_edge_values(env):
    """All-zero, all q-1 and single-impulse operands (linearity anchors)."""
    ntt_amd, gu, orc = env
    logN = 12
    moduli = M16[:4]
    qs = np.array([m[0] for m in moduli], dtype=np.uint32)[:, None]
    N = 1 << logN
    cases = [np.zeros((4, N), np.uint32), np.broadcast_to(qs - 1, (4, N)).copy()]
    imp = np.zeros((4, N), np.uint32)
    imp[:, 0] = 1
    cases.append(imp)
    for x in cases:
        for domain in (0, 1):
            fwd, back = run_dropin(env, logN, moduli, domain, x)
            assert np.array_equal(fwd, orc.forward(x, logN, moduli, domain))
            assert np.array_equal(back, x)


def test_montgomery_input_above_q_is_reduced(env):
    """The fused R^2 conversion accepts any 32-bit word (v04:229-232): the
    forward transform of x and of x + q agree."""
    ntt_amd, gu, orc = env
    logN = 12
    moduli = M16[:2]
    x = gu.random_operand(logN, moduli, 3)
    qs = np.array([m[0] for m in moduli], dtype=np.uint64)[:, None]
    xp = (x.astype(np.uint64) + qs).astype(np.uint32)
    a, _ = run_dropin(env, logN, moduli, 1, x)
    b, _ = run_dropin(env, logN, moduli, 1, xp)
    assert np.array_equal(a, b)


def test_launch_errors_are_status_codes(env):
    ntt_amd, gu, orc = env
    moduli = M16[:2]
    d = gu.DropIn(12, moduli, 1)
    A = gu.to_dev(np.zeros((2, 1 << 12), np.uint32))
    bad = ntt_amd.TransformConfig.make((5, 3, 0), (5, 3, 0))
    with pytest.raises(ntt_amd.NTTError) as e:
        d.forward(A, bad)
    assert e.value.status == 5
    with pytest.raises(ntt_amd.NTTError) as e:
        ntt_amd.launch_ntt(A, d.tw, d.q, d.q_inv, d.r2, 0, 12)
    assert e.value.status == 1
    ctx = ntt_amd.Context(0)
    with pytest.raises(ntt_amd.NTTError) as e:
        ctx.plan(17, [(974258177, 3)])        # q != 1 mod 2^18
    assert e.value.status == 2


def test_autotune_result_is_valid_and_exact(env, tmp_path):
    ntt_amd, gu, orc = env
    logN, limbs, count = 12, 4, 8
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    ctx = ntt_amd.Context(0)
    _print("case autotune")
    r = ctx.autotune(logN, moduli, count=count, reps=2)
    cfg = ntt_amd.TransformConfig.make(*r["cfg"][:2], *r["cfg"][2:])
    assert ntt_amd.validate_config(logN, cfg) == 0
    assert r["candidates"] > 10 and r["fwd_ms"] > 0 and r["inv_ms"] > 0
    plan = ctx.plan(logN, moduli, 1)          # picks the tuned entry for count operands
    xs = np.stack([gu.random_operand(logN, moduli, 40 + k) for k in range(count)])
    A = gu.to_dev(xs)
    plan.forward(A, count)
    gu.sync()
    got = gu.to_host(A).reshape(xs.shape)
    for k in range(count):
        assert np.array_equal(got[k], orc.forward(xs[k], logN, moduli, 1))
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(xs.shape), xs)
    path = str(tmp_path / "tune.txt")
    ctx.save_tuning(path)
    ctx2 = ntt_amd.Context(0)
    ctx2.load_tuning(path)
    ctx2.save_tuning(path + "2")
    assert open(path).read() == open(path + "2").read()


@pytest.mark.parametrize("logN", [10, 11, 12, 13, 14, 15, 16])
@pytest.mark.parametrize("engine,eb,arith,logw", [(2, 7, 1, 16), (2, 7, 0, 16), (2, 6, 1, 16), (2, 6, 0, 16),
                                                  (2, 7, 1, 15), (2, 7, 0, 15), (2, 6, 1, 15), (2, 6, 0, 15),
                                                  (1, 0, -1, 16)])
def test_engines_all_sizes(env, logN, engine, eb, arith, logw):
    """Fused limb kernel (both window widths; lazy and canonical arithmetic in
    both directions; 2^16 or, for logN <= 15, 2^15 words per workgroup) and
    two-pass kernels, bit-exact; limb counts that leave the last fused work
    item partly empty.  The inverse is checked on its own input too."""
    ntt_amd, gu, orc = env
    lpw = 1 << (16 - logN)
    for limbs in sorted({1, 3, lpw + 1}):
        moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
        for domain in (1, 0):
            ctx = ntt_amd.Context(0)
            plan = ctx.plan(logN, moduli, domain)
            plan.set_engine(engine)
            plan.set_fused_shape(eb)
            plan.set_fused_arith(arith)
            plan.set_fused_words(logw)
            assert plan.engine(1) == engine
            if engine == 2:
                for direction in (0, 1):
                    assert plan.fused_variant(1, direction) == (eb, arith == 1)
            x = gu.random_operand(logN, moduli, 1000 + limbs)
            A = gu.to_dev(x)
            _print("case engine", engine, eb, arith, logw, logN, limbs, domain)
            plan.forward(A)
            gu.sync()
            assert np.array_equal(gu.to_host(A).reshape(x.shape), orc.forward(x, logN, moduli, domain))
            plan.inverse(A)
            gu.sync()
            assert np.array_equal(gu.to_host(A).reshape(x.shape), x)
            y = gu.random_operand(logN, moduli, 2000 + limbs)
            B = gu.to_dev(y)
            plan.inverse(B)
            gu.sync()
            assert np.array_equal(gu.to_host(B).reshape(y.shape), orc.inverse(y, logN, moduli, domain))


def test_fused_many_operands_period(env):
    """count operands as one job: limb l uses the table of limb l % limbs."""
    ntt_amd, gu, orc = env
    logN, limbs, count = 13, 5, 7
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, 1)
    plan.set_engine(2)
    plan.set_fused_shape(6)
    xs = np.stack([gu.random_operand(logN, moduli, 70 + k) for k in range(count)])
    A = gu.to_dev(xs)
    _print("case fused-many")
    plan.forward(A, count)
    gu.sync()
    got = gu.to_host(A).reshape(xs.shape)
    for k in range(count):
        assert np.array_equal(got[k], orc.forward(xs[k], logN, moduli, 1))
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(xs.shape), xs)


@pytest.mark.parametrize("engine", [2, 1])
def test_engines_canonical_arithmetic(env, engine):
    """default_moduli() has q >= 2^30, so plans use canonical (not lazy)
    arithmetic; both engines stay bit-exact."""
    ntt_amd, gu, orc = env
    logN = 16
    moduli = [tuple(m) for m in ntt_amd.default_moduli()]
    for domain in (1, 0):
        ctx = ntt_amd.Context(0)
        plan = ctx.plan(logN, moduli, domain)
        plan.set_engine(engine)
        x = orc.fill_random_limbs(logN, moduli, 1234)
        A = gu.to_dev(x)
        _print("case canonical engine", engine, domain)
        plan.forward(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), orc.forward(x, logN, moduli, domain))
        plan.inverse(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), x)


@pytest.mark.parametrize("logN,engine", [(12, 2), (12, 1), (16, 2), (16, 1), (17, 1)])
@pytest.mark.parametrize("domain", [1, 0])
def test_plan_constants_prologue_epilogue(env, logN, engine, domain):
    """Fused per-limb constant multiplies: forward = NTT(c * x), inverse =
    e * INTT(y), bit-exact against the oracle applied to the scaled operand."""
    ntt_amd, gu, orc = env
    limbs = 5
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    qs = np.array([q for q, _ in moduli], np.uint64)[:, None]
    rng = np.random.default_rng(logN * 10 + engine + domain)
    c = [int(rng.integers(0, q)) for q, _ in moduli]
    e = [int(rng.integers(0, 1 << 32)) for _ in moduli]      # reduced mod q by the library
    x = gu.random_operand(logN, moduli, 77)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, domain)
    plan.set_engine(engine)
    plan.set_constants(c, e)
    A = gu.to_dev(x)
    _print("case constants", logN, engine, domain)
    plan.forward(A)
    gu.sync()
    xs = (x.astype(np.uint64) * np.array(c, np.uint64)[:, None] % qs).astype(np.uint32)
    fwd = gu.to_host(A).reshape(x.shape)
    assert np.array_equal(fwd, orc.forward(xs, logN, moduli, domain))
    plan.inverse(A)
    gu.sync()
    ev = np.array([v % int(q) for v, (q, _) in zip(e, moduli)], np.uint64)[:, None]
    exp = (orc.inverse(fwd, logN, moduli, domain).astype(np.uint64) * ev % qs).astype(np.uint32)
    assert np.array_equal(gu.to_host(A).reshape(x.shape), exp)
    # back to plain transforms
    plan.set_constants(None, None)
    A.upload(x)
    plan.forward(A)
    plan.inverse(A)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(x.shape), x)


def test_graph_capture_of_plan_launches(env):
    """A restore copy plus forward and inverse plan launches captured with
    ntt_rt_graph_*: every replay (with eager work between replays) equals the
    eager result."""
    ntt_amd, gu, orc = env
    logN, limbs = 14, 6
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    x = gu.random_operand(logN, moduli, 31)
    src, A = gu.to_dev(x), gu.to_dev(x)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, 1)
    s = ntt_amd.Stream()
    exp = orc.forward(x, logN, moduli, 1)
    for engine in (ntt_amd.ENGINE_FUSED, ntt_amd.ENGINE_TWO_PASS):
        plan.set_engine(engine)
        with ntt_amd.device.Graph.capture(s) as g:
            A.copy_from(src, stream=s)
            plan.forward(A, 1, s)
        _print("case graph", engine)
        for r in range(3):
            A.zero()
            gu.sync()
            g.launch(s)
            s.synchronize()
            assert np.array_equal(gu.to_host(A).reshape(x.shape), exp), (engine, r)
            plan.inverse(A, 1, s)              # eager work between replays
            s.synchronize()
            assert np.array_equal(gu.to_host(A).reshape(x.shape), x)
        g.close()
