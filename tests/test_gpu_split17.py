"""GPU parity of the split fused kernels: a limb of 2^17 words per workgroup,
transformed as two 2^16-word halves in turn (the forward computes the
cross-half stage on load and parks half 1 in the output; the inverse parks
half 0's result there and combines it with half 1's).  Bit-exact against the
oracle (integer arithmetic: tolerance 0): both arithmetics (lazy, canonical),
both domains, limb counts 1..8, several operands per call, repeated launches,
edge values, the prologue/epilogue constants, hipGraph replays, and agreement
with the two-pass engine at 48 limbs.  The out-of-place forms at logN 17 run
in tests/test_gpu_moddown.py on both engines."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LOGN = 17


@pytest.fixture(scope="module")
def env(oracle_mod):
    import ntt_amd
    from tests import gpu_util
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    return ntt_amd, gpu_util, oracle_mod


def _print(*a):
    print(*a, flush=True)


def _plan(ntt_amd, moduli, domain, arith=-1):
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(LOGN, moduli, domain)
    plan.set_engine(ntt_amd.ENGINE_FUSED)
    plan.set_fused_arith(arith)
    return ctx, plan


@pytest.mark.parametrize("limbs,count", [(1, 1), (3, 1), (8, 1), (2, 3)])
@pytest.mark.parametrize("arith", [1, 0])
@pytest.mark.parametrize("domain", [1, 0])
def test_split_matches_oracle(env, limbs, count, arith, domain):
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, limbs)]
    ctx, plan = _plan(ntt_amd, moduli, domain, arith)
    assert plan.engine(count) == ntt_amd.ENGINE_FUSED
    xs = np.stack([gu.random_operand(LOGN, moduli, 300 + 11 * k + limbs) for k in range(count)])
    A = gu.to_dev(xs)
    _print("case split", limbs, count, arith, domain)
    plan.forward(A, count)
    gu.sync()
    got = gu.to_host(A).reshape(xs.shape)
    for k in range(count):
        assert np.array_equal(got[k], orc.forward(xs[k], LOGN, moduli, domain)), k
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(xs.shape), xs)
    # the inverse on its own input, three launches in a row
    for r in range(3):
        ys = np.stack([gu.random_operand(LOGN, moduli, 900 + 13 * k + r) for k in range(count)])
        B = gu.to_dev(ys)
        plan.inverse(B, count)
        gu.sync()
        got = gu.to_host(B).reshape(ys.shape)
        for k in range(count):
            assert np.array_equal(got[k], orc.inverse(ys[k], LOGN, moduli, domain)), (r, k)


def test_split_canonical_moduli(env):
    """q >= 2^30 (lazy arithmetic impossible): the canonical split kernels."""
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, 3, 1 << 31)]
    assert all(q > (1 << 30) for q, _ in moduli)
    for domain in (1, 0):
        ctx, plan = _plan(ntt_amd, moduli, domain)
        x = gu.random_operand(LOGN, moduli, 17)
        A = gu.to_dev(x)
        _print("case split canonical", domain)
        plan.forward(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), orc.forward(x, LOGN, moduli, domain))
        plan.inverse(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), x)


def test_split_edge_values(env):
    """All-zero, all q - 1 and alternating 0 / q - 1 limbs."""
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, 3)]
    N = 1 << LOGN
    qm = np.array([q - 1 for q, _ in moduli], np.uint32)[:, None]
    x = np.stack([np.zeros(N, np.uint32), np.full(N, qm[1, 0], np.uint32),
                  np.where(np.arange(N) % 2 == 0, 0, qm[2, 0]).astype(np.uint32)])
    for arith in (1, 0):
        ctx, plan = _plan(ntt_amd, moduli, 1, arith)
        A = gu.to_dev(x)
        _print("case split edges", arith)
        plan.forward(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), orc.forward(x, LOGN, moduli, 1))
        plan.inverse(A)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), x)


def test_split_constants(env):
    """Prologue / epilogue constants on the split kernels."""
    ntt_amd, gu, orc = env
    limbs = 4
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, limbs)]
    qs = np.array([q for q, _ in moduli], np.uint64)[:, None]
    rng = np.random.default_rng(5)
    c = [int(rng.integers(0, q)) for q, _ in moduli]
    e = [int(rng.integers(0, q)) for q, _ in moduli]
    x = gu.random_operand(LOGN, moduli, 71)
    for domain in (1, 0):
        ctx, plan = _plan(ntt_amd, moduli, domain)
        plan.set_constants(c, e)
        A = gu.to_dev(x)
        _print("case split constants", domain)
        plan.forward(A)
        gu.sync()
        xs = (x.astype(np.uint64) * np.array(c, np.uint64)[:, None] % qs).astype(np.uint32)
        fwd = gu.to_host(A).reshape(x.shape)
        assert np.array_equal(fwd, orc.forward(xs, LOGN, moduli, domain))
        plan.inverse(A)
        gu.sync()
        exp = (orc.inverse(fwd, LOGN, moduli, domain).astype(np.uint64) * np.array(e, np.uint64)[:, None]
               % qs).astype(np.uint32)
        assert np.array_equal(gu.to_host(A).reshape(x.shape), exp)


def test_split_graph_replays(env):
    """Forward + inverse split launches captured in a hipGraph: every replay
    (with eager launches of the same plan between replays) is exact."""
    ntt_amd, gu, orc = env
    limbs = 5
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, limbs)]
    x = gu.random_operand(LOGN, moduli, 23)
    y = gu.random_operand(LOGN, moduli, 24)
    ctx, plan = _plan(ntt_amd, moduli, 1)
    src, A, B = gu.to_dev(x), gu.to_dev(x), gu.to_dev(y)
    s = ntt_amd.Stream()
    plan.inverse(B, 1, s)
    s.synchronize()
    exp_inv = orc.inverse(y, LOGN, moduli, 1)
    assert np.array_equal(gu.to_host(B).reshape(y.shape), exp_inv)
    with ntt_amd.device.Graph.capture(s) as g:
        A.copy_from(src, stream=s)
        plan.forward(A, 1, s)
        plan.inverse(A, 1, s)
    exp = orc.forward(x, LOGN, moduli, 1)
    _print("case split graph")
    for r in range(3):
        A.zero()
        gu.sync()
        g.launch(s)
        s.synchronize()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), x), r
        plan.forward(A, 1, s)          # eager work between replays
        s.synchronize()
        assert np.array_equal(gu.to_host(A).reshape(x.shape), exp), r
        B.upload(y)
        plan.inverse(B, 1, s)
        s.synchronize()
        assert np.array_equal(gu.to_host(B).reshape(y.shape), exp_inv), r
    g.close()


def test_split_agrees_with_two_pass_many_operands(env):
    """17 x 48 limbs (the C3 moduli set's shape), 3 operands: the split and
    two-pass engines give the same words, and the roundtrip is the identity."""
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, 48)]
    xs = np.stack([gu.random_operand(LOGN, moduli, 500 + k) for k in range(3)])
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(LOGN, moduli, 1)
    outs = []
    for engine in (ntt_amd.ENGINE_FUSED, ntt_amd.ENGINE_TWO_PASS):
        plan.set_engine(engine)
        A = gu.to_dev(xs)
        _print("case split vs two-pass", engine)
        plan.forward(A, 3)
        gu.sync()
        outs.append(gu.to_host(A).reshape(xs.shape))
        plan.inverse(A, 3)
        gu.sync()
        assert np.array_equal(gu.to_host(A).reshape(xs.shape), xs), engine
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0][1], orc.forward(xs[1], LOGN, moduli, 1))


@pytest.mark.parametrize("arith", [1, 0])
@pytest.mark.parametrize("domain", [1, 0])
def test_split_on_the_fly_twiddles(env, arith, domain):
    """The split kernels with on-the-fly twiddles (tw[i] = tw[i % 64] msb[i / 64]
    on the 2^17-word table, the later rounds' words staged in LDS), both
    directions, several limbs and operands (the staged words are refreshed per
    limb's table)."""
    ntt_amd, gu, orc = env
    limbs, count = 5, 2
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, limbs)]
    ctx, plan = _plan(ntt_amd, moduli, domain, arith)
    plan.set_twiddle_mode(ntt_amd.TWIDDLE_OT)
    for d in (ntt_amd.FORWARD, ntt_amd.INVERSE):
        assert plan.fused_variant_full(count, d)[3] is True
    xs = np.stack([gu.random_operand(LOGN, moduli, 40 + k + 10 * arith + domain) for k in range(count)])
    A = gu.to_dev(xs)
    _print("case split OT", arith, domain)
    plan.forward(A, count)
    gu.sync()
    fwd = gu.to_host(A).reshape(xs.shape)
    for k in range(count):
        assert np.array_equal(fwd[k], orc.forward(xs[k], LOGN, moduli, domain)), k
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(xs.shape), xs)
    ys = np.stack([gu.random_operand(LOGN, moduli, 80 + k + 10 * arith + domain) for k in range(count)])
    B = gu.to_dev(ys)
    plan.inverse(B, count)
    gu.sync()
    got = gu.to_host(B).reshape(ys.shape)
    for k in range(count):
        assert np.array_equal(got[k], orc.inverse(ys[k], LOGN, moduli, domain)), k


def test_tuned_engine_per_direction(env, tmp_path):
    """The autotuner picks each direction's engine (split or two-pass); plans
    follow it per direction, results stay exact, and the choice survives a
    save/load of the tuning file."""
    ntt_amd, gu, orc = env
    limbs, count = 6, 4
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, limbs)]
    ctx = ntt_amd.Context(0)
    _print("case tuned engines 17")
    r = ctx.autotune(LOGN, moduli, count=count, reps=2)
    eng = {"two_pass": ntt_amd.ENGINE_TWO_PASS, "fused": ntt_amd.ENGINE_FUSED}
    plan = ctx.plan(LOGN, moduli, 1)
    assert plan.engine(count, ntt_amd.FORWARD) == eng[r["engine"]]
    assert plan.engine(count, ntt_amd.INVERSE) == eng[r["engine_inv"]]
    xs = np.stack([gu.random_operand(LOGN, moduli, 60 + k) for k in range(count)])
    A = gu.to_dev(xs)
    plan.forward(A, count)
    gu.sync()
    got = gu.to_host(A).reshape(xs.shape)
    for k in range(count):
        assert np.array_equal(got[k], orc.forward(xs[k], LOGN, moduli, 1)), k
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(gu.to_host(A).reshape(xs.shape), xs)
    path = str(tmp_path / "tune17.txt")
    ctx.save_tuning(path)
    ctx2 = ntt_amd.Context(0)
    ctx2.load_tuning(path)
    plan2 = ctx2.plan(LOGN, moduli, 1)
    for d in (ntt_amd.FORWARD, ntt_amd.INVERSE):
        assert plan2.engine(count, d) == plan.engine(count, d)


@pytest.mark.parametrize("direction_first", [0, 1])
def test_split_launch_with_two_pass_tail(env, direction_first):
    """48 limbs x 6 operands (288 limbs on 256 CUs): the whole operands that fit
    one round (5) run on the split kernels and the last operand on the two-pass
    kernels in the same call; every operand matches the oracle, in place and
    out of place."""
    ntt_amd, gu, orc = env
    limbs, count = 48, 6
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(LOGN, limbs)]
    ctx, plan = _plan(ntt_amd, moduli, 1)
    xs = np.stack([gu.random_operand(LOGN, moduli, 700 + 3 * k + direction_first) for k in range(count)])
    ref = orc.forward if direction_first == 0 else orc.inverse
    run = plan.forward if direction_first == 0 else plan.inverse
    _print("case split tail", direction_first)
    A = gu.to_dev(xs)
    run(A, count)
    gu.sync()
    got = gu.to_host(A).reshape(xs.shape)
    for k in range(count):
        assert np.array_equal(got[k], ref(xs[k], LOGN, moduli, 1)), k
    # out of place, the source untouched
    src = gu.to_dev(xs)
    dst = ntt_amd.DeviceBuffer(xs.nbytes)
    (plan.forward_to if direction_first == 0 else plan.inverse_to)(dst, src, count)
    gu.sync()
    assert np.array_equal(gu.to_host(dst).reshape(xs.shape), got)
    assert np.array_equal(gu.to_host(src).reshape(xs.shape), xs)
