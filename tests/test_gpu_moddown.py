"""GPU parity of the out-of-place transforms and the fused mod-down form
(ntt_plan_forward_to / ntt_plan_inverse_to / ntt_plan_forward_mod_down),
bit-exact against the oracle (integer arithmetic: tolerance 0) on both
engines, both domains, lazy-capable and canonical-only moduli, several
operands per call, plus the exact-division identity at full size, the
argument checks, and hipGraph replays of a mod-down and an on-the-fly-twiddle
launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(10, 3), (13, 5), (16, 4), (17, 3)]


@pytest.fixture(scope="module")
def env(oracle_mod):
    import ntt_amd
    from tests import gpu_util
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    return ntt_amd, gpu_util, oracle_mod


def _print(*a):
    print(*a, flush=True)


def _plan(ntt_amd, logN, moduli, domain, engine):
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, domain)
    plan.set_engine(ntt_amd.ENGINE_FUSED if engine == "fused" else ntt_amd.ENGINE_TWO_PASS)
    return ctx, plan


def _engines(logN):
    # logN 17 on the fused engine: the split kernels (the mod-down form there
    # runs on the two-pass kernels)
    return ["fused", "two_pass"] if logN <= 17 else ["two_pass"]


@pytest.mark.parametrize("logN,limbs", CASES)
@pytest.mark.parametrize("domain", [0, 1])
def test_out_of_place_transforms(env, logN, limbs, domain):
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    x = np.stack([gu.random_operand(logN, moduli, 10 + k) for k in range(2)])    # two operands
    for engine in _engines(logN):
        _print("case out-of-place", logN, limbs, domain, engine)
        ctx, plan = _plan(ntt_amd, logN, moduli, domain, engine)
        src = gu.to_dev(x)
        dst = ntt_amd.DeviceBuffer(x.nbytes)
        dst.zero()
        plan.forward_to(dst, src, 2)
        gu.sync()
        fwd = gu.to_host(dst).reshape(x.shape)
        for k in range(2):
            assert np.array_equal(fwd[k], orc.forward(x[k], logN, moduli, domain)), (engine, k)
        assert np.array_equal(gu.to_host(src).reshape(x.shape), x)          # input untouched
        back = ntt_amd.DeviceBuffer(x.nbytes)
        plan.inverse_to(back, dst, 2)
        gu.sync()
        assert np.array_equal(gu.to_host(back).reshape(x.shape), x)
        assert np.array_equal(gu.to_host(dst).reshape(x.shape), fwd)
        # dst == src is the in-place call
        plan.forward_to(src, src, 2)
        gu.sync()
        assert np.array_equal(gu.to_host(src).reshape(x.shape), fwd)


@pytest.mark.parametrize("logN,limbs", CASES)
@pytest.mark.parametrize("domain", [0, 1])
def test_mod_down_matches_oracle(env, logN, limbs, domain):
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    x = np.stack([gu.random_operand(logN, moduli, 40 + k) for k in range(3)])
    sub = np.stack([gu.random_operand(logN, moduli, 60 + k) for k in range(3)])
    rng = np.random.default_rng(logN)
    f = [int(rng.integers(0, 1 << 32)) for _ in moduli]          # reduced mod q by the library
    for engine in _engines(logN):
        _print("case mod-down", logN, limbs, domain, engine)
        ctx, plan = _plan(ntt_amd, logN, moduli, domain, engine)
        plan.set_mod_down_factors(f)
        src, sb = gu.to_dev(x), gu.to_dev(sub)
        dst = ntt_amd.DeviceBuffer(x.nbytes)
        plan.forward_mod_down(dst, src, sb, 3)
        gu.sync()
        got = gu.to_host(dst).reshape(x.shape)
        for k in range(3):
            exp = orc.mod_down(orc.forward(x[k], logN, moduli, domain), sub[k], f, moduli)
            assert np.array_equal(got[k], exp), (engine, k)
        assert np.array_equal(gu.to_host(src).reshape(x.shape), x)
        assert np.array_equal(gu.to_host(sb).reshape(x.shape), sub)
        # in place on the input (dst == src)
        plan.forward_mod_down(src, src, sb, 3)
        gu.sync()
        assert np.array_equal(gu.to_host(src).reshape(x.shape), got)


@pytest.mark.parametrize("engine", ["fused", "two_pass"])
def test_mod_down_canonical_moduli(env, engine):
    """default_moduli() has q >= 2^30: the canonical mod-down kernels."""
    ntt_amd, gu, orc = env
    logN = 14
    moduli = [tuple(m) for m in ntt_amd.default_moduli()][:6]
    x = gu.random_operand(logN, moduli, 5)
    sub = gu.random_operand(logN, moduli, 6)
    f = [3 + 7 * l for l in range(len(moduli))]
    ctx, plan = _plan(ntt_amd, logN, moduli, 1, engine)
    plan.set_mod_down_factors(f)
    dst = ntt_amd.DeviceBuffer(x.nbytes)
    _print("case mod-down canonical", engine)
    plan.forward_mod_down(dst, gu.to_dev(x), gu.to_dev(sub))
    gu.sync()
    exp = orc.mod_down(orc.forward(x, logN, moduli, 1), sub, f, moduli)
    assert np.array_equal(gu.to_host(dst).reshape(x.shape), exp)


def test_mod_down_exact_division_full_size(env):
    """16 limbs of 2^16 words x 8 operands: with P coprime to every q and
    f = P^-1 mod q, mod_down(src = x, sub = NTT(x + P y)) = NTT(y); the
    reference values come from the library's own forward transform, whose
    parity is tested separately."""
    ntt_amd, gu, orc = env
    logN, limbs, count = 16, 16, 8
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    qs = np.array([q for q, _ in moduli], dtype=np.uint64)[:, None]
    P = 1000003 * 999983
    xs = np.stack([gu.random_operand(logN, moduli, 100 + k) for k in range(count)])
    ys = np.stack([gu.random_operand(logN, moduli, 200 + k) for k in range(count)])
    Pm = np.array([P % int(q) for q, _ in moduli], dtype=np.uint64)[:, None]
    xpy = ((xs.astype(np.uint64) + Pm * ys.astype(np.uint64)) % qs).astype(np.uint32)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, 1)
    plan.set_mod_down_factors([pow(P, -1, int(q)) for q, _ in moduli])
    sub, ny = gu.to_dev(xpy), gu.to_dev(ys)
    plan.forward(sub, count)
    plan.forward(ny, count)
    dst = ntt_amd.DeviceBuffer(xs.nbytes)
    _print("case mod-down exact division")
    plan.forward_mod_down(dst, gu.to_dev(xs), sub, count)
    gu.sync()
    assert np.array_equal(gu.to_host(dst), gu.to_host(ny))


def test_inverse_to_with_epilogue(env):
    """INTTForModDown: dst = INTT(src) * c per limb, out of place."""
    ntt_amd, gu, orc = env
    for logN, engine in ((15, "fused"), (17, "two_pass")):
        moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, 3)]
        x = gu.random_operand(logN, moduli, 9)
        c = [11, 1 << 31, 12345]
        ctx, plan = _plan(ntt_amd, logN, moduli, 1, engine)
        plan.set_constants(None, c)
        dst = ntt_amd.DeviceBuffer(x.nbytes)
        _print("case inverse-to epilogue", logN, engine)
        plan.inverse_to(dst, gu.to_dev(x))
        gu.sync()
        qs = np.array([q for q, _ in moduli], dtype=np.uint64)[:, None]
        cv = np.array([v % int(q) for v, (q, _) in zip(c, moduli)], dtype=np.uint64)[:, None]
        exp = (orc.inverse(x, logN, moduli, 1).astype(np.uint64) * cv % qs).astype(np.uint32)
        assert np.array_equal(gu.to_host(dst).reshape(x.shape), exp)


def test_mod_down_argument_errors(env):
    ntt_amd, gu, orc = env
    logN = 12
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, 2)]
    x = gu.random_operand(logN, moduli, 1)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, 1)
    a, b, big = gu.to_dev(x), gu.to_dev(x), ntt_amd.DeviceBuffer(2 * x.nbytes)
    with pytest.raises(ntt_amd.NTTError):
        plan.forward_mod_down(a, b, b)              # factors not set
    plan.set_mod_down_factors([1, 1])
    with pytest.raises(ntt_amd.NTTError):
        plan.forward_mod_down(a, b, a)              # sub == dst
    with pytest.raises(ntt_amd.NTTError):
        plan.forward_to(big.ptr + 256, big.ptr)     # partial overlap of src and dst
    with pytest.raises(ntt_amd.NTTError):
        plan.forward_to(a, b.ptr + 4)               # src not 16-byte aligned
    plan.set_mod_down_factors(None)
    with pytest.raises(ntt_amd.NTTError):
        plan.forward_mod_down(a, b, b)              # factors released


def test_mod_down_and_ot_graph_replays(env):
    """A mod-down launch and an on-the-fly-twiddle forward captured in one
    hipGraph; three replays with eager work and synchronizations between
    them, each checked against the oracle."""
    ntt_amd, gu, orc = env
    logN, limbs = 14, 4
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    x = gu.random_operand(logN, moduli, 71)
    sub = gu.random_operand(logN, moduli, 72)
    f = [5, 7, 11, 13]
    ctx = ntt_amd.Context(0)
    plan = ctx.plan(logN, moduli, 1)
    plan.set_engine(ntt_amd.ENGINE_FUSED)
    plan.set_mod_down_factors(f)
    ot = ctx.plan(logN, moduli, 1)
    ot.set_engine(ntt_amd.ENGINE_FUSED)
    ot.set_twiddle_mode(ntt_amd.TWIDDLE_OT)
    src, sb = gu.to_dev(x), gu.to_dev(sub)
    dst, fw = ntt_amd.DeviceBuffer(x.nbytes), ntt_amd.DeviceBuffer(x.nbytes)
    s = ntt_amd.Stream()
    with ntt_amd.device.Graph.capture(s) as g:
        plan.forward_mod_down(dst, src, sb, 1, s)
        fw.copy_from(src, stream=s)
        ot.forward(fw, 1, s)
    exp_md = orc.mod_down(orc.forward(x, logN, moduli, 1), sub, f, moduli)
    exp_fw = orc.forward(x, logN, moduli, 1)
    for r in range(3):
        dst.zero()
        fw.zero()
        gu.sync()
        _print("case graph mod-down/ot", r)
        g.launch(s)
        s.synchronize()
        assert np.array_equal(gu.to_host(dst).reshape(x.shape), exp_md), r
        assert np.array_equal(gu.to_host(fw).reshape(x.shape), exp_fw), r
        plan.forward(fw, 1, s)          # eager work between replays
        s.synchronize()
    g.close()


@pytest.mark.parametrize("logN,limbs", [(12, 5), (16, 4), (17, 3)])
@pytest.mark.parametrize("domain", [0, 1])
@pytest.mark.parametrize("padding", [False, True])
def test_mod_down_sub_range(env, logN, limbs, domain, padding):
    """Cheddar's src2 limb range and padding (NTT_cheddar.cu:566-590): sub
    holds limbs [lo, hi) only (optionally multiplied by per-limb factors);
    limbs outside the range store (0 - NTT) * f.  Two operands per call."""
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    lo, hi = 1, limbs - 1
    x = np.stack([gu.random_operand(logN, moduli, 80 + k) for k in range(2)])
    sub = np.stack([gu.random_operand(logN, moduli[lo:hi], 90 + k) for k in range(2)])
    f = [11 + 5 * l for l in range(limbs)]
    g = [7919 + 13 * k for k in range(hi - lo)] if padding else None
    for engine in _engines(logN):
        _print("case mod-down range", logN, limbs, domain, padding, engine)
        ctx, plan = _plan(ntt_amd, logN, moduli, domain, engine)
        plan.set_mod_down_factors(f)
        plan.set_mod_down_sub(lo, hi, g)
        dst = ntt_amd.DeviceBuffer(x.nbytes)
        plan.forward_mod_down(dst, gu.to_dev(x), gu.to_dev(sub), 2)
        gu.sync()
        got = gu.to_host(dst).reshape(x.shape)
        for k in range(2):
            exp = orc.mod_down(orc.forward(x[k], logN, moduli, domain), sub[k], f, moduli,
                               sub_start=lo, sub_end=hi, sub_factors=g)
            assert np.array_equal(got[k], exp), (engine, k)
        # back to the whole range: the plain form again
        plan.set_mod_down_sub(0, limbs, None)
        full = np.stack([gu.random_operand(logN, moduli, 95 + k) for k in range(2)])
        plan.forward_mod_down(dst, gu.to_dev(x), gu.to_dev(full), 2)
        gu.sync()
        got = gu.to_host(dst).reshape(x.shape)
        assert np.array_equal(got[1], orc.mod_down(orc.forward(x[1], logN, moduli, domain), full[1], f, moduli))


@pytest.mark.parametrize("logN,limbs", [(11, 6), (16, 5), (17, 4)])
@pytest.mark.parametrize("skip", [(0, 2), (2, 3), (-2, 0)])
def test_forward_skip_range(env, logN, limbs, skip):
    """NTTForModUp's skipped limb range (NTT_cheddar.cu:815-887): limbs outside
    [skip_start, skip_end) transformed in one launch, limbs inside untouched;
    in place and out of place, two operands."""
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli(logN, limbs)]
    s0, s1 = skip if skip[0] >= 0 else (limbs + skip[0], limbs)      # (-2, 0): the last two limbs
    x = np.stack([gu.random_operand(logN, moduli, 110 + k) for k in range(2)])
    fill = np.full(x.shape, 0x5a5a5a5a, dtype=np.uint32)
    for engine in _engines(logN):
        for domain in (0, 1):
            _print("case forward-skip", logN, limbs, skip, engine, domain)
            ctx, plan = _plan(ntt_amd, logN, moduli, domain, engine)
            exp = np.stack([orc.forward(x[k], logN, moduli, domain) for k in range(2)])
            exp[:, s0:s1] = fill[:, s0:s1]
            dst = gu.to_dev(fill)
            plan.forward_skip(dst, gu.to_dev(x), s0, s1, 2)
            gu.sync()
            assert np.array_equal(gu.to_host(dst).reshape(x.shape), exp), (engine, "out of place")
            A = gu.to_dev(x)
            plan.forward_skip(A, None, s0, s1, 2)
            gu.sync()
            exp[:, s0:s1] = x[:, s0:s1]
            assert np.array_equal(gu.to_host(A).reshape(x.shape), exp), (engine, "in place")


@pytest.mark.parametrize("logN,limbs", [(13, 4), (16, 3), (17, 2)])
def test_inverse_from_limb_subset(env, logN, limbs):
    """INTTForModDown over a limb subset of its source (NTT_cheddar.cu:974-1049):
    dst = INTT(limbs [first, first + L) of source operands of W limbs) * e_l,
    two operands, both engines and domains."""
    ntt_amd, gu, orc = env
    W, first = limbs + 3, 2
    wide = [tuple(m) for m in ntt_amd.generate_moduli(logN, W)]
    moduli = wide[first:first + limbs]
    src = np.stack([gu.random_operand(logN, wide, 130 + k) for k in range(2)])
    e = [101 + 3 * l for l in range(limbs)]
    for engine in _engines(logN):
        for domain in (0, 1):
            _print("case inverse-from", logN, limbs, engine, domain)
            ctx, plan = _plan(ntt_amd, logN, moduli, domain, engine)
            plan.set_constants(None, e)
            dst = ntt_amd.DeviceBuffer(4 * 2 * limbs << logN)
            S = gu.to_dev(src)
            plan.inverse_from(dst, S, W, first, 2)
            gu.sync()
            got = gu.to_host(dst).reshape(2, limbs, 1 << logN)
            for k in range(2):
                inv = orc.inverse(src[k, first:first + limbs], logN, moduli, domain).astype(object)
                exp = np.array([inv[l] * e[l] % moduli[l][0] for l in range(limbs)], dtype=np.uint32)
                assert np.array_equal(got[k], exp), (engine, domain, k)
            assert np.array_equal(gu.to_host(S).reshape(src.shape), src)      # source untouched
            with pytest.raises(ntt_amd.NTTError):
                plan.inverse_from(dst, S, W, W - limbs + 1, 2)                  # range past the source
