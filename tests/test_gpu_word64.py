"""GPU parity of the 64-bit word transforms (ntt_plan64_*): bit-exact against
the oracle's 64-bit restatement (integer arithmetic: tolerance 0) for
62-bit primes, sizes from a single pass to two passes of 2^17, several
operands per launch, and the modulus checks."""
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


def _operand(moduli, logN, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, int(q), 1 << logN, dtype=np.uint64) for q, _ in moduli])


@pytest.mark.parametrize("logN,limbs", [(4, 2), (9, 3), (10, 2), (12, 5), (16, 3), (17, 2)])
def test_word64_matches_oracle(env, logN, limbs):
    ntt_amd, gu, orc = env
    moduli = [tuple(m) for m in ntt_amd.generate_moduli64(logN, limbs)]
    count = 2
    x = np.stack([_operand(moduli, logN, 10 * logN + k) for k in range(count)])
    ctx = ntt_amd.Context(0)
    plan = ctx.plan64(logN, moduli)
    A = ntt_amd.DeviceBuffer.from_numpy(x)
    _print("case word64", logN, limbs)
    plan.forward(A, count)
    gu.sync()
    fwd = A.numpy(np.uint64).reshape(x.shape)
    for k in range(count):
        assert np.array_equal(fwd[k], orc.forward64(x[k], logN, moduli)), k
    plan.inverse(A, count)
    gu.sync()
    assert np.array_equal(A.numpy(np.uint64).reshape(x.shape), x)
    B = ntt_amd.DeviceBuffer.from_numpy(fwd)
    plan.inverse(B, count)
    gu.sync()
    for k in range(count):
        assert np.array_equal(B.numpy(np.uint64).reshape(x.shape)[k], orc.inverse64(fwd[k], logN, moduli)), k
    plan.close()


def test_word64_small_primes_equal_32bit_plain_variant(env):
    """For q < 2^31 the 64-bit path computes the 32-bit plain variant's words."""
    ntt_amd, gu, orc = env
    logN = 13
    m32 = [tuple(m) for m in ntt_amd.generate_moduli(logN, 3)]
    x = gu.random_operand(logN, m32, 4)
    ctx = ntt_amd.Context(0)
    plan = ctx.plan64(logN, m32)      # a primitive root is a quadratic non-residue
    A = ntt_amd.DeviceBuffer.from_numpy(x.astype(np.uint64))
    _print("case word64 small primes")
    plan.forward(A)
    gu.sync()
    assert np.array_equal(A.numpy(np.uint64).reshape(x.shape), orc.forward(x, logN, m32, 0).astype(np.uint64))


def test_word64_modulus_checks(env):
    ntt_amd, gu, orc = env
    ctx = ntt_amd.Context(0)
    q, g = tuple(ntt_amd.generate_moduli64(12, 1)[0])
    for bad in ((q + 2, g), (q, 4), ((1 << 62) + 1, 3), (q - 1, g)):
        with pytest.raises(ntt_amd.NTTError):
            ctx.plan64(12, [bad])
