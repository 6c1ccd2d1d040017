"""Workload generators of the batched and served paths (host only)."""
import numpy as np

from ntt_amd import workloads
from tests import mathref


def test_batched_uneven_ranges_and_determinism():
    a = workloads.batched_uneven(0)
    b = workloads.batched_uneven(0)
    assert a == b and len(a) == 256
    assert {j.logN for j in a} == set(range(12, 18))
    assert min(j.limbs for j in a) >= 1 and max(j.limbs for j in a) <= 48
    assert {j.direction for j in a} == {0, 1}
    # about 1 GB of operands in total (4 bytes per word)
    total = sum(4 * j.words for j in a)
    assert 0.5e9 < total < 2e9
    assert workloads.batched_uneven(1) != a


def test_moduli_for_prefix_and_congruence():
    for logN in (12, 17):
        m48 = workloads.moduli_for(logN, 48)
        assert workloads.moduli_for(logN, 5) == m48[:5]
        qs = [q for q, _ in m48]
        assert qs == sorted(qs, reverse=True) and len(set(qs)) == 48
        for q, g in m48[:4]:
            assert q < 1 << 30 and (q - 1) % (2 << logN) == 0 and mathref.is_prime(q)
            assert g == mathref.smallest_primitive_root(q)


def test_serving_stream_rules():
    reqs = workloads.serving_stream(3, 2000, rate_per_s=1000.0)
    t = np.array([r.arrival_s for r in reqs])
    assert np.all(np.diff(t) >= 0)
    # Poisson: mean gap 1/rate within a few percent at n = 2000
    assert abs(np.mean(np.diff(t)) - 1e-3) < 1e-4
    assert {r.job.logN for r in reqs} == set(range(13, 18))
    lb = np.array([r.job.limbs for r in reqs])
    assert lb.min() >= 4 and lb.max() <= 64
    # log-uniform: the median of L is near sqrt(4 * 64) = 16
    assert 13 <= np.median(lb) <= 19
    assert workloads.serving_stream(3, 2000, 1000.0) == reqs
