"""Synthetic workload descriptions for the batched and served paths.

Both are drawn from numpy's PCG64 generator (`numpy.random.default_rng(seed)`),
in the order stated per function, so a seed fixes the workload exactly.
Moduli come from per-logN prime lists: the largest primes q < 2^30 with
q = 1 mod 2N, each with its smallest primitive root (`generate_moduli`); a job
with L limbs takes the first L primes of its size's list, so jobs of one size
share twiddle tables.
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from . import FORWARD, INVERSE, generate_moduli

MAX_LIMBS_PER_SIZE = 64


@functools.lru_cache(maxsize=None)
def prime_list(logN: int, count: int = MAX_LIMBS_PER_SIZE) -> Tuple[Tuple[int, int], ...]:
    """The `count` largest primes q < 2^30 with q = 1 mod 2^(logN+1), with roots."""
    return tuple((int(q), int(g)) for q, g in generate_moduli(logN, count))


def moduli_for(logN: int, limbs: int) -> List[Tuple[int, int]]:
    return list(prime_list(logN, max(limbs, MAX_LIMBS_PER_SIZE))[:limbs])


@dataclass(frozen=True)
class Job:
    logN: int
    limbs: int
    direction: int           # FORWARD / INVERSE

    @property
    def words(self) -> int:
        return self.limbs << self.logN

    @property
    def butterflies(self) -> float:
        return 0.5 * (1 << self.logN) * self.logN * self.limbs


def batched_uneven(seed: int = 0, jobs: int = 256, logn=(12, 17), limbs=(1, 48)) -> List[Job]:
    """Per job, in order: logN ~ U{logn}, L ~ U{limbs}, direction ~ Bernoulli(1/2)
    (1 = inverse).  Drawn as three vectors of length `jobs`."""
    rng = np.random.default_rng(seed)
    ln = rng.integers(logn[0], logn[1] + 1, jobs)
    lb = rng.integers(limbs[0], limbs[1] + 1, jobs)
    dr = rng.integers(0, 2, jobs)
    return [Job(int(a), int(b), INVERSE if c else FORWARD) for a, b, c in zip(ln, lb, dr)]


@dataclass(frozen=True)
class Request:
    arrival_s: float
    job: Job
    seed: int


def serving_stream(seed: int, requests: int, rate_per_s: float, logn=(13, 17), limbs=(4, 64)) -> List[Request]:
    """Poisson arrivals at `rate_per_s`.  Per request, drawn as vectors in this
    order: inter-arrival ~ Exp(rate); logN ~ U{logn} (uniform in log N);
    L = round(exp(U[log lo, log hi])) (log-uniform, rounded to nearest);
    direction ~ Bernoulli(1/2); operand seed = a 63-bit draw."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, requests)
    ln = rng.integers(logn[0], logn[1] + 1, requests)
    lb = np.rint(np.exp(rng.uniform(math.log(limbs[0]), math.log(limbs[1]), requests))).astype(int)
    lb = np.clip(lb, limbs[0], limbs[1])
    dr = rng.integers(0, 2, requests)
    sd = rng.integers(0, 1 << 63, requests)
    t = np.cumsum(gaps)
    return [Request(float(t[i]), Job(int(ln[i]), int(lb[i]), INVERSE if dr[i] else FORWARD), int(sd[i]))
            for i in range(requests)]
