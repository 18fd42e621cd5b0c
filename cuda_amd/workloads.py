"""Seeded synthetic workloads of the batched and served configurations.

Every random draw comes from splitmix64 (Steele, Lea, Flood 2014; the 64-bit finaliser
of java.util.SplittableRandom), so the exponent lists and arrival times are identical on
every host for a given seed.

* Batched uneven tests: 64 LL tests, prime exponents drawn log-uniform in
  [1,000,000, 20,000,000], 10,000 iterations each, NTT lengths by the reference's length
  rule (ntt_lengths.cu:88-98).
* Served mix: Poisson arrivals (rate 4/s, 2,000 requests); 80% full LL tests with prime
  p uniform in [50,000, 500,000] (p - 2 iterations), 20% 100,000-iteration PRP chunks
  with prime p uniform in [5,000,000, 40,000,000].
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform in [0, 1) with 53 random bits."""
        return (self.next() >> 11) / float(1 << 53)


def is_prime(x: int) -> bool:
    if x < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13):
        if x % q == 0:
            return x == q
    i = 17
    while i * i <= x:
        if x % i == 0:
            return False
        i += 2
    return True


def next_prime(x: int) -> int:
    while not is_prime(x):
        x += 1
    return x


def batched_exponents(seed: int = 0, count: int = 64, lo: int = 1_000_000,
                      hi: int = 20_000_000) -> List[int]:
    """Prime exponents log-uniform in [lo, hi]: p = next_prime(lo * (hi/lo)^u)."""
    rng = SplitMix64(seed)
    return [next_prime(int(lo * math.exp(math.log(hi / lo) * rng.uniform()))) for _ in range(count)]


@dataclass
class Request:
    index: int
    arrival: float      # seconds after the start
    p: int
    prp: bool
    iterations: int


def served_requests(seed: int = 0, count: int = 2000, rate: float = 4.0,
                    ll_range=(50_000, 500_000), prp_range=(5_000_000, 40_000_000),
                    prp_fraction: float = 0.2, prp_iterations: int = 100_000) -> List[Request]:
    """The served mix: exponential inter-arrival times (Poisson process of the given rate)."""
    rng = SplitMix64(seed ^ 0x5EEDC0DE)
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        if rng.uniform() < prp_fraction:
            p = next_prime(prp_range[0] + int(rng.uniform() * (prp_range[1] - prp_range[0])))
            out.append(Request(i, t, p, True, prp_iterations))
        else:
            p = next_prime(ll_range[0] + int(rng.uniform() * (ll_range[1] - ll_range[0])))
            out.append(Request(i, t, p, False, p - 2))
    return out


# Requests of the served mix whose results are pinned by big-integer fixtures: the first
# twelve LL requests in arrival order (they include p = 52,919, whose length is 2^11), one
# more LL request of length 2^11, and three PRP chunks (the first two in arrival order and
# the smallest exponent) cut to their first SAMPLE_PRP_ITERATIONS iterations.
SAMPLE_PRP_ITERATIONS = 10_000


def served_sample(seed: int = 0) -> List[Request]:
    reqs = served_requests(seed)
    ll = [r for r in reqs if not r.prp]
    prp = [r for r in reqs if r.prp]
    pick = ll[:12]
    pick += [r for r in ll[12:] if r.p <= 53_248 and r not in pick][:1]
    chunks = prp[:2] + [min(prp, key=lambda r: r.p)]
    pick += [Request(r.index, r.arrival, r.p, True, SAMPLE_PRP_ITERATIONS) for r in chunks]
    return sorted(pick, key=lambda r: r.index)
