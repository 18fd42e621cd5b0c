"""CPU: the seeded workload generators and the one-at-a-time queue model."""
import math

from cuda_amd import secured_log2_length
from cuda_amd.serve import fifo, percentile, scaled
from cuda_amd.workloads import (Request, SplitMix64, batched_exponents, is_prime, next_prime,
                                served_requests)


def test_splitmix64_known_answers():
    # splitmix64 from state 0: published first outputs of the generator
    r = SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_batched_exponents_shape():
    ps = batched_exponents(0)
    assert len(ps) == 64 and ps == batched_exponents(0)
    assert all(is_prime(p) and 1_000_000 <= p <= 20_000_100 for p in ps)
    logs = sorted({secured_log2_length(p) for p in ps})
    assert logs[0] >= 16 and logs[-1] <= 20
    # log-uniform: about half of the draws below the geometric midpoint
    mid = math.sqrt(1e6 * 2e7)
    assert 20 <= sum(p < mid for p in ps) <= 44
    assert batched_exponents(1) != ps


def test_served_requests_shape():
    rs = served_requests(0)
    assert len(rs) == 2000
    assert all(b.arrival > a.arrival for a, b in zip(rs, rs[1:]))
    assert abs(rs[-1].arrival / 2000 - 0.25) < 0.03  # rate 4/s
    prp = [r for r in rs if r.prp]
    assert 0.16 < len(prp) / len(rs) < 0.24
    assert all(r.iterations == 100_000 and 5_000_000 <= r.p <= 40_000_100 for r in prp)
    ll = [r for r in rs if not r.prp]
    assert all(r.iterations == r.p - 2 and 50_000 <= r.p <= 500_100 and is_prime(r.p) for r in ll)


def test_next_prime():
    assert next_prime(70_000_000) == 70_000_027
    assert next_prime(140_000_000) == 140_000_041


def test_fifo_queue_model():
    rs = [Request(i, a, 0, False, 10) for i, a in enumerate([0.0, 1.0, 1.5, 10.0])]
    assert fifo(rs, [2.0, 2.0, 2.0, 1.0]) == [2.0, 4.0, 6.0, 11.0]
    assert percentile([1, 2, 3, 4], 50) == 2 and percentile([1, 2, 3, 4], 99) == 4
    assert scaled(Request(0, 0, 0, True, 100_000), 0.05) == 5000
