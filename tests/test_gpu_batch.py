"""GPU: batched tests (all members advance in one hipGraph launch per iteration) leave
every member exactly where the same test run alone on the plain path leaves it.

Bar: bit-exact residues (full digit vectors), across several graph replays with the
program's own eager work between them, membership changes, a member re-planned by the
tuner, and against the GMP fixtures of the batched workload where they exist.
"""
import sys

import pytest

from helpers import golden_cases

import cuda_amd
from cuda_amd.workloads import batched_exponents

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# small path (n <= 4096), LL and PRP at one length, several large lengths
MIX = [(9689, False), (86243, True), (86243, False), (200003, False), (400009, True),
       (1000003, False), (2000003, True), (4000037, False)]


def test_batch_matches_solo_across_replays():
    say("batch mix", MIX)
    batch = cuda_amd.Batch()
    members = [cuda_amd.LucasPRP(p, prp=prp) for p, prp in MIX]
    solo = [cuda_amd.LucasPRP(p, prp=prp) for p, prp in MIX]
    for t in members:
        batch.add(t)
    for rnd, k in enumerate((1, 7, 50, 3)):
        batch.iterate(k)
        for s in solo:  # eager work on other streams between the replays
            s.iterate(k)
        batch.sync()
        for t, s, (p, prp) in zip(members, solo, MIX):
            assert t.get_digits() == s.get_digits(), (p, prp, rnd)
        if rnd == 1:  # a member leaves, works alone, and joins again
            batch.remove(members[3])
            members[3].iterate(5)
            solo[3].iterate(5)
            batch.add(members[3])
    batch.close()
    for t in members + solo:
        t.close()


def test_member_used_alone_after_batch_iterate_is_ordered():
    """No sync between the batch launch and the member's own iteration."""
    batch = cuda_amd.Batch()
    t, s = cuda_amd.LucasPRP(1000003), cuda_amd.LucasPRP(1000003)
    batch.add(t)
    batch.iterate(20)
    t.iterate(3)
    s.iterate(23)
    assert t.res64() == s.res64()
    batch.close()


def test_retuned_member_rebuilds_the_graph():
    p = 4000037
    batch = cuda_amd.Batch()
    t, s = cuda_amd.LucasPRP(p), cuda_amd.LucasPRP(p)
    batch.add(t)
    batch.iterate(10)
    t.tune(20)  # may change the member's plan and buffers
    batch.iterate(10)
    s.iterate(20)
    batch.sync()
    assert t.get_digits() == s.get_digits()
    batch.close()


def test_destroyed_member_leaves_the_batch():
    batch = cuda_amd.Batch()
    a, b = cuda_amd.LucasPRP(200003), cuda_amd.LucasPRP(400009)
    batch.add(a)
    batch.add(b)
    b.close()
    assert len(batch) == 1
    batch.iterate(5)
    s = cuda_amd.LucasPRP(200003)
    s.iterate(5)
    assert a.res64() == s.res64()
    batch.close()


def test_batched_workload_sample_matches_gmp():
    """A sample of the batched workload's exponents (seed 0) at 10,000 iterations against
    GMP, run as one batch."""
    fixtures = {c["p"]: int(c["res64"]["10000"], 16) for c in golden_cases()
                if c.get("batched") and c["mode"] == "LL" and c["iters"] == 10000}
    ps = [p for p in sorted(set(batched_exponents(0))) if p in fixtures][:12]
    if not ps:
        pytest.skip("no GMP fixtures of the batched workload yet (make_fixtures.py --batched)")
    say("batched sample", ps)
    batch = cuda_amd.Batch()
    ts = [cuda_amd.LucasPRP(p) for p in ps]
    for t in ts:
        batch.add(t)
    for _ in range(10):
        batch.iterate(1000)
    batch.sync()
    for p, t in zip(ps, ts):
        assert t.res64() == fixtures[p], p
    batch.close()


def test_batch_of_longest_lengths_matches_solo():
    """Two 2^25 tests (8192-word rows, batched row kernel) and one 2^26 test in a batch."""
    import numpy as np
    from helpers import random_digits_np
    ps = [637534199, 637534157, 1207959503]
    say("batch longest", ps)
    batch = cuda_amd.Batch()
    members = [cuda_amd.LucasPRP(p) for p in ps]
    starts = [random_digits_np(p, t.n, i) for i, (p, t) in enumerate(zip(ps, members))]
    for t, d in zip(members, starts):
        t.set_digits(d)
        batch.add(t)
    batch.iterate(4)
    batch.sync()
    for t, p, d in zip(members, ps, starts):
        s = cuda_amd.LucasPRP(p)
        s.set_digits(d)
        s.iterate(4)
        a = np.frombuffer(t.get_digits_buffer(), dtype=np.uint64)
        b = np.frombuffer(s.get_digits_buffer(), dtype=np.uint64)
        assert np.array_equal(a, b), p
        s.close()
    batch.close()
    for t in members:
        t.close()
