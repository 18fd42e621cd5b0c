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


def test_members_of_different_row_shapes_are_not_grouped(monkeypatch):
    """At 2^24 and 2^25 the row kernel has 32- and 16-words-per-thread variants (row3=1 and
    row3=2) with different table layouts; members of each shape, in one batch, must each
    equal their solo run."""
    import numpy as np
    from helpers import random_digits_np
    cases = [(318767093, "row3=1"), (318767093, "row3=2"), (637534199, "row3=1"), (637534199, "row3=2")]
    batch = cuda_amd.Batch()
    members, starts = [], []
    for i, (p, plan) in enumerate(cases):
        monkeypatch.setenv("LNT_PLAN", plan)
        t = cuda_amd.LucasPRP(p)
        say("member", p, plan, t.describe())
        d = random_digits_np(p, t.n, 100 + i)
        t.set_digits(d)
        batch.add(t)
        members.append(t)
        starts.append(d)
    monkeypatch.delenv("LNT_PLAN")
    assert len({tuple(t.describe()) for t in members}) == 4
    batch.iterate(3)
    batch.sync()
    for (p, plan), t, d in zip(cases, members, starts):
        s = cuda_amd.LucasPRP(p)  # default plan, run alone
        s.set_digits(d)
        s.iterate(3)
        a = np.frombuffer(t.get_digits_buffer(), dtype=np.uint64)
        b = np.frombuffer(s.get_digits_buffer(), dtype=np.uint64)
        assert np.array_equal(a, b), (p, plan)
        s.close()
    batch.close()
    for t in members:
        t.close()


def test_member_state_changed_between_batch_steps():
    """set_digits, a product and a copy on LL members already in a batch clear their
    pending "-2"; the next batch step must not subtract it (each equals a solo run)."""
    p = 1000003
    batch = cuda_amd.Batch()
    a, b, c = cuda_amd.LucasPRP(p), cuda_amd.LucasPRP(p), cuda_amd.LucasPRP(p)
    sa, sb, sc = cuda_amd.LucasPRP(p), cuda_amd.LucasPRP(p), cuda_amd.LucasPRP(p)
    for t in (a, b, c):
        batch.add(t)
    batch.iterate(5)
    for s in (sa, sb, sc):
        s.iterate(5)
    batch.sync()
    d = sa.get_digits()
    d[3] ^= 1
    a.set_digits(d)
    sa.set_digits(d)
    b.mul(b)
    sb.mul(sb)
    c.copy_from(a)
    sc.copy_from(sa)
    batch.iterate(4)
    for s in (sa, sb, sc):
        s.iterate(4)
    batch.sync()
    for t, s, what in ((a, sa, "set_digits"), (b, sb, "mul"), (c, sc, "copy")):
        assert t.get_digits() == s.get_digits(), what
    batch.close()


def test_membership_changes_reuse_the_graph():
    """Members joining or leaving a kernel-shape group with free slots rewrite its parameter
    slots; the iteration graph is captured again only for a new shape (or a full group).
    Every member still equals its solo run."""
    p1, p2 = 1000003, 1000033  # both 2^16
    batch = cuda_amd.Batch()
    a, b, c = cuda_amd.LucasPRP(p1), cuda_amd.LucasPRP(p2), cuda_amd.LucasPRP(p1, prp=True)
    sa, sb, sc = cuda_amd.LucasPRP(p1), cuda_amd.LucasPRP(p2), cuda_amd.LucasPRP(p1, prp=True)
    for t in (a, b):
        batch.add(t)
    batch.iterate(5)
    assert batch.stats()["graph_captures"] == 1
    batch.remove(a)
    d = cuda_amd.LucasPRP(p2)
    sd = cuda_amd.LucasPRP(p2)
    batch.add(d)  # same shape: a free slot
    batch.iterate(6)
    assert batch.stats()["graph_captures"] == 1
    batch.add(c)  # PRP: another group, a new graph
    batch.iterate(7)
    assert batch.stats()["graph_captures"] == 2
    sa.iterate(5)
    sb.iterate(18)
    sd.iterate(13)
    sc.iterate(7)
    a.sync()
    for t, s, what in ((a, sa, "left"), (b, sb, "stayed"), (d, sd, "joined"), (c, sc, "new shape")):
        assert t.get_digits() == s.get_digits(), what
    batch.close()
