"""GPU parity: the HIP engine (through the C ABI) against GMP fixtures and the oracle.

Bar: bit-exact residues.  Sizes the oracle finishes in seconds are compared digit for
digit; the full-size configurations are compared through GMP res64 checkpoints.
"""
import random
import sys

import pytest

from helpers import OracleTest, digits_of, golden, golden_cases, value_of, widths

import cuda_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# every fixture exponent up to 44,501: lengths 8 (p = 131), 16 (p = 307) ... 2048 (p = 44,497
# and 44,501), the single-workgroup kernels of every length below 4096
SMALL_KAT = [c for c in golden_cases() if c["p"] <= 44501 and "block" not in c]


@pytest.mark.parametrize("case", SMALL_KAT, ids=lambda c: f"{c['mode']}{c['p']}")
def test_small_exponents_match_gmp(case):
    p, prp = case["p"], case["mode"] == "PRP"
    say("case", case["mode"], p)
    t = cuda_amd.LucasPRP(p, prp=prp)
    done = 0
    for k, r in sorted((int(k), int(v, 16)) for k, v in case["res64"].items()):
        t.iterate(k - done)
        done = k
        assert t.res64() == r, (p, case["mode"], k)
    t.close()


def test_m86243_full_ll_every_10000():
    """Plain single LL test of M86243 at n = 4096: res64 at every 10,000 iterations and
    residue 0 at the end."""
    g = golden(86243, "LL")
    t = cuda_amd.LucasPRP(86243)
    assert t.n == 4096
    done = 0
    for k in list(range(10000, 86241, 10000)) + [86241]:
        say("M86243 iteration", k)
        t.iterate(k - done)
        done = k
        assert t.res64() == g[k], k
    d = t.get_digits()
    assert set(d) == {0}


@pytest.mark.parametrize("p,n", [
    (86243, 8192), (200003, 16384), (400009, 32768), (1000003, 65536),
    (2000003, 131072), (4000037, 262144), (8000009, 524288), (16000057, 1 << 20),
    (33000011, 1 << 21), (70000027, 1 << 22),
])
@pytest.mark.parametrize("prp", [False, True])
def test_large_random_state_matches_oracle(p, n, prp):
    say("large random", p, n, prp)
    rnd = random.Random(p + n + prp)
    o = OracleTest(p, n, prp, n0=1 if n <= 65536 else 1024)
    w = o.widths()
    d = [rnd.getrandbits(wi) for wi in w]
    o.set_digits(d)
    t = cuda_amd.LucasPRP(p, n, prp=prp)
    t.set_digits(d)
    assert t.get_digits() == d  # untouched residue reads back as given
    k = 2 if n >= (1 << 20) else 4
    o.iterate(k)
    t.iterate(k)
    assert t.get_digits() == o.digits()
    t.close()


# The tuner's candidate grid (lnt_tune): row length 2^lc, column tile K, row and column
# variants.  Shapes a length does not support fall back to its default plan, so the
# distinct plans are those lnt_describe reports.
TUNER_GRID = [f"lc={lc},K={K},row3={r3},col3={c3}" for lc in range(7, 15) for K in (16, 8, 4)
              for r3 in (0, 1, 2) for c3 in (0, 1)]


@pytest.mark.parametrize("p,n", [(2000003, 1 << 17), (16000057, 1 << 20), (70000027, 1 << 22),
                                 (140000041, 1 << 23)])
def test_every_tunable_plan_matches_oracle(p, n, monkeypatch):
    """Every plan lnt_tune can choose at this length, forced through LNT_PLAN, against the
    oracle digit for digit from a random residue (LL, two iterations)."""
    import numpy as np
    from helpers import random_digits_np
    d = random_digits_np(p, n, seed=n + 5)
    o = OracleTest(p, n, False, n0=1024)
    o.set_digits_np(d)
    o.iterate(2)
    want = o.digits_np()
    seen = set()
    for plan in TUNER_GRID:
        monkeypatch.setenv("LNT_PLAN", plan)
        t = cuda_amd.LucasPRP(p, n)
        desc = tuple(t.describe())
        if desc in seen:
            t.close()
            continue
        seen.add(desc)
        say("plan", n, plan, ";".join(desc))
        t.set_digits(d.tolist())
        t.iterate(2)
        got = np.asarray(t.get_digits_buffer(), dtype=np.uint64)
        assert np.array_equal(got, want), (n, plan, desc)
        t.close()
    say("distinct plans", n, len(seen))
    assert len(seen) >= 3


@pytest.mark.parametrize("p,n,plans", [(44497, 2048, ["", "lc=7"]),
                                       (86243, 4096, ["", "lc=7", "lc=8", "lc=8,K=8", "row3=0"])])
def test_splits_at_2048_and_4096_match_oracle(p, n, plans, monkeypatch):
    """n = 2048 (three launches by default) and 4096 (one workgroup by default, or the
    three-kernel path when a row length is asked for): every plan against the oracle digit
    for digit from a random residue (LL, three iterations)."""
    rnd = random.Random(p + n)
    o = OracleTest(p, n, False)
    d = [rnd.getrandbits(wi) for wi in o.widths()]
    o.set_digits(d)
    o.iterate(3)
    want = o.digits()
    seen = set()
    for plan in plans:
        monkeypatch.setenv("LNT_PLAN", plan)
        t = cuda_amd.LucasPRP(p, n)
        desc = tuple(t.describe())
        say("plan", n, plan or "default", ";".join(desc))
        seen.add(desc)
        t.set_digits(d)
        t.iterate(3)
        assert t.get_digits() == want, (n, plan, desc)
        t.close()
    assert len(seen) >= (1 if n == 2048 else 3)


@pytest.mark.parametrize("p,n", [(318767093, 1 << 24), (600000001, 1 << 25), (1200000001, 1 << 26)])
def test_longest_lengths_match_oracle(p, n):
    """n = 2^24, 2^25 and 2^26 (rows of 8192, 8192 and 16384 words, mixed-radix 3-phase row
    kernel, over 2048- and 4096-word 3-phase columns) at exponents near the top of each
    length: two iterations from a random state, digit for digit against the oracle with
    exact digit widths (the reference's double-precision widths round at 2^25 and 2^26)."""
    import numpy as np
    from helpers import random_digits_np
    say("longest lengths", p, n)
    d = random_digits_np(p, n, p)
    t = cuda_amd.LucasPRP(p)
    assert t.n == n and any(k.startswith("lnt_k_row3m") for k in t.describe())
    t.set_digits(d)
    t.iterate(2)
    got = np.frombuffer(t.get_digits_buffer(), dtype=np.uint64)
    t.close()
    o = OracleTest(p, n, False, n0=1024, exact_widths=True)
    o.set_digits_np(d)
    o.iterate(2)
    assert np.array_equal(got, o.digits_np())


@pytest.mark.parametrize("p,plans", [(637534199, ["row3=1", "row3=2", "lc=14,row3=1", "lc=14,row3=2"]),
                                     (1207959503, ["row3=1", "row3=2"])])
def test_longest_length_row_shapes_agree(p, plans, monkeypatch):
    """Every row shape for 8192 / 16384-word rows (32 or 16 words per thread) and the other
    split of 2^25 give the same digits as the default plan, from one random state."""
    import numpy as np
    from helpers import random_digits_np
    t = cuda_amd.LucasPRP(p)
    n = t.n
    d = random_digits_np(p, n, 11)
    t.set_digits(d)
    t.iterate(3)
    want = np.frombuffer(t.get_digits_buffer(), dtype=np.uint64).copy()
    t.close()
    for pl in plans:
        say("row shape", p, pl)
        monkeypatch.setenv("LNT_PLAN", pl)
        t = cuda_amd.LucasPRP(p)
        t.set_digits(d)
        t.iterate(3)
        assert np.array_equal(np.frombuffer(t.get_digits_buffer(), dtype=np.uint64), want), pl
        t.close()


def test_longest_length_squaring_matches_gmp():
    """One LL squaring at n = 2^25, p = 600,000,001, against GMP's mpz product."""
    import numpy as np
    from helpers import gmp_square_digits, random_digits_np
    p, n = 600000001, 1 << 25
    say("GMP squaring", p, n)
    d = random_digits_np(p, n, 7)
    t = cuda_amd.LucasPRP(p)
    t.set_digits(d)
    t.iterate(1)
    got = np.frombuffer(t.get_digits_buffer(), dtype=np.uint64)
    t.close()
    assert np.array_equal(got, gmp_square_digits(d, p, n))


def test_c2_res64_checkpoints():
    """p = 70,000,027, n = 2^22 (the throughput configuration): res64 at every 1,000
    iterations that GMP has been run to."""
    g = golden(70000027, "LL")
    if not g:
        pytest.skip("no GMP fixture for p=70000027 yet")
    t = cuda_amd.LucasPRP(70000027)
    assert t.n == 1 << 22
    done = 0
    for k in sorted(g):
        say("C2 iteration", k)
        t.iterate(k - done)
        done = k
        assert t.res64() == g[k], k


def test_c3_prp_res64_checkpoints():
    g = golden(140000041, "PRP")
    if not g:
        pytest.skip("no GMP fixture for p=140000041 yet")
    t = cuda_amd.LucasPRP(140000041, prp=True)
    assert t.n == 1 << 23
    done = 0
    for k in sorted(g):
        say("C3 iteration", k)
        t.iterate(k - done)
        done = k
        assert t.res64() == g[k], k


@pytest.mark.parametrize("p,n", [(521, 32), (86243, 4096), (400009, 32768)])
def test_edge_values(p, n):
    """0, 1, 2, M-1 and the all-ones form of zero, one iteration each (wrap-around carry
    and borrow, and the reference's PRP all-ones convention)."""
    M = (1 << p) - 1
    w = widths(p, n)
    for v in (0, 1, 2, 3, M - 1, M - 2):
        for prp in (False, True):
            t = cuda_amd.LucasPRP(p, n, prp=prp)
            t.set_digits(digits_of(v, p, n))
            t.iterate(1)
            assert value_of(t.get_digits(), p, n) == (v * v - (0 if prp else 2)) % M, (v, prp)
            t.close()
    ones = [(1 << wi) - 1 for wi in w]
    t = cuda_amd.LucasPRP(p, n)
    t.set_digits(ones)
    t.iterate(1)
    assert value_of(t.get_digits(), p, n) == M - 2
    t = cuda_amd.LucasPRP(p, n, prp=True)
    t.set_digits(ones)
    t.iterate(2)
    assert t.get_digits() == ones and t.res64() == 0


def test_tune_keeps_residue():
    t = cuda_amd.LucasPRP(2000003, 1 << 17)
    u = cuda_amd.LucasPRP(2000003, 1 << 17)
    t.iterate(50)
    u.iterate(50)
    t.tune(20)
    t.iterate(50)
    u.iterate(50)
    assert t.res64() == u.res64()
    assert t.get_digits() == u.get_digits()


def test_device_field_ops_equal_host_forms():
    """The device forms of the field arithmetic (32-bit carry chains) against the host
    forms, which the oracle tests pin to Python integers mod P."""
    from cuda_amd.lnt import field_selftest
    assert field_selftest(1 << 18, seed=7) == 0


@pytest.mark.parametrize("prp", [False, True])
def test_graph_replays_match_oracle(prp):
    """Long iteration runs replay a captured graph of 16 iterations: a run of 37 from a set
    residue (one eager iteration without the LL "-2", two replays, four eager), then 16
    more (one replay), against the oracle; the graph is rebuilt after the tuner changes
    the context."""
    import numpy as np
    from helpers import random_digits_np
    p, n = 2000003, 1 << 17
    d = random_digits_np(p, n, seed=77 + prp)
    o = OracleTest(p, n, prp, n0=256)
    o.set_digits_np(d)
    t = cuda_amd.LucasPRP(p, n, prp=prp)
    t.set_digits(d.tolist())
    for k in (37, 16):
        o.iterate(k)
        t.iterate(k)
        got = np.asarray(t.get_digits_buffer(), dtype=np.uint64)
        assert np.array_equal(got, o.digits_np()), (prp, k)
    t.tune(20)
    o.iterate(40)
    t.iterate(40)
    got = np.asarray(t.get_digits_buffer(), dtype=np.uint64)
    assert np.array_equal(got, o.digits_np()), (prp, "after tune")
    t.close()
