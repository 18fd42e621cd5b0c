"""GPU: residue products and the Gerbicz-checked PRP path (the base-3 PRP configuration:
block B, a check every L blocks, one injected single-word corruption that the check must
detect and roll back).  Residues and the checkpoint product D = s_0 s_B s_2B ... are
compared with GMP (tests/golden/res64.json)."""
import random
import sys

import pytest

from helpers import digits_of, golden_cases, value_of

import cuda_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def gz_fixture(p, block):
    for c in golden_cases():
        if c["p"] == p and c.get("block") == block:
            return ({int(k): int(v, 16) for k, v in c["res64"].items()},
                    {int(k): int(v, 16) for k, v in c["dres64"].items()})
    return None, None


@pytest.mark.parametrize("p,n,prp", [(521, 32, True), (86243, 4096, False), (86243, 8192, True),
                                      (1000003, 65536, True), (2000003, 131072, False)])
def test_product_matches_python(p, n, prp):
    say("product", p, n, prp)
    rnd = random.Random(p ^ n)
    M = (1 << p) - 1
    a, b = rnd.randrange(M), rnd.randrange(M)
    x = cuda_amd.LucasPRP(p, n, prp=prp)
    y = cuda_amd.LucasPRP(p, n, prp=prp)
    x.set_digits(digits_of(a, p, n))
    y.set_digits(digits_of(b, p, n))
    x.iterate(1)
    y.iterate(1)
    c = 0 if prp else 2
    a, b = (a * a - c) % M, (b * b - c) % M
    x.mul(y)
    assert value_of(x.get_digits(), p, n) == a * b % M
    x.iterate(1)
    assert value_of(x.get_digits(), p, n) == ((a * b % M) ** 2 - c) % M


def run_checked(p, n, iters, block, per_check, inject_at=None, word=17, mask=1 << 9):
    t = cuda_amd.LucasPRP(p, n, prp=True)
    t.gerbicz_begin(block, per_check)
    if inject_at is not None:
        t.gerbicz_iterate(inject_at)
        say("inject at", inject_at)
        t.inject_error(word, mask)
        t.gerbicz_iterate(iters - inject_at)
    else:
        t.gerbicz_iterate(iters)
    return t


def test_gerbicz_small_path_full_prp_with_corruption():
    """PRP of M86243 (prime: final residue 9) with one corrupted word mid-block."""
    res, dres = gz_fixture(86243, 1000)
    t = run_checked(86243, 4096, 86243, 1000, 5, inject_at=12517)
    st = t.gerbicz_stats()
    say(st)
    assert st["rollbacks"] == 1 and st["checks"] >= 17
    d = t.get_digits()
    assert d[0] == 9 and set(d[1:]) == {0}
    assert st["checkpoint_res64"] == dres[st["verified_iteration"]]


def test_gerbicz_large_path_with_corruption():
    res, dres = gz_fixture(2000003, 200)
    t = run_checked(2000003, 1 << 17, 6000, 200, 3, inject_at=777, word=54321, mask=1 << 3)
    st = t.gerbicz_stats()
    say(st)
    assert st["rollbacks"] == 1 and st["verified_iteration"] == 6000
    assert t.res64() == res[6000]
    assert st["checkpoint_res64"] == dres[6000]
    clean = run_checked(2000003, 1 << 17, 6000, 200, 3)
    assert clean.gerbicz_stats()["rollbacks"] == 0
    assert clean.get_digits() == t.get_digits()


def test_gerbicz_c3_size_against_gmp():
    """p = 140,000,041, n = 2^23, block 1000, one corruption at iteration 517: res64 and the
    checkpoint product at every verified checkpoint GMP has been run to."""
    res = {}
    for c in golden_cases():
        if c["p"] == 140000041 and c["mode"] == "PRP":
            res.update({int(k): int(v, 16) for k, v in c["res64"].items()})
    _, dres = gz_fixture(140000041, 1000)
    last = max(k for k in res if k % 2000 == 0) if res else 0
    if last == 0:
        pytest.skip("no GMP fixture for p=140000041 yet")
    t = cuda_amd.LucasPRP(140000041, prp=True)
    assert t.n == 1 << 23
    t.gerbicz_begin(1000, 2)
    t.gerbicz_iterate(517)
    t.inject_error(4321, 1 << 11)
    done = 517
    for k in range(2000, last + 1, 2000):
        t.gerbicz_iterate(k - done)
        done = k
        st = t.gerbicz_stats()
        say("C3 checked", k, st)
        assert st["verified_iteration"] == k
        assert t.res64() == res[k], k
        if dres and k in dres:
            assert st["checkpoint_res64"] == dres[k], k
    assert t.gerbicz_stats()["rollbacks"] == 1


def test_product_and_gerbicz_at_2p25():
    """n = 2^25 (8192-word rows in the forward and multiply modes): x * x through lnt_mul
    against GMP's square, and a checked PRP run with one corruption that is rolled back to
    the residue of the plain path."""
    import numpy as np
    from helpers import gmp_square_digits, random_digits_np
    p, n = 637534199, 1 << 25
    say("product 2^25", p)
    d = random_digits_np(p, n, 3)
    x, y = cuda_amd.LucasPRP(p, prp=True), cuda_amd.LucasPRP(p, prp=True)
    x.set_digits(d)
    y.set_digits(d)
    x.mul(y)
    got = np.frombuffer(x.get_digits_buffer(), dtype=np.uint64)
    x.close()
    y.close()
    assert np.array_equal(got, gmp_square_digits(d, p, n, minus_two=False))
    say("gerbicz 2^25")
    t = run_checked(p, n, 60, 10, 2, inject_at=17, word=123457, mask=1 << 5)
    st = t.gerbicz_stats()
    say(st)
    assert st["rollbacks"] == 1 and st["verified_iteration"] == 60
    s = cuda_amd.LucasPRP(p, prp=True)
    s.iterate(60)
    assert t.res64() == s.res64()
    assert np.array_equal(np.frombuffer(t.get_digits_buffer(), dtype=np.uint64),
                          np.frombuffer(s.get_digits_buffer(), dtype=np.uint64))
