"""Pins the oracle (oracle/lnt_oracle.c) to what mathematics and GMP fix.

Every oracle function is checked against something independent of its own formula:
Python-int field arithmetic, an exact-integer form of the length bound, closed-form
properties of the IBDWT tables, GMP residues (tests/golden/res64.json), Python-int
brute force on small exponents, and known Mersenne primes.  Planted faults in the
oracle must make these checks fail.
"""
import ctypes
import math
import random

import pytest

from helpers import (P_MOD, OracleTest, digits_of, golden, golden_cases, ll_python, oracle,
                     value_of, widths)

EDGE = [0, 1, 2, P_MOD - 1, P_MOD - 2, (1 << 32) - 1, 1 << 32, (1 << 32) + 1, 1 << 63,
        P_MOD >> 1, (P_MOD >> 1) + 1, 0xFFFFFFFF00000000]


def test_field_ops_match_python_ints():
    L = oracle()
    rnd = random.Random(7)
    pairs = [(a, b) for a in EDGE for b in EDGE] + [(rnd.randrange(P_MOD), rnd.randrange(P_MOD))
                                                     for _ in range(20000)]
    for a, b in pairs:
        assert L.lo_add(a, b) == (a + b) % P_MOD
        assert L.lo_sub(a, b) == (a - b) % P_MOD
        assert L.lo_mul(a, b) == (a * b) % P_MOD
    for a in EDGE[1:] + [rnd.randrange(1, P_MOD) for _ in range(200)]:
        e = rnd.randrange(1 << 64)
        assert L.lo_pow(a, e) == pow(a, e, P_MOD)
        assert L.lo_inv(a) == pow(a, -1, P_MOD)


def secured_exact(p):
    """ntt_lengths.cu:88-98 as exact integers: smallest L with 2^(L+5) >= p, bumped once if
    2 * 2^L * (2^ceil(p/2^L) - 1)^2 >= P."""
    L = 0
    while (32 << L) < p:
        L += 1
    w = -(-p // (1 << L))
    m = (1 << w) - 1
    if m * m * (1 << (L + 1)) >= P_MOD:
        L += 1
    return L


def test_secured_length_dense_sweep():
    L = oracle()
    for p in list(range(33, 60000)) + random.Random(3).sample(range(60000, 1 << 31), 3000):
        assert L.lo_secured_length(p) == secured_exact(p), p
    # the configurations' lengths
    assert L.lo_secured_length(86243) == 12
    assert L.lo_secured_length(70000027) == 22
    assert L.lo_secured_length(140000041) == 23


def test_candidates_format():
    buf = ctypes.create_string_buffer(1024)
    oracle().lo_candidates(70000027, buf, 1024)
    assert buf.value.decode().split(";") == ["4194304.000000", "2048.000000:2048.000000"]
    oracle().lo_candidates(140000041, buf, 1024)
    assert buf.value.decode().split(";") == ["8388608.000000", "2048.000000:4096.000000",
                                             "4096.000000:2048.000000", "256:256:128", "512:256:64"]


@pytest.mark.parametrize("p,n", [(86243, 4096), (9689, 512), (400009, 1 << 15), (1000003, 1 << 16)])
def test_tables_closed_form(p, n):
    L = oracle()
    w = (ctypes.c_uint64 * n)()
    uw = (ctypes.c_uint64 * n)()
    wd = (ctypes.c_uint8 * n)()
    L.lo_tables(p, n, w, uw, wd)
    B = pow(7, (P_MOD - 1) // 192 // n * 5, P_MOD)
    assert pow(B, n, P_MOD) == 2  # B is an n-th root of two
    e = [-(-p * i // n) for i in range(n + 1)]
    assert list(wd) == [e[i + 1] - e[i] for i in range(n)]
    assert sum(wd) == p and set(wd) <= {p // n, -(-p // n)}
    for i in range(n):
        assert w[i] == pow(B, n * e[i] - p * i, P_MOD)  # a_i = 2^(e_i - p i / n)
        assert w[i] * uw[i] % P_MOD * n % P_MOD == 1


@pytest.mark.parametrize("p,n", [(318767093, 1 << 24), (83886071, 1 << 22), (280000001, 1 << 25)])
def test_reference_width_expression_exact_up_to_2p53(p, n):
    """The reference's double expression ceil(1.0*p*(i+1)/n) - ceil(1.0*p*i/n) equals the
    integer form while p * n <= 2^53 (every length up to 2^24 at its largest exponent)."""
    import numpy as np
    from helpers import widths_np
    L = oracle()
    a, b = np.empty(n, np.uint8), np.empty(n, np.uint8)
    L.lo_tables(p, n, None, None, a.ctypes.data_as(ctypes.c_void_p))
    L.lo_exact_widths(p, n, b.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(a, b) and np.array_equal(b, widths_np(p, n).astype(np.uint8))


@pytest.mark.parametrize("p,n,where", [(600000001, 1 << 25, [23116288, 23116289]),
                                       (560000017, 1 << 26, [41518320, 41518321]),
                                       (1200000001, 1 << 26, [17396736, 17396737, 34793473, 34793474,
                                                              52190210, 52190211])])
def test_reference_width_expression_rounds_above_2p53(p, n, where):
    """Above p * n = 2^53 (n >= 2^25) the double products round: the reference's expression
    moves a digit boundary by one bit where p*i mod n is 1, 2 or 3, while the weights (exact
    integers) assume the exact boundary.  The exact widths are the integer form."""
    import numpy as np
    from helpers import widths_np
    L = oracle()
    a, b = np.empty(n, np.uint8), np.empty(n, np.uint8)
    L.lo_tables(p, n, None, None, a.ctypes.data_as(ctypes.c_void_p))
    L.lo_exact_widths(p, n, b.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(b, widths_np(p, n).astype(np.uint8))
    bad = np.nonzero(a != b)[0].tolist()
    assert bad == where
    assert all(min((p * i) % n, (p * (i + 1)) % n) <= 3 for i in bad)


@pytest.mark.parametrize("case", [c for c in golden_cases() if c["p"] < 30000],
                         ids=lambda c: f"{c['mode']}{c['p']}")
def test_oracle_matches_gmp_small(case):
    p, prp = case["p"], case["mode"] == "PRP"
    n = 1 << oracle().lo_secured_length(p)
    for n0 in ((1, n // 32) if n >= 1024 else (1,)):
        o = OracleTest(p, n, prp, n0)
        done = 0
        for k, r in sorted((int(k), int(v, 16)) for k, v in case["res64"].items()):
            o.iterate(k - done)
            done = k
            assert o.res64() == r, (p, case["mode"], n0, k)


def test_known_mersenne_primes_reach_zero_and_nine():
    for c in golden_cases():
        if c["p"] in (521, 607, 1279, 2203, 9689, 44497, 86243) and c["iters"] in (c["p"], c["p"] - 2):
            final = int(c["res64"][str(c["iters"])], 16)
            assert final == (9 if c["mode"] == "PRP" else 0)


def test_oracle_ll_86243_against_gmp_20k():
    g = golden(86243, "LL", 86241)
    o = OracleTest(86243, 4096, False, 64)
    done = 0
    for k in range(1000, 20001, 1000):
        o.iterate(k - done)
        done = k
        assert o.res64() == g[k], k


@pytest.mark.parametrize("p,n", [(107, 4), (521, 32), (1279, 64), (2203, 128), (4423, 256)])
def test_brute_force_random_states(p, n):
    """Random residues (not only the LL orbit) against Python-int squaring."""
    rnd = random.Random(p)
    M = (1 << p) - 1
    for prp in (False, True):
        for trial in range(5):
            v = rnd.randrange(M)
            o = OracleTest(p, n, prp)
            o.set_digits(digits_of(v, p, n))
            o.iterate(3)
            want = ll_python(p, 3, v, prp)
            assert value_of(o.digits(), p, n) % M == want
            assert o.res64() == want & ((1 << 64) - 1)


@pytest.mark.parametrize("p,n", [(521, 32), (4423, 256)])
def test_carry_and_minus_two_edges(p, n):
    """V = 0, 1, 2, M-1 and the all-ones form of 0: wrap-around carries and borrows."""
    M = (1 << p) - 1
    w = widths(p, n)
    for v in (0, 1, 2, 3, M - 1, M - 2):
        for prp in (False, True):
            o = OracleTest(p, n, prp)
            o.set_digits(digits_of(v, p, n))
            o.iterate(1)
            want = (v * v - (0 if prp else 2)) % M
            assert value_of(o.digits(), p, n) == want, (v, prp)
    ones = [(1 << wi) - 1 for wi in w]
    o = OracleTest(p, n, False)
    o.set_digits(ones)
    o.iterate(1)
    assert value_of(o.digits(), p, n) == M - 2  # (2^p - 1)^2 - 2 = -2
    o = OracleTest(p, n, True)
    o.set_digits(ones)
    o.iterate(1)
    assert o.digits() == ones and o.res64() == 0  # r2b keeps the all-ones form of 0


@pytest.mark.parametrize("fault", [1, 2, 3, 4])
def test_planted_faults_are_caught(fault):
    """A twiddle index off by one, a dropped wrap carry, a missing -2 or a shifted weight
    in the oracle each break agreement with GMP."""
    g = golden(9689, "LL", 9687)
    for n0 in (1, 16):
        o = OracleTest(9689, 512, False, n0)
        o.fault(fault)
        o.iterate(2421)
        assert o.res64() != g[2421]
        o2 = OracleTest(9689, 512, False, n0)
        o2.iterate(2421)
        assert o2.res64() == g[2421]
