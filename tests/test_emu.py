"""The HIP kernel bodies run on the CPU (tests/emu: OpenMP threads as workgroup lanes,
bounds-checked LDS and global indexing) against the oracle, bit for bit."""
import ctypes
import os
import random

import pytest

from helpers import OracleTest, emu, golden

os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")


class Emu:
    def __init__(self, p, n, prp=False, force_lc=0, K=0, row3=-1, col3=-1):
        self.L = emu()
        self.n = n
        self.h = self.L.emu_create3(p, n, int(prp), force_lc, K, row3, col3)
        assert self.h, (p, n)

    def __del__(self):
        try:
            self.L.emu_destroy(self.h)
        except Exception:
            pass

    def iterate(self, k):
        assert self.L.emu_iterate(self.h, k) == 0

    def digits(self):
        b = (ctypes.c_uint64 * self.n)()
        self.L.emu_get_digits(self.h, b)
        return list(b)

    def set_digits(self, d):
        self.L.emu_set_digits(self.h, (ctypes.c_uint64 * self.n)(*d))

    def res64(self):
        return int(self.L.emu_res64(self.h))


CASES = [  # (p, n, force_lc): one workgroup at 8..1024 and 4096, large path at 2048 and
           # 2^13..2^17 with several splits
    (131, 8, 0), (521, 32, 0), (1279, 64, 0), (2203, 128, 0), (4423, 256, 0), (9689, 512, 0),
    (19937, 1024, 0), (44497, 2048, 0), (86243, 4096, 0),
    (44497, 2048, 7), (86243, 4096, 7), (86243, 4096, 8),  # the large path at n = 2048 / 4096
    (86243, 8192, 0), (200003, 16384, 0), (400009, 32768, 0), (400009, 32768, 9),
    (1000003, 65536, 0), (1000003, 65536, 10), (2000003, 131072, 0),
    (1000003, 65536, 12), (2000003, 131072, 12),  # 4096-word rows over 16 / 32-row columns
    (2000003, 131072, 13),  # 8192-word rows (mixed radix; 16384-word rows: test below)
]


def test_emulated_8192_word_rows_16_words_per_thread():
    """8192-word rows as 16 x 32 x 16 (512 threads, half of them idle in phase 2)."""
    p, n = 2000003, 131072
    rnd = random.Random(p + 1)
    o = OracleTest(p, n, True)
    d = [rnd.getrandbits(wi) for wi in o.widths()]
    o.set_digits(d)
    e = Emu(p, n, True, 13, 0, 2)
    e.set_digits(d)
    o.iterate(3)
    e.iterate(3)
    assert e.digits() == o.digits()


@pytest.mark.parametrize("p,n,lc", CASES)
@pytest.mark.parametrize("prp", [False, True])
def test_emulated_kernels_match_oracle_random_state(p, n, lc, prp):
    rnd = random.Random(p * 7 + n + prp)
    o = OracleTest(p, n, prp)
    w = o.widths()
    d = [rnd.getrandbits(wi) for wi in w]
    o.set_digits(d)
    e = Emu(p, n, prp, lc)
    e.set_digits(d)
    k = 3 if n >= 32768 else 6
    o.iterate(k)
    e.iterate(k)
    assert e.digits() == o.digits()


@pytest.mark.parametrize("row3", [1, 2])
def test_emulated_16384_word_rows_match_oracle(row3):
    """Mixed-radix 3-phase rows of 16384 words over 16-row columns: 32 x 32 x 16 (row3=1)
    and 16 x 32 x 32 with half the threads idle in phases 2 and 3 (row3=2)."""
    p, n = 4000037, 262144
    rnd = random.Random(p)
    o = OracleTest(p, n, False)
    d = [rnd.getrandbits(wi) for wi in o.widths()]
    o.set_digits(d)
    e = Emu(p, n, False, 14, 0, row3)
    e.set_digits(d)
    o.iterate(3)
    e.iterate(3)
    assert e.digits() == o.digits()


COL3_CASES = [  # (p, n, lc, K): 3-phase columns, R = 512, 1024, 2048, 4096
    (1000003, 65536, 7, 8), (1000003, 65536, 7, 16), (2000003, 131072, 7, 8), (2000003, 131072, 7, 16),
    (4000037, 262144, 7, 4), (4000037, 262144, 7, 8), (8000009, 524288, 7, 4),
    # 8 bits per digit: carried digits of arbitrary inputs may exceed 2^31, so the weight
    # multiply takes the field-encoded digit (no biased form)
    (524309, 65536, 7, 16), (1048583, 131072, 7, 8),
]


@pytest.mark.parametrize("p,n,lc,K", COL3_CASES)
def test_emulated_3phase_columns_match_oracle(p, n, lc, K):
    rnd = random.Random(p + K)
    prp = K == 16
    o = OracleTest(p, n, prp)
    d = [rnd.getrandbits(wi) for wi in o.widths()]
    o.set_digits(d)
    e = Emu(p, n, prp, lc, K, -1, 1)
    e.set_digits(d)
    o.iterate(2)
    e.iterate(2)
    assert e.digits() == o.digits()


@pytest.mark.parametrize("p,n", [(9689, 512), (86243, 4096)])
def test_emulated_orbit_matches_gmp(p, n):
    g = golden(p, "LL")
    e = Emu(p, n)
    done = 0
    for k in sorted(g):
        if k > (3000 if n <= 512 else 1000):
            break
        e.iterate(k - done)
        done = k
        assert e.res64() == g[k], k


def test_emulated_full_small_mersenne_prime():
    e = Emu(4423, 256)
    e.iterate(4421)
    assert e.res64() == 0 and set(e.digits()) == {0}
    e = Emu(4423, 256, prp=True)
    e.iterate(4423)
    d = e.digits()
    assert d[0] == 9 and set(d[1:]) == {0}


def test_emulated_large_path_orbit_matches_oracle():
    """The row/column path along the LL orbit from s0 = 4 (small values early, full-size
    values after ~17 squarings), digit for digit against the oracle's 2-factor chain."""
    o = OracleTest(86243, 8192, False, n0=64)
    e = Emu(86243, 8192)
    for k in (5, 20, 40):
        o.iterate(k)
        e.iterate(k)
        assert e.digits() == o.digits(), k


@pytest.mark.parametrize("p,n,prp,lc", [(521, 32, True, 0), (9689, 512, False, 0), (86243, 4096, True, 0),
                                         (86243, 8192, True, 0), (400009, 32768, False, 0),
                                         (1000003, 131072, True, 13), (1000003, 131072, False, -13)])
def test_emulated_product_matches_python(p, n, prp, lc):
    """The forward / multiply kernel modes (used by the Gerbicz check): x <- x * y mod M_p
    from pre-carry states with a pending LL -2, then one more squaring."""
    from helpers import digits_of, value_of
    L = emu()
    L.emu_mul.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    rnd = random.Random(n + p)
    M = (1 << p) - 1
    a, b = rnd.randrange(M), rnd.randrange(M)
    # lc < 0: split -lc with 16-word mixed-radix rows
    ea, eb = (Emu(p, n, prp, -lc, 0, 2), Emu(p, n, prp, -lc, 0, 2)) if lc < 0 else (Emu(p, n, prp, lc), Emu(p, n, prp, lc))
    ea.set_digits(digits_of(a, p, n))
    eb.set_digits(digits_of(b, p, n))
    ea.iterate(1)
    eb.iterate(2)
    c = 0 if prp else 2
    a = (a * a - c) % M
    for _ in range(2):
        b = (b * b - c) % M
    assert L.emu_mul(ea.h, eb.h) == 0
    assert value_of(ea.digits(), p, n) == a * b % M
    ea.iterate(1)
    assert value_of(ea.digits(), p, n) == ((a * b % M) ** 2 - c) % M


def test_plans_for_the_longest_lengths():
    """n = 2^25 and 2^26 split into 4096-word columns and 8192 / 16384-word rows (mixed-radix
    3-phase rows 16 x 32 x 16 and 16 x 32 x 32 by default); longer transforms and 2-phase
    rows of those lengths are rejected."""
    L = emu()
    L.emu_plan.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    out = (ctypes.c_uint32 * 8)()
    assert L.emu_plan(280000001, 1 << 25, 0, -1, out) == 0
    assert list(out) == [8192, 4096, 0, 16, 32, 16, 16, 4]
    assert L.emu_plan(280000001, 1 << 25, 0, 1, out) == 0
    assert list(out) == [8192, 4096, 0, 32, 16, 16, 16, 4]
    assert L.emu_plan(560000017, 1 << 26, 0, -1, out) == 0
    assert list(out) == [16384, 4096, 0, 16, 32, 32, 16, 4]
    assert L.emu_plan(560000017, 1 << 26, 0, 1, out) == 0
    assert list(out) == [16384, 4096, 0, 32, 32, 16, 16, 4]
    assert L.emu_plan(560000017, 1 << 26, 0, 2, out) == 0
    assert list(out) == [16384, 4096, 0, 16, 32, 32, 16, 4]
    assert L.emu_plan(318767093, 1 << 24, 0, -1, out) == 0
    assert list(out) == [8192, 2048, 0, 16, 32, 16, 16, 8]  # 2^24: 8192-word rows too
    assert L.emu_plan(140000041, 1 << 24, 12, -1, out) == 0 and out[0] == 4096 and out[2] == 16
    assert L.emu_plan(140000041, 1 << 24, 12, 2, out) == -1  # 16-word rows only above 4096 words
    assert L.emu_plan(560000017, 1 << 26, 0, 0, out) == -1  # no 2-phase rows of 16384 words
    assert L.emu_plan(1100000017, 1 << 27, 0, -1, out) == -1
    assert L.emu_secured(560000017) <= 26


def _levels_py(p, n):
    """The carry bound of lnt_host.hpp (carry_levels), restated with Python integers."""
    L = n.bit_length() - 1
    wmin, wmax = p >> L, (p + n - 1) >> L
    half_p = ((1 << 64) - (1 << 32)) // 2
    for m in range(1, 9):
        D = (1 << (wmax - 1)) + 2
        for _ in range(200):
            V0 = 2 * n * D * D + 2
            if V0 > half_p:
                break
            V = V0
            for _ in range(m):
                V = (1 << (wmax - 1)) + (V >> wmin) + 1
            if V + 1 <= D:
                return m
            D = V + 1
    return 0


def _simulate_worst_case(p, n, m):
    """Brute-force check of the bound on digit magnitudes: the largest pre-carry values the
    bound allows, pushed through m balanced levels with the worst-case carries (all of
    the same sign), must come back within the fixed point."""
    L = n.bit_length() - 1
    wmin, wmax = p >> L, (p + n - 1) >> L
    D = (1 << (wmax - 1)) + 2
    for _ in range(200):
        V = 2 * n * D * D + 2
        vals = [V] * (m + 1)
        for _ in range(m):
            out = [0] * len(vals)
            for j, v in enumerate(vals):
                w = wmin
                h = (v + (1 << (w - 1))) >> w
                lo = v - (h << w)
                out[j] += lo
                if j + 1 < len(out):
                    out[j + 1] += h
            vals = out
        Dn = max(abs(v) for v in vals[m:]) + (1 << (wmax - 1))
        if Dn <= D:
            return D
        D = Dn
    return None


def test_carry_levels_exact_at_the_top_exponent_of_every_length():
    """For every length 2^3 .. 2^26, the largest exponent the length rule assigns to it and
    a sample below: the engine's carry level count (exact integers) equals the Python
    restatement, is at least 1, keeps 2 n D^2 + 2 <= (P-1)/2, and a brute-force push of the
    worst-case values through that many levels stays within the bound."""
    L = emu()
    L.emu_carry_levels.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.emu_carry_levels.restype = ctypes.c_uint32
    L.emu_secured.argtypes = [ctypes.c_uint32]
    L.emu_secured.restype = ctypes.c_uint32
    for log_n in range(3, 27):
        n = 1 << log_n
        # largest p with secured length log_n (bisection on the monotone length rule)
        lo, hi = 33, 2 ** 32 - 1
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if L.emu_secured(mid) <= log_n:
                lo = mid
            else:
                hi = mid - 1
        top = lo
        assert L.emu_secured(top) == log_n
        for p in (top, top - 1, top - 1000, (top * 3) // 4):
            if p < 2 * n or L.emu_secured(p) != log_n:
                continue
            m = L.emu_carry_levels(p, n)
            assert m == _levels_py(p, n), (p, n)
            assert 1 <= m <= 8, (p, n)
            assert _simulate_worst_case(p, n, m) is not None, (p, n, m)


def test_plan_fallback_is_the_default_plan():
    """A context whose options (LNT_PLAN or a tuned entry) do not fit the length falls back
    to the default plan built into the same Plan object: every field must be the default's
    (a stale field once kept a 1024-row column kernel on 32-row columns), and every plan
    must pass the launchers' consistency check."""
    L = emu()
    L.emu_plan_retry.argtypes = [ctypes.c_uint32] * 6 + [ctypes.c_void_p, ctypes.c_void_p]
    L.emu_plan_retry.restype = ctypes.c_int
    L.emu_plan_retry.argtypes[4] = ctypes.c_int
    L.emu_plan_retry.argtypes[5] = ctypes.c_int
    L.emu_secured.argtypes = [ctypes.c_uint32]
    L.emu_secured.restype = ctypes.c_uint32
    out = (ctypes.c_uint32 * 10)()
    ok = ctypes.c_int()
    for log_n in range(13, 27):
        n = 1 << log_n
        p = 2 * n + 1
        while L.emu_secured(p) < log_n or any(p % q == 0 for q in range(2, 2000) if q * q <= p):
            p += 2
        assert L.emu_plan_retry(p, n, 0, 0, -1, -1, out, ctypes.byref(ok)) == 1
        default = tuple(out)
        for lc in range(7, 15):
            for K in (16, 8, 4):
                for r3 in (0, 1, 2):
                    for c3 in (0, 1):
                        used = L.emu_plan_retry(p, n, lc, K, r3, c3, out, ctypes.byref(ok))
                        assert used in (0, 1) and ok.value == 1, (n, lc, K, r3, c3)
                        if used == 0:
                            assert tuple(out) == default, (n, lc, K, r3, c3, tuple(out), default)


def test_biased_weight_multiply_matches_python():
    """gl_mul_biased(d + 2^31, w, -2^31 w mod P) = d w mod P for every 32-bit biased digit,
    edge operands and random ones (the column kernels' weight multiply on carried digits)."""
    import random
    L = emu()
    L.emu_gl_mul_biased.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]
    L.emu_gl_mul_biased.restype = ctypes.c_uint64
    P = (1 << 64) - (1 << 32) + 1
    rng = random.Random(31)
    ws = [0, 1, 2, P - 1, P - 2, (1 << 32) - 1, 1 << 32, (1 << 63), P >> 1] + [rng.randrange(P) for _ in range(200)]
    ds = [0, 1, (1 << 31) - 1, 1 << 31, (1 << 31) + 1, (1 << 32) - 1] + [rng.randrange(1 << 32) for _ in range(200)]
    for w in ws:
        wb = (-(1 << 31) * w) % P
        for db in ds:
            assert L.emu_gl_mul_biased(db, w, wb) == ((db - (1 << 31)) * w) % P, (db, w)


def test_biased_digit_bound():
    """digits_fit_31: after the plan's carry levels every digit of any input below 2^63 stays
    below 2^31, checked by iterating the level bound in Python; true at the 2^22 headline
    exponent, false where few digit bits leave large carries."""
    L = emu()
    L.emu_digits_fit_31.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.emu_digits_fit_31.restype = ctypes.c_int
    L.emu_carry_levels.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.emu_carry_levels.restype = ctypes.c_uint32
    seen = set()
    for p, n in [(70000027, 1 << 22), (140000041, 1 << 23), (2000003, 1 << 17), (1000003, 1 << 16),
                 (637534199, 1 << 25), (524309, 1 << 16), (655373, 1 << 16), (600011, 1 << 16),
                 (1048583, 1 << 17)]:
        m = L.emu_carry_levels(p, n)
        if not m:
            continue
        L2 = n.bit_length() - 1
        wmin, wmax = p >> L2, (p + n - 1) >> L2
        V = 1 << 63
        for _ in range(m):
            V = (1 << (wmax - 1)) + (V >> wmin) + 1
        want = int(V < (1 << 31))
        assert L.emu_digits_fit_31(p, n) == want, (p, n)
        seen.add(want)
    assert L.emu_digits_fit_31(70000027, 1 << 22) == 1
    assert seen == {0, 1}
