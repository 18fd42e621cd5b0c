"""Test helpers: the oracle (oracle/_build/liblnt_oracle.so), the CPU emulation of the
kernels (tests/emu), golden fixtures.  Test infrastructure only."""
import ctypes
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liblnt_oracle.so")
EMU_SRC = os.path.join(HERE, "emu", "lnt_emu.cpp")
EMU_SO = os.path.join(HERE, "emu", "_build", "libemu.so")
GOLDEN = os.path.join(HERE, "golden", "res64.json")
P_MOD = (1 << 64) - (1 << 32) + 1

_oracle = None
_emu = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                           capture_output=True)
        L = ctypes.CDLL(ORACLE_SO)
        u32, u64, vp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
        for f in ("lo_add", "lo_sub", "lo_mul", "lo_pow"):
            getattr(L, f).argtypes = [u64, u64]
            getattr(L, f).restype = u64
        L.lo_inv.argtypes = [u64]
        L.lo_inv.restype = u64
        L.lo_secured_length.argtypes = [u32]
        L.lo_secured_length.restype = u32
        L.lo_candidates.argtypes = [u32, ctypes.c_char_p, ctypes.c_int]
        L.lo_tables.argtypes = [u32, u32, vp, vp, vp]
        L.lo_exact_widths.argtypes = [u32, u32, vp]
        L.lo_set_exact_widths.argtypes = [vp]
        L.lo_create.argtypes = [u32, u32, u32, ctypes.c_int]
        L.lo_create.restype = vp
        L.lo_destroy.argtypes = [vp]
        L.lo_iterate.argtypes = [vp, u32]
        L.lo_set_digits.argtypes = [vp, vp]
        L.lo_get_digits.argtypes = [vp, vp]
        L.lo_get_widths.argtypes = [vp, vp]
        L.lo_res64.argtypes = [vp]
        L.lo_res64.restype = u64
        L.lo_set_fault.argtypes = [vp, ctypes.c_int]
        L.lo_n.argtypes = [vp]
        L.lo_n.restype = u32
        L.lo_threads.restype = ctypes.c_int
        _oracle = L
    return _oracle


class OracleTest:
    """The oracle's LucasPRP restatement (n0 = 1: one transform, else the 2-factor chain)."""

    def __init__(self, p, n, prp=False, n0=1, exact_widths=False):
        L = oracle()
        self.L = L
        self.n = n
        self.h = L.lo_create(p, n0, n // n0, int(prp))
        if exact_widths:  # integer digit widths (differ from the reference's only when p*n > 2^53)
            L.lo_set_exact_widths(self.h)

    def __del__(self):
        try:
            self.L.lo_destroy(self.h)
        except Exception:
            pass

    def iterate(self, k):
        self.L.lo_iterate(self.h, k)

    def res64(self):
        return int(self.L.lo_res64(self.h))

    def digits(self):
        b = (ctypes.c_uint64 * self.n)()
        self.L.lo_get_digits(self.h, b)
        return list(b)

    def set_digits(self, d):
        b = (ctypes.c_uint64 * self.n)(*d)
        self.L.lo_set_digits(self.h, b)

    def digits_np(self):
        import numpy as np
        b = np.empty(self.n, dtype=np.uint64)
        self.L.lo_get_digits(self.h, b.ctypes.data_as(ctypes.c_void_p))
        return b

    def set_digits_np(self, d):
        import numpy as np
        d = np.ascontiguousarray(d, dtype=np.uint64)
        self.L.lo_set_digits(self.h, d.ctypes.data_as(ctypes.c_void_p))

    def widths(self):
        b = (ctypes.c_uint8 * self.n)()
        self.L.lo_get_widths(self.h, b)
        return list(b)

    def fault(self, kind):
        self.L.lo_set_fault(self.h, kind)


def emu():
    """Build (if needed) and load the CPU emulation of the HIP kernels."""
    global _emu
    if _emu is None:
        deps = [EMU_SRC] + [os.path.join(ROOT, "cuda_amd", "csrc", f)
                            for f in os.listdir(os.path.join(ROOT, "cuda_amd", "csrc"))
                            if f.endswith(".hpp")]
        if not os.path.exists(EMU_SO) or any(os.path.getmtime(d) > os.path.getmtime(EMU_SO)
                                             for d in deps):
            os.makedirs(os.path.dirname(EMU_SO), exist_ok=True)
            tmp = f"{EMU_SO}.{os.getpid()}.tmp"  # parallel test workers may build at once
            subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o",
                            tmp, EMU_SRC], check=True)
            os.replace(tmp, EMU_SO)
        L = ctypes.CDLL(EMU_SO)
        u32, vp = ctypes.c_uint32, ctypes.c_void_p
        L.emu_create.argtypes = [u32, u32, ctypes.c_int, u32]
        L.emu_create.restype = vp
        L.emu_create3.argtypes = [u32, u32, ctypes.c_int, u32, u32, ctypes.c_int, ctypes.c_int]
        L.emu_create3.restype = vp
        L.emu_destroy.argtypes = [vp]
        L.emu_iterate.argtypes = [vp, u32]
        L.emu_get_digits.argtypes = [vp, vp]
        L.emu_set_digits.argtypes = [vp, vp]
        L.emu_res64.argtypes = [vp]
        L.emu_res64.restype = ctypes.c_uint64
        L.emu_levels.argtypes = [vp]
        L.emu_levels.restype = u32
        _emu = L
    return _emu


def golden_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def golden(p, mode, iters=None):
    """All GMP res64 checkpoints recorded for (p, mode), merged over fixture cases."""
    out = {}
    for c in golden_cases():
        if c["p"] == p and c["mode"] == mode and (iters is None or c["iters"] == iters):
            out.update({int(k): int(v, 16) for k, v in c["res64"].items()})
    return out or None


def widths(p, n):
    e = [-(-p * i // n) for i in range(n + 1)]
    return [e[i + 1] - e[i] for i in range(n)]


def widths_np(p, n):
    import numpy as np
    i = np.arange(n + 1, dtype=np.int64)
    e = -((-p * i) // n)
    return np.diff(e).astype(np.uint64)


def random_digits_np(p, n, seed):
    """Seeded random natural digits (each below 2^width) as a numpy uint64 array."""
    import numpy as np
    w = widths_np(p, n)
    r = np.random.default_rng(seed).integers(0, 1 << 63, n, dtype=np.uint64)
    return r & ((np.uint64(1) << w) - np.uint64(1))


def digits_of(value, p, n):
    """Natural digits of an integer value (0 <= value < 2^p), in linear time."""
    b = bin(value)[2:].zfill(p)[::-1]  # bit i at index i
    e = [-(-p * i // n) for i in range(n + 1)]
    return [int(b[e[i]:e[i + 1]][::-1] or "0", 2) for i in range(n)]


def value_of(d, p, n):
    """Integer value of natural digits (each below 2^width), in linear time."""
    w = widths(p, n)
    parts = [format(di, "b").zfill(wi)[-wi:] if wi else "" for di, wi in zip(d, w)]
    return int("".join(reversed(parts)) or "0", 2)


def ll_python(p, k, s=4, prp=False):
    """s_k by Python integers (small p only)."""
    M = (1 << p) - 1
    for _ in range(k):
        s = s * s if prp else s * s - 2
        s %= M
    return s


# --------------------------------------------------------------- GMP (libgmp.so.10)
# One squaring mod M_p through GMP's documented mpz entry points (no gmp.h in the image;
# mpz_t is the documented {int alloc; int size; limb *d} struct), for checks at lengths
# where Python integers are too slow.  Test infrastructure only.
class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def _bit_offsets(p, n):
    import numpy as np
    w = widths_np(p, n).astype(np.int64)
    return w, np.concatenate([[0], np.cumsum(w)])


def gmp_square_digits(d, p, n, minus_two=True):
    """Canonical digits (numpy uint64, widths ceil(p(i+1)/n) - ceil(p i/n)) of
    x^2 - 2 (or x^2) mod 2^p - 1, x the value of the canonical digits d."""
    import numpy as np
    G = ctypes.CDLL("libgmp.so.10")
    vp = ctypes.c_void_p
    G.__gmpz_import.argtypes = [vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                ctypes.c_size_t, vp]
    G.__gmpz_export.argtypes = [vp, vp, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, vp]
    G.__gmpz_export.restype = vp
    G.__gmpz_sizeinbase.argtypes = [vp, ctypes.c_int]
    G.__gmpz_sizeinbase.restype = ctypes.c_size_t
    G.__gmpz_cmp.argtypes = [vp, vp]
    w, e = _bit_offsets(p, n)
    wmax = int(w.max())
    bits = np.zeros(p, np.uint8)
    for b in range(wmax):
        m = w > b
        bits[e[:-1][m] + b] = ((d[m] >> np.uint64(b)) & np.uint64(1)).astype(np.uint8)
    by = np.packbits(bits, bitorder="little")
    x, s, hi, lo, M = _Mpz(), _Mpz(), _Mpz(), _Mpz(), _Mpz()
    for z in (x, s, hi, lo, M):
        G.__gmpz_init(ctypes.byref(z))
    G.__gmpz_import(ctypes.byref(x), len(by), -1, 1, 0, 0, by.ctypes.data_as(vp))
    G.__gmpz_mul(ctypes.byref(s), ctypes.byref(x), ctypes.byref(x))
    G.__gmpz_setbit(ctypes.byref(M), ctypes.c_ulong(p))
    G.__gmpz_sub_ui(ctypes.byref(M), ctypes.byref(M), ctypes.c_ulong(1))
    if minus_two:
        G.__gmpz_add(ctypes.byref(s), ctypes.byref(s), ctypes.byref(M))  # stays >= 0
        G.__gmpz_sub_ui(ctypes.byref(s), ctypes.byref(s), ctypes.c_ulong(2))
    while G.__gmpz_sizeinbase(ctypes.byref(s), 2) > p:  # fold 2^p = 1
        G.__gmpz_tdiv_q_2exp(ctypes.byref(hi), ctypes.byref(s), ctypes.c_ulong(p))
        G.__gmpz_tdiv_r_2exp(ctypes.byref(lo), ctypes.byref(s), ctypes.c_ulong(p))
        G.__gmpz_add(ctypes.byref(s), ctypes.byref(lo), ctypes.byref(hi))
    if G.__gmpz_cmp(ctypes.byref(s), ctypes.byref(M)) == 0:  # M_p itself is 0
        G.__gmpz_sub(ctypes.byref(s), ctypes.byref(s), ctypes.byref(M))
    buf = np.zeros((p + 7) // 8 + 8, np.uint8)
    cnt = ctypes.c_size_t()
    G.__gmpz_export(buf.ctypes.data_as(vp), ctypes.byref(cnt), -1, 1, 0, 0, ctypes.byref(s))
    for z in (x, s, hi, lo, M):
        G.__gmpz_clear(ctypes.byref(z))
    ob = np.unpackbits(buf, bitorder="little")[:p]
    out = np.zeros(n, np.uint64)
    for b in range(wmax):
        m = w > b
        out[m] |= ob[e[:-1][m] + b].astype(np.uint64) << np.uint64(b)
    return out

