"""GPU: lnt_iterate_static (no device allocation; work buffer and tables in the library's
fixed device array) equals lnt_iterate on the same residues, bit for bit.

Calls in a row on one stream, the later ones smaller (2^16 large path, then 2^13, then
the single-workgroup 4096-word path), one call continuing from the previous call's output
on cached tables, and a size beyond the static maximum rejected with an error.
"""
import ctypes
import random
import sys

import pytest

import cuda_amd

pytestmark = pytest.mark.gpu

P_MOD = (1 << 64) - (1 << 32) + 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def widths(p, n):
    e = [-(-p * i // n) for i in range(n + 1)]
    return [e[i + 1] - e[i] for i in range(n)]


def canonical(words, p, n):
    """Canonical natural digits of the signed digit vector encoded by field elements."""
    w = widths(p, n)
    v = [x if x <= (P_MOD - 1) // 2 else x - P_MOD for x in words]
    d = list(v)
    carry = 0
    for _ in range(8):  # the wrapped carry (2^p = 1) settles in a pass or two
        for i in range(n):
            t = d[i] + carry
            d[i] = t & ((1 << w[i]) - 1)
            carry = t >> w[i]
        if carry == 0:
            break
    assert carry == 0
    if all(di == (1 << wi) - 1 for di, wi in zip(d, w)):
        d = [0] * n  # 2^p - 1 is 0
    return d


@pytest.fixture(scope="module")
def lib():
    L = cuda_amd.load()
    L.lnt_iterate_static.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_uint32, ctypes.c_void_p]
    L.lnt_iterate_static.restype = ctypes.c_int
    L.lnt_last_error.restype = ctypes.c_char_p
    L.lnt_last_error.argtypes = [ctypes.c_void_p]
    return L


class Hip:
    """The HIP runtime liblnt is linked against (one runtime per process), for the
    caller's side: device buffers and a stream."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        self.h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        self.h.hipFree.argtypes = [vp]
        self.h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        self.h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.h.hipStreamSynchronize.argtypes = [vp]
        self.h.hipStreamDestroy.argtypes = [vp]

    def malloc(self, nbytes):
        ptr = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(ptr), nbytes) == 0
        return ptr

    def upload(self, dptr, words):
        import numpy as np
        a = np.ascontiguousarray(np.asarray(words, dtype=np.uint64))
        assert self.h.hipMemcpy(dptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0

    def download(self, dptr, n):
        import numpy as np
        a = np.empty(n, dtype=np.uint64)
        assert self.h.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), dptr, a.nbytes, 2) == 0
        return [int(v) for v in a]

    def stream(self):
        s = ctypes.c_void_p()
        assert self.h.hipStreamCreate(ctypes.byref(s)) == 0
        return s


def test_static_entry_point_equals_lnt_iterate(lib):
    hip = Hip()
    stream = hip.stream()
    cases = [(1000003, 1 << 16, False, 7), (132049, 1 << 13, True, 5), (86243, 4096, False, 300)]
    prev = None
    for p, n, prp, k in cases:
        say("static", p, n, "PRP" if prp else "LL", k)
        rnd = random.Random(p * 7 + n)
        d = [rnd.getrandbits(wi) for wi in widths(p, n)]
        t = cuda_amd.LucasPRP(p, n, prp=prp)
        t.set_digits(d)
        t.iterate(k)
        want = t.get_digits()
        x = hip.malloc(8 * n)
        hip.upload(x, d)
        rc = lib.lnt_iterate_static(p, int(prp), x, n, k, stream)
        assert rc == 0, lib.lnt_last_error(None)
        assert hip.h.hipStreamSynchronize(stream) == 0
        assert canonical(hip.download(x, n), p, n) == want, (p, n, prp, k)
        # the output is the next call's input (same p and n: the cached tables)
        t.iterate(3)
        rc = lib.lnt_iterate_static(p, int(prp), x, n, 3, stream)
        assert rc == 0, lib.lnt_last_error(None)
        assert hip.h.hipStreamSynchronize(stream) == 0
        assert canonical(hip.download(x, n), p, n) == t.get_digits(), (p, n, prp, "continued")
        hip.h.hipFree(x)
        t.close()
        assert prev is None or n < prev
        prev = n
    hip.h.hipStreamDestroy(stream)


def test_static_entry_point_rejects_sizes_beyond_its_maximum(lib):
    hip = Hip()
    n = 2 * lnt_static_max()
    x = hip.malloc(64)
    rc = lib.lnt_iterate_static(70000027, 0, x, n, 1, None)
    assert rc == 4  # LNT_INVALID_ARGUMENT, before any launch
    assert b"static maximum" in lib.lnt_last_error(None)
    hip.h.hipFree(x)


def lnt_static_max():
    """LNT_STATIC_MAX_N as the header defines it, "(1u << k)"."""
    import os
    import re
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "lnt_api.h")
    m = re.search(r"#define LNT_STATIC_MAX_N \(1u << (\d+)\)", open(hdr).read())
    assert m, "LNT_STATIC_MAX_N not found"
    return 1 << int(m.group(1))
