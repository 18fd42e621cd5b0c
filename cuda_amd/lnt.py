"""Python binding of liblnt (include/lnt_api.h) over ctypes.

The engine is the HIP library built in-tree (cuda_amd/liblnt.so).  There is no CPU
fallback: if the library is missing or cannot reach a GPU, every call raises.

`LucasPRP` mirrors the reference test driver (lucas_prp.cuh:15-39): one object per
exponent and NTT length, seeded with 4 (LL) or 3 (PRP) (lucas_prp.cu:127-134), advanced
one squaring per iteration (lucas_prp.cu:227-229), with the residue available as the
reference's natural digit vector (lucas_prp.cu:254-263) and as a 64-bit residue.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence, Tuple

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblnt.so")

LNT_OK = 0
LNT_UNSUPPORTED = 1
LNT_DEVICE_ERROR = 2
LNT_OUT_OF_MEMORY = 3
LNT_INVALID_ARGUMENT = 4

MODE_LL = 0
MODE_PRP = 1

# Every entry point declared in include/lnt_api.h.
EXPORTS = (
    "lnt_create", "lnt_set_digits", "lnt_iterate", "lnt_get_digits", "lnt_res64",
    "lnt_sync", "lnt_tune", "lnt_destroy", "lnt_last_error", "lnt_secured_log2_length",
    "lnt_length", "lnt_describe", "lnt_profile", "lnt_mul", "lnt_copy", "lnt_inject_error",
    "lnt_gerbicz_begin", "lnt_gerbicz_iterate", "lnt_gerbicz_stats", "lnt_batch_create",
    "lnt_batch_add", "lnt_batch_remove", "lnt_batch_size", "lnt_batch_iterate", "lnt_batch_sync",
    "lnt_batch_destroy", "lnt_batch_last_error", "lnt_server_create", "lnt_server_submit",
    "lnt_server_poll", "lnt_server_stats", "lnt_server_now", "lnt_server_destroy",
    "lnt_server_last_error", "lnt_selftest", "lnt_tune_export", "lnt_tune_import",
    "lnt_tune_results", "lnt_iterate_static", "lnt_valu_rates", "lnt_batch_stats",
)


class LntError(RuntimeError):
    """A failed liblnt call; `code` is the lnt_status value."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


class ServerConfig(ctypes.Structure):
    """lnt_server_config"""
    _fields_ = [("device", ctypes.c_int), ("hbm_budget", ctypes.c_uint64), ("max_tests", ctypes.c_uint32),
                ("step_ms", ctypes.c_float), ("gerbicz_block", ctypes.c_uint32),
                ("gerbicz_check", ctypes.c_uint32), ("tune_loops", ctypes.c_uint32)]


class Result(ctypes.Structure):
    """lnt_result"""
    _fields_ = [("tag", ctypes.c_uint64), ("p", ctypes.c_uint32), ("prp", ctypes.c_int32),
                ("iterations", ctypes.c_uint64), ("res64", ctypes.c_uint64), ("is_zero", ctypes.c_int32),
                ("status", ctypes.c_int32), ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("checks", ctypes.c_uint64), ("rollbacks", ctypes.c_uint64),
                ("check_res64", ctypes.c_uint64)]


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load liblnt.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `make -C cuda_amd` or __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp, u32, u64p = ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)
    lib.lnt_create.argtypes = [ctypes.c_int, u32, ctypes.POINTER(u32), ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(vp)]
    lib.lnt_set_digits.argtypes = [vp, u64p, ctypes.c_size_t]
    lib.lnt_iterate.argtypes = [vp, u32]
    lib.lnt_get_digits.argtypes = [vp, u64p, ctypes.c_size_t]
    lib.lnt_res64.argtypes = [vp, u64p]
    lib.lnt_sync.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    lib.lnt_tune.argtypes = [vp, ctypes.c_int]
    lib.lnt_tune_results.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    lib.lnt_destroy.argtypes = [vp]
    lib.lnt_destroy.restype = None
    lib.lnt_last_error.argtypes = [vp]
    lib.lnt_last_error.restype = ctypes.c_char_p
    lib.lnt_secured_log2_length.argtypes = [u32]
    lib.lnt_secured_log2_length.restype = u32
    lib.lnt_length.argtypes = [vp]
    lib.lnt_length.restype = u32
    lib.lnt_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    lib.lnt_profile.argtypes = [vp, u32, ctypes.c_char_p, ctypes.c_size_t,
                                ctypes.POINTER(ctypes.c_float), ctypes.c_int]
    lib.lnt_mul.argtypes = [vp, vp]
    lib.lnt_copy.argtypes = [vp, vp]
    lib.lnt_inject_error.argtypes = [vp, u32, ctypes.c_uint64]
    lib.lnt_gerbicz_begin.argtypes = [vp, u32, u32]
    lib.lnt_gerbicz_iterate.argtypes = [vp, u32]
    lib.lnt_gerbicz_stats.argtypes = [vp, u64p, ctypes.c_int]
    lib.lnt_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.lnt_batch_add.argtypes = [vp, vp]
    lib.lnt_batch_remove.argtypes = [vp, vp]
    lib.lnt_batch_size.argtypes = [vp]
    lib.lnt_batch_stats.argtypes = [vp, u64p, ctypes.c_int]
    lib.lnt_batch_iterate.argtypes = [vp, u32]
    lib.lnt_batch_sync.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    lib.lnt_batch_destroy.argtypes = [vp]
    lib.lnt_batch_destroy.restype = None
    lib.lnt_batch_last_error.argtypes = [vp]
    lib.lnt_batch_last_error.restype = ctypes.c_char_p
    lib.lnt_selftest.argtypes = [ctypes.c_int, u32, ctypes.c_uint64, u64p]
    lib.lnt_valu_rates.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p,
                                   ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.lnt_tune_export.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.lnt_tune_import.argtypes = [ctypes.c_char_p]
    lib.lnt_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    lib.lnt_server_submit.argtypes = [vp, u32, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
    lib.lnt_server_poll.argtypes = [vp, ctypes.POINTER(Result), ctypes.c_int, ctypes.c_int]
    lib.lnt_server_stats.argtypes = [vp, u64p, ctypes.c_int]
    lib.lnt_server_now.argtypes = [vp]
    lib.lnt_server_now.restype = ctypes.c_double
    lib.lnt_server_destroy.argtypes = [vp]
    lib.lnt_server_destroy.restype = None
    lib.lnt_server_last_error.argtypes = [vp]
    lib.lnt_server_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def tune_export() -> str:
    """The tuner's choices so far, one line per device and length."""
    lib = load()
    buf = ctypes.create_string_buffer(1 << 16)
    rc = lib.lnt_tune_export(buf, len(buf))
    if rc != LNT_OK:
        raise LntError(rc, lib.lnt_last_error(None).decode())
    return buf.value.decode()


def tune_import(text: str) -> None:
    """Merge saved tuner choices (the format of tune_export) for later contexts."""
    lib = load()
    rc = lib.lnt_tune_import(text.encode())
    if rc != LNT_OK:
        raise LntError(rc, lib.lnt_last_error(None).decode())


def field_selftest(count: int = 1 << 16, seed: int = 1, device: int = 0) -> int:
    """Device field ops against the host forms; returns the number of mismatches."""
    lib = load()
    bad = ctypes.c_uint64()
    rc = lib.lnt_selftest(device, count, seed, ctypes.byref(bad))
    if rc != LNT_OK:
        raise LntError(rc, lib.lnt_last_error(None).decode())
    return int(bad.value)


def valu_rates(device: int = 0) -> dict:
    """Field operations per CU per measured clock on the device ({name: rate}, plus "ghz")."""
    lib = load()
    cap = 16
    rates = (ctypes.c_double * cap)()
    names = ctypes.create_string_buffer(512)
    ghz = ctypes.c_double()
    k = lib.lnt_valu_rates(device, rates, names, len(names), ctypes.byref(ghz), cap)
    if k < 0:
        raise LntError(-k, "lnt_valu_rates failed")
    out = dict(zip(names.value.decode().split(";"), [float(rates[i]) for i in range(k)]))
    out["ghz"] = float(ghz.value)
    return out


def secured_log2_length(p: int) -> int:
    """log2 of the NTT length the reference picks for exponent p (ntt_lengths.cu:88-98)."""
    return int(load().lnt_secured_log2_length(p))


def parse_length(description: str) -> List[int]:
    """Factors of an NTT length string such as "2048:2048", "4K" or "1M"
    (NttLengths::StringToLength / StringToLengths, ntt_lengths.cu:32-75)."""
    out = []
    for part in description.split(":"):
        part = part.strip()
        factor = 1
        if part[-1:].upper() == "K":
            factor, part = 1024, part[:-1]
        if part[-1:].upper() == "M":
            factor, part = 1024 * 1024, part[:-1]
        out.append(int(float(part)) * factor)
    return out


class LucasPRP:
    """One Lucas-Lehmer (prp=False) or base-3 PRP (prp=True) test of M_p on the GPU."""

    def __init__(self, p: int, length: Optional[str | int | Sequence[int]] = None,
                 prp: bool = False, device: int = 0):
        self._lib = load()
        self.p = int(p)
        self.prp = bool(prp)
        if length is None:
            factors: List[int] = []
        elif isinstance(length, str):
            factors = parse_length(length)
        elif isinstance(length, int):
            factors = [length]
        else:
            factors = [int(f) for f in length]
        arr = (ctypes.c_uint32 * max(1, len(factors)))(*factors) if factors else None
        h = ctypes.c_void_p()
        rc = self._lib.lnt_create(device, self.p, arr, len(factors),
                                  MODE_PRP if self.prp else MODE_LL, ctypes.byref(h))
        if rc != LNT_OK:
            raise LntError(rc, self._lib.lnt_last_error(None).decode())
        self._h = h
        self.n = int(self._lib.lnt_length(h))
        self.iteration = 0

    # -- helpers -----------------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != LNT_OK:
            raise LntError(rc, self._lib.lnt_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lnt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the hot path ----------------------------------------------------------------
    def iterate(self, k: int = 1) -> None:
        """Enqueue k iterations (asynchronous)."""
        self._check(self._lib.lnt_iterate(self._h, int(k)))
        self.iteration += int(k)

    def sync(self) -> float:
        """Wait; GPU milliseconds since the previous sync."""
        ms = ctypes.c_float()
        self._check(self._lib.lnt_sync(self._h, ctypes.byref(ms)))
        return float(ms.value)

    def get_digits(self) -> List[int]:
        buf = (ctypes.c_uint64 * self.n)()
        self._check(self._lib.lnt_get_digits(self._h, buf, self.n))
        return list(buf)

    def get_digits_buffer(self):
        buf = (ctypes.c_uint64 * self.n)()
        self._check(self._lib.lnt_get_digits(self._h, buf, self.n))
        return buf

    def set_digits(self, digits: Sequence[int]) -> None:
        """Digits as a sequence of ints, or (no copy) a contiguous numpy uint64 array."""
        if len(digits) != self.n:
            raise ValueError(f"expected {self.n} digits, got {len(digits)}")
        if getattr(digits, "dtype", None) is not None and str(digits.dtype) == "uint64" \
                and digits.flags["C_CONTIGUOUS"]:
            buf = digits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        else:
            buf = (ctypes.c_uint64 * self.n)(*digits)
        self._check(self._lib.lnt_set_digits(self._h, buf, self.n))

    def res64(self) -> int:
        out = ctypes.c_uint64()
        self._check(self._lib.lnt_res64(self._h, ctypes.byref(out)))
        return int(out.value)

    def tune(self, loops: int = 50) -> None:
        self._check(self._lib.lnt_tune(self._h, int(loops)))

    def tune_results(self) -> List[Tuple[str, float]]:
        """(candidate, ms for `loops` iterations) of every candidate the last tune() timed."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self._lib.lnt_tune_results(self._h, buf, len(buf)))
        out = []
        for line in buf.value.decode().splitlines():
            name, _, ms = line.rpartition(" = ")
            out.append((name, float(ms)))
        return out

    # -- products, copies, faults, Gerbicz ------------------------------------------
    def mul(self, other: "LucasPRP") -> None:
        """self <- self * other mod M_p (same p, length, mode and split)."""
        self._check(self._lib.lnt_mul(self._h, other._h))

    def copy_from(self, other: "LucasPRP") -> None:
        self._check(self._lib.lnt_copy(self._h, other._h))
        self.iteration = other.iteration

    def inject_error(self, word: int, xor_mask: int) -> None:
        """Flip bits of one residue word on the device (a single-word corruption)."""
        self._check(self._lib.lnt_inject_error(self._h, int(word), int(xor_mask)))

    def gerbicz_begin(self, block: int = 1000, blocks_per_check: int = 10) -> None:
        """Start the Gerbicz-checked PRP from the current residue (PRP tests only)."""
        self._check(self._lib.lnt_gerbicz_begin(self._h, int(block), int(blocks_per_check)))

    def gerbicz_iterate(self, k: int) -> None:
        """k checked iterations; failed checks roll back to the last verified checkpoint and
        redo the iterations since."""
        self._check(self._lib.lnt_gerbicz_iterate(self._h, int(k)))
        self.iteration += int(k)

    def gerbicz_stats(self) -> dict:
        out = (ctypes.c_uint64 * 5)()
        self._check(self._lib.lnt_gerbicz_stats(self._h, out, 5))
        return {"iteration": int(out[0]), "verified_iteration": int(out[1]), "checks": int(out[2]),
                "rollbacks": int(out[3]), "checkpoint_res64": int(out[4])}

    def describe(self) -> List[str]:
        buf = ctypes.create_string_buffer(1024)
        self._check(self._lib.lnt_describe(self._h, buf, len(buf)))
        return buf.value.decode().split(";")

    def profile(self, iters: int) -> List[Tuple[str, float]]:
        """Average GPU ms per launch of each kernel of the iteration, each launched `iters`
        times back to back (the residue is left unchanged)."""
        names = ctypes.create_string_buffer(1024)
        ms = (ctypes.c_float * 8)()
        cnt = self._lib.lnt_profile(self._h, int(iters), names, len(names), ms, 8)
        if cnt < 0:
            raise LntError(-cnt, self._lib.lnt_last_error(self._h).decode())
        if cnt == 0:
            raise LntError(LNT_DEVICE_ERROR, self._lib.lnt_last_error(self._h).decode())
        return list(zip(names.value.decode().split(";"), [float(ms[i]) for i in range(cnt)]))


class Batch:
    """Independent tests advanced together: one hipGraph launch per iteration for all
    members (members of one shape share batched kernel launches)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        rc = self._lib.lnt_batch_create(device, ctypes.byref(h))
        if rc != LNT_OK:
            raise LntError(rc, self._lib.lnt_last_error(None).decode())
        self._h = h
        self.members: List[LucasPRP] = []

    def _check(self, rc: int) -> None:
        if rc != LNT_OK:
            raise LntError(rc, self._lib.lnt_batch_last_error(self._h).decode())

    def add(self, t: LucasPRP) -> None:
        self._check(self._lib.lnt_batch_add(self._h, t._h))
        self.members.append(t)

    def remove(self, t: LucasPRP) -> None:
        self._check(self._lib.lnt_batch_remove(self._h, t._h))
        self.members.remove(t)

    def __len__(self) -> int:
        return int(self._lib.lnt_batch_size(self._h))

    def stats(self) -> dict:
        out = (ctypes.c_uint64 * 3)()
        self._check(self._lib.lnt_batch_stats(self._h, out, 3))
        return {"members": int(out[0]), "groups": int(out[1]), "graph_captures": int(out[2])}

    def iterate(self, k: int) -> None:
        """Every member advances k iterations (asynchronous; a member's own later work is
        ordered after it)."""
        self.members = [t for t in self.members if getattr(t, "_h", None)]  # closed members left
        self._check(self._lib.lnt_batch_iterate(self._h, int(k)))
        for t in self.members:
            t.iteration += int(k)

    def sync(self) -> float:
        ms = ctypes.c_float()
        self._check(self._lib.lnt_batch_sync(self._h, ctypes.byref(ms)))
        return float(ms.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lnt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop (lnt_server_*): submit tests at any time; a scheduler thread in the
    library admits them by device footprint and runs every admitted test in one batch."""

    def __init__(self, hbm_budget: int = 200 << 30, max_tests: int = 0, step_ms: float = 5.0,
                 gerbicz_block: int = 1000, gerbicz_check: int = 0, device: int = 0, tune_loops: int = 0):
        self._lib = load()
        cfg = ServerConfig(device, hbm_budget, max_tests, step_ms, gerbicz_block, gerbicz_check, tune_loops)
        h = ctypes.c_void_p()
        rc = self._lib.lnt_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != LNT_OK:
            raise LntError(rc, self._lib.lnt_last_error(None).decode())
        self._h = h

    def submit(self, p: int, prp: bool, iterations: int, tag: int) -> None:
        rc = self._lib.lnt_server_submit(self._h, p, int(prp), iterations, tag)
        if rc != LNT_OK:
            raise LntError(rc, self._lib.lnt_server_last_error(self._h).decode())

    def poll(self, wait_ms: int = 0, cap: int = 256) -> List[dict]:
        buf = (Result * cap)()
        k = self._lib.lnt_server_poll(self._h, buf, cap, wait_ms)
        if k < 0:
            raise LntError(-k, "lnt_server_poll: bad argument")
        return [{f: getattr(buf[i], f) for f, _ in Result._fields_} for i in range(k)]

    def stats(self) -> dict:
        out = (ctypes.c_uint64 * 9)()
        self._lib.lnt_server_stats(self._h, out, 9)
        keys = ("queued", "running", "finished", "bytes_admitted", "steps", "iterations", "admitted",
                "graph_captures", "lengths_tuned")
        return {k: int(v) for k, v in zip(keys, out)}

    def now(self) -> float:
        return float(self._lib.lnt_server_now(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lnt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
