"""cuda_amd -- MI355X (gfx950) Lucas-Lehmer / PRP engine for Mersenne numbers.

The hot path (one squaring mod 2^p - 1 per iteration, as an irrational-base weighted
NTT over GF(2^64 - 2^32 + 1)) is hand-written HIP in cuda_amd/csrc, exported through the
C ABI in include/lnt_api.h and bound here with ctypes.
"""
from .lnt import Batch, LucasPRP, LntError, load, parse_length, secured_log2_length  # noqa: F401

__all__ = ["Batch", "LucasPRP", "LntError", "load", "parse_length", "secured_log2_length"]
