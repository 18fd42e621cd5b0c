"""Host-side logic of the product library and its C ABI surface (no GPU needed)."""
import ctypes
import os
import random
import re

import pytest

from helpers import ROOT, oracle

import cuda_amd
from cuda_amd import lnt


def header_symbols():
    src = open(os.path.join(ROOT, "include", "lnt_api.h")).read()
    return sorted(set(re.findall(r"\b(lnt_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = cuda_amd.load()
    declared = header_symbols()
    assert set(declared) == set(lnt.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_secured_length_matches_oracle_dense():
    lib = cuda_amd.load()
    L = oracle()
    ps = list(range(33, 40000)) + random.Random(11).sample(range(40000, (1 << 32) - 1), 4000)
    for p in ps:
        assert lib.lnt_secured_log2_length(p) == L.lo_secured_length(p), p


def test_parse_length_strings():
    assert cuda_amd.parse_length("2048:2048") == [2048, 2048]
    assert cuda_amd.parse_length("4K") == [4096]
    assert cuda_amd.parse_length("1M") == [1 << 20]
    assert cuda_amd.parse_length("2048.000000:4096.000000") == [2048, 4096]


def test_create_without_gpu_fails_loudly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(cuda_amd.LntError):
        cuda_amd.LucasPRP(86243)


def test_tuner_choices_round_trip_as_text():
    """lnt_tune_import / lnt_tune_export (host-only): lines in, the same lines out; a
    malformed line is rejected."""
    from cuda_amd.lnt import LntError, tune_export, tune_import
    tune_import("# saved choices\n0 4194304 12 16 -1 -1\n0 65536 12 16 -1 -1\n")  # the default plans
    out = tune_export().splitlines()
    assert "0 4194304 12 16 -1 -1" in out and "0 65536 12 16 -1 -1" in out
    with pytest.raises(LntError):
        tune_import("0 1000 12 16 1 1\n")  # not a power of two
