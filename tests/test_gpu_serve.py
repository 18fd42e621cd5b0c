"""GPU: the serving loop (continuous batching over lnt_server_*) returns for every request
exactly the residue the same test gives alone on the plain path, with arrivals while
others run, admission limited by the device-memory budget, and Gerbicz-checked PRP.
"""
import sys
import time

import pytest

import cuda_amd
from cuda_amd.lnt import Server

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def alone(p, prp, iters, block=0, check=0):
    t = cuda_amd.LucasPRP(p, prp=prp)
    if prp and block:
        t.gerbicz_begin(block, check)
        t.gerbicz_iterate(iters)
    else:
        t.iterate(iters)
    v = t.res64()
    zero = (not prp) and not any(t.get_digits())
    t.close()
    return v, zero


# (p, prp, iterations): full LL of two Mersenne primes (residue 0), partial LL and PRP
REQS = [(9689, False, 9687), (86243, False, 86241), (9697, False, 9695), (200003, False, 3000),
        (400009, True, 2000), (1000003, False, 1500), (2000003, True, 2000), (86249, True, 5000),
        (131071, False, 2500), (4000037, False, 700), (521, False, 519), (1279, True, 1279)]


def collect(srv, n, timeout_s=120):
    got, t0 = {}, time.time()
    while len(got) < n:
        assert time.time() - t0 < timeout_s, f"served {len(got)} of {n}"
        for r in srv.poll(wait_ms=100):
            got[r["tag"]] = r
    return got


def test_served_equals_alone_with_staggered_arrivals():
    srv = Server(hbm_budget=8 << 30, step_ms=2.0, gerbicz_block=500, gerbicz_check=2)
    for i, (p, prp, k) in enumerate(REQS):
        say("submit", i, p, prp, k)
        srv.submit(p, prp, k, i)
        time.sleep(0.02 * (i % 3))  # some arrive while others run
    got = collect(srv, len(REQS))
    srv.close()
    for i, (p, prp, k) in enumerate(REQS):
        r = got[i]
        assert r["status"] == 0, (p, prp)
        v, zero = alone(p, prp, k, 500, 2)
        assert r["res64"] == v, (p, prp, k)
        assert bool(r["is_zero"]) == zero, (p, prp)
        if prp and k >= 1000:
            assert r["checks"] >= 1, (p, r)
    assert got[0]["is_zero"] and got[1]["is_zero"]  # M9689 and M86243 are prime


def test_budget_limits_admission():
    # room for two 2^16-word tests at a time (5 words per digit of 8 bytes + 256 KiB each)
    srv = Server(hbm_budget=2 * (8 * 65536 * 5 + (256 << 10)), step_ms=1.0, gerbicz_block=0)
    ps = [1000003, 1000033, 1000037, 1000039, 1000081]
    for i, p in enumerate(ps):
        srv.submit(p, False, 400, i)
    peak = 0
    got, t0 = {}, time.time()
    while len(got) < len(ps):
        assert time.time() - t0 < 60
        peak = max(peak, srv.stats()["running"])
        for r in srv.poll(wait_ms=5):
            got[r["tag"]] = r
    srv.close()
    assert peak <= 2
    for i, p in enumerate(ps):
        assert got[i]["res64"] == alone(p, False, 400)[0]


def test_server_tunes_each_length_once():
    """With tune_loops the first test of each NTT length is tuned before it joins; every
    result still equals the test alone (res64 pinned by the plain path)."""
    reqs = [(200003, False, 1500), (200017, False, 1200), (1000003, False, 800), (1000033, True, 900),
            (86243, False, 3000)]
    srv = Server(hbm_budget=8 << 30, step_ms=2.0, gerbicz_block=0, tune_loops=5)
    for i, (p, prp, k) in enumerate(reqs):
        srv.submit(p, prp, k, i)
    got = collect(srv, len(reqs))
    st = srv.stats()
    srv.close()
    lengths = {cuda_amd.secured_log2_length(p) for p, _, _ in reqs}
    assert st["lengths_tuned"] == len(lengths)
    for i, (p, prp, k) in enumerate(reqs):
        assert got[i]["status"] == 0, (p, got[i])
        assert got[i]["res64"] == alone(p, prp, k)[0], (p, prp, k)


def test_served_mix_sample_matches_gmp():
    """The served mix (cuda_amd.workloads.served_requests, seed 0): its GMP-pinned sample of
    full LL tests (p in [50k, 500k], two of them at length 2^11) and PRP chunks (p in
    [5M, 40M], their first 10,000 iterations) arrives with its Poisson spacing (gaps
    scaled by 1/4, at most 0.25 s) into one server with Gerbicz checks on; every res64,
    every LL verdict and every PRP chunk's Gerbicz checksum equals GMP's."""
    from helpers import golden_cases
    from cuda_amd.workloads import served_sample
    fixtures = {(c["p"], c["mode"], c["iters"]): c for c in golden_cases() if c.get("served")}
    reqs = served_sample(0)
    assert all((r.p, "PRP" if r.prp else "LL", r.iterations) in fixtures for r in reqs)
    assert any(cuda_amd.secured_log2_length(r.p) == 11 for r in reqs if not r.prp)
    srv = Server(hbm_budget=16 << 30, step_ms=5.0, gerbicz_block=1000, gerbicz_check=2)
    prev = reqs[0].arrival
    for r in reqs:
        time.sleep(min(0.25, (r.arrival - prev) / 4))
        prev = r.arrival
        say("submit", r.index, r.p, "PRP" if r.prp else "LL", r.iterations)
        srv.submit(r.p, r.prp, r.iterations, r.index)
    got = collect(srv, len(reqs), timeout_s=240)
    srv.close()
    for r in reqs:
        res = got[r.index]
        c = fixtures[(r.p, "PRP" if r.prp else "LL", r.iterations)]
        assert res["status"] == 0, (r.p, res)
        assert res["res64"] == int(c["res64"][str(r.iterations)], 16), (r.index, r.p)
        if r.prp:
            assert res["checks"] == r.iterations // 2000 and res["rollbacks"] == 0, (r.p, res)
            assert res["check_res64"] == int(c["dres64"][str(r.iterations)], 16), (r.index, r.p)
        else:
            assert not res["is_zero"]  # none of the sampled exponents gives a Mersenne prime
        say("served", r.index, r.p, f"{res['done_s'] - res['submit_s']:.2f} s")
