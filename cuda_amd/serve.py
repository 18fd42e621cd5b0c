"""Served mixed workload: Poisson arrivals into the serving loop (lnt_server_*), reporting
tests/hour and p50/p99 time-to-result, against the same requests served one at a time
on the plain path (a single FIFO server: each test runs alone from set-up to res64).

  python -m cuda_amd.serve [--requests 2000] [--rate 4] [--seed 0] [--iter-scale S]
                           [--budget-gb 200] [--verify 24] [--json out.json]

The request stream is cuda_amd.workloads.served_requests (80% full LL tests with p in
[50k, 500k], 20% 100,000-iteration PRP chunks with p in [5M, 40M], arrivals at rate 4/s).
--iter-scale multiplies every request's iteration count (1.0 = the stream as specified).

The one-at-a-time baseline is a FIFO queue whose service times come from the plain path
measured on this GPU: per length class, the set-up time of a test and its time per
iteration over whole Gerbicz check periods.  A sample of requests (--verify) is also run
alone for real, in arrival order: each one's res64 must equal its served res64, and its
measured service time checks the modelled one.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from typing import Dict, List

from .lnt import LucasPRP, Server, secured_log2_length
from .workloads import Request, served_requests


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def scaled(r: Request, s: float) -> int:
    return max(1, int(round(r.iterations * s)))


def percentile(xs: List[float], q: float) -> float:
    xs = sorted(xs)
    if not xs:
        return float("nan")
    i = min(len(xs) - 1, max(0, int(math.ceil(q / 100.0 * len(xs))) - 1))
    return xs[i]


def run_served(reqs: List[Request], s: float, budget: int, step_ms: float, block: int,
               check: int, tune_loops: int = 0) -> Dict[int, dict]:
    """Submit each request at its arrival time; collect every result.  With tune_loops the
    server tunes the first test of each length (its time counts in the served figures)."""
    srv = Server(hbm_budget=budget, step_ms=step_ms, gerbicz_block=block, gerbicz_check=check,
                 tune_loops=tune_loops)
    t0 = time.perf_counter()
    base = srv.now()
    out: Dict[int, dict] = {}
    i, last_log, peak = 0, 0.0, 0
    while len(out) < len(reqs):
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i].arrival <= now:
            r = reqs[i]
            srv.submit(r.p, r.prp, scaled(r, s), r.index)
            i += 1
        wait = 50 if i >= len(reqs) else max(1, min(50, int((reqs[i].arrival - now) * 1000)))
        for res in srv.poll(wait_ms=wait):
            out[res["tag"]] = res
        st = srv.stats()
        peak = max(peak, st["running"])
        if now - last_log > 20:
            last_log = now
            log(f"serve: t={now:.0f}s submitted={i} finished={len(out)} queued={st['queued']} "
                f"running={st['running']} steps={st['steps']} GB={st['bytes_admitted'] / 2**30:.1f}")
    st = srv.stats()
    srv.close()
    for res in out.values():  # times relative to the arrival clock
        for k in ("submit_s", "admit_s", "done_s"):
            res[k] -= base
    out[-1] = {"peak_running": peak, "steps": st["steps"], "lengths_tuned": st["lengths_tuned"]}
    return out


def plain_classes(reqs: List[Request], s: float, block: int, check: int,
                  tune_loops: int = 0) -> Dict[tuple, dict]:
    """Plain-path cost per (log2 n, mode): set-up seconds and seconds per iteration (with
    tune_loops, on the plan the tuner picks for the length, tuned once beforehand)."""
    classes: Dict[tuple, int] = {}
    for r in reqs:
        classes.setdefault((secured_log2_length(r.p), r.prp), r.p)
    out = {}
    for (ln, prp), p in sorted(classes.items()):
        if tune_loops:
            t = LucasPRP(p, prp=prp)
            t.tune(tune_loops)
            t.close()
        t0 = time.perf_counter()
        t = LucasPRP(p, prp=prp)
        if prp and block:
            t.gerbicz_begin(block, check)
        t.res64()
        setup = time.perf_counter() - t0
        k = block * check if prp and block else 2000
        t.sync()
        t0 = time.perf_counter()
        (t.gerbicz_iterate if prp and block else t.iterate)(k)
        t.res64()
        per = (time.perf_counter() - t0) / k
        t.close()
        out[(ln, prp)] = {"setup_s": setup, "s_per_iter": per}
        log(f"serve: plain class 2^{ln} {'PRP' if prp else 'LL'}: set-up {setup * 1e3:.1f} ms, "
            f"{per * 1e6:.1f} us/iter")
    return out


def plain_service(r: Request, s: float, classes) -> float:
    c = classes[(secured_log2_length(r.p), r.prp)]
    return c["setup_s"] + scaled(r, s) * c["s_per_iter"]


def fifo(reqs: List[Request], service: List[float]) -> List[float]:
    """Completion times of a single FIFO server."""
    free, done = 0.0, []
    for r, sv in zip(reqs, service):
        start = max(r.arrival, free)
        free = start + sv
        done.append(free)
    return done


def run_alone(r: Request, s: float, block: int, check: int):
    t0 = time.perf_counter()
    t = LucasPRP(r.p, prp=r.prp)
    k = scaled(r, s)
    if r.prp and block:
        t.gerbicz_begin(block, check)
        t.gerbicz_iterate(k)
    else:
        t.iterate(k)
    v = t.res64()
    dt = time.perf_counter() - t0
    t.close()
    return v, dt


def served_mix(requests: int = 2000, rate: float = 4.0, seed: int = 0, iter_scale: float = 1.0,
               budget_gb: float = 200.0, step_ms: float = 5.0, gerbicz_block: int = 1000,
               gerbicz_check: int = 0, verify: int = 24, tune_loops: int = 20) -> dict:
    """Serve the first `requests` of the mix at their arrival times and model the same stream
    one at a time on the plain path; returns the report (tests/hour, p50/p99 both ways)."""
    reqs = served_requests(seed, requests, rate)
    prp_iters = max(scaled(r, iter_scale) for r in reqs if r.prp) if any(r.prp for r in reqs) else 0
    block = gerbicz_block if prp_iters >= gerbicz_block else 0
    check = gerbicz_check or max(1, min(10, prp_iters // max(1, gerbicz_block)))
    log(f"serve: {len(reqs)} requests over {reqs[-1].arrival:.0f} s, iteration scale {iter_scale}, "
        f"Gerbicz block {block} x {check}")

    classes = plain_classes(reqs, iter_scale, block, check, tune_loops)
    service = [plain_service(r, iter_scale, classes) for r in reqs]
    pdone = fifo(reqs, service)
    p_ttr = [d - r.arrival for d, r in zip(pdone, reqs)]

    got = run_served(reqs, iter_scale, int(budget_gb * 2**30), step_ms, block, check, tune_loops)
    meta = got.pop(-1)
    ttr = [got[r.index]["done_s"] - got[r.index]["submit_s"] for r in reqs]
    errors = sum(1 for v in got.values() if v["status"] != 0)
    makespan = max(v["done_s"] for v in got.values()) - reqs[0].arrival
    p_makespan = pdone[-1] - reqs[0].arrival

    # a sample run alone on the plain path: same res64, and a check of the service model
    sample = sorted(set(list(range(min(verify // 2, len(reqs))))
                        + [int(i * len(reqs) / max(1, verify // 2)) for i in range(verify // 2)]))
    equal, model_err = 0, []
    for i in sample:
        r = reqs[i]
        v, dt = run_alone(r, iter_scale, block, check)
        equal += int(v == got[r.index]["res64"])
        model_err.append((service[i] - dt) / dt)
    return {
        "workload": f"{len(reqs)} requests, Poisson rate {rate}/s (seed {seed}): 80% full LL "
                    f"p in [50k, 500k], 20% PRP chunks of 100,000 iterations p in [5M, 40M]; "
                    f"iterations x {iter_scale}; Gerbicz block {block} x {check}; HBM budget "
                    f"{budget_gb:g} GB; each length tuned ({tune_loops} loops per candidate) "
                    f"on first use" if tune_loops else f"{budget_gb:g} GB; default plans",
        "requests": len(reqs), "rate": rate, "iter_scale": iter_scale,
        "arrival_window_s": round(reqs[-1].arrival - reqs[0].arrival, 2),
        "served": {"tests_per_hour": round(len(reqs) / makespan * 3600, 1),
                   "makespan_s": round(makespan, 2),
                   "p50_s": round(percentile(ttr, 50), 3), "p99_s": round(percentile(ttr, 99), 3),
                   "mean_s": round(sum(ttr) / len(ttr), 3), "peak_running": meta["peak_running"],
                   "steps": meta["steps"], "lengths_tuned": meta["lengths_tuned"], "errors": errors,
                   "ll_prime_found": sum(1 for v in got.values() if v["is_zero"] and not v["prp"])},
        "plain_one_at_a_time": {"tests_per_hour": round(len(reqs) / p_makespan * 3600, 1),
                                "makespan_s": round(p_makespan, 2),
                                "p50_s": round(percentile(p_ttr, 50), 3),
                                "p99_s": round(percentile(p_ttr, 99), 3),
                                "busy_s": round(sum(service), 2),
                                "model": "FIFO over measured plain-path set-up + per-iteration time per length class"},
        "tests_per_hour_ratio": round(p_makespan / makespan, 3),
        "p99_ratio_plain_over_served": round(percentile(p_ttr, 99) / max(1e-9, percentile(ttr, 99)), 3),
        "verified_alone": {"count": len(sample), "res64_equal": equal,
                           "service_model_error_median": round(sorted(model_err)[len(model_err) // 2], 3)
                           if model_err else None},
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--iter-scale", type=float, default=1.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--step-ms", type=float, default=5.0)
    ap.add_argument("--gerbicz-block", type=int, default=1000)
    ap.add_argument("--gerbicz-check", type=int, default=0, help="blocks per check (0: up to 10 per chunk)")
    ap.add_argument("--verify", type=int, default=24, help="requests also run alone for real")
    ap.add_argument("--tune-loops", type=int, default=20,
                    help="tune each NTT length on first use with this many loops per plan (0: default plans)")
    ap.add_argument("--json", default="")
    a = ap.parse_args(argv)
    out = served_mix(a.requests, a.rate, a.seed, a.iter_scale, a.budget_gb, a.step_ms, a.gerbicz_block,
                     a.gerbicz_check, a.verify, a.tune_loops)
    line = json.dumps(out)
    print(line, flush=True)
    if a.json:
        with open(a.json, "w") as f:
            f.write(line + "\n")
    ok = out["served"]["errors"] == 0 and out["verified_alone"]["res64_equal"] == out["verified_alone"]["count"]
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
