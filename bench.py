#!/usr/bin/env python3
"""Throughput of the per-iteration squaring mod 2^p - 1 on one MI355X.

Workload (the reference's single-test throughput configuration): LL test of M_p with
p = smallest prime >= 70,000,000 + (seed mod 10^5) (p = 70,000,027 for seed 0), NTT length
2^22 u64 words as the reference's length selection assigns (~16.7 bits/word), start
residue 4.  A step is one LL iteration (one squaring and -2) over the whole residue, which
stays resident in HBM.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--seed S] [--cpu-seconds T]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/iter @2^22-word NTT (+% HBM roofline)"
HBM_PEAK_GBS = 8000.0
# integer VALU: 256 CUs x 64 lanes x 2.4 GHz (one 32-bit lane-op per lane per clock)
VALU_PEAK = 256 * 64 * 2.4e9


def is_prime(x: int) -> bool:
    if x < 2:
        return False
    if x % 2 == 0:
        return x == 2
    i = 3
    while i * i <= x:
        if x % i == 0:
            return False
        i += 2
    return True


def next_prime(x: int) -> int:
    while not is_prime(x):
        x += 1
    return x


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p: int, n: int, seconds: float):
    """The oracle's C restatement of the same weighted NTT squaring (OpenMP over rows and
    columns), timed on this host for a bounded sample of iterations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import OracleTest, oracle  # test infrastructure: the CPU baseline leg
    L = oracle()
    o = OracleTest(p, n, False, n0=1 << (int(math.log2(n)) // 2))
    t0 = time.perf_counter()
    o.iterate(2)
    per = (time.perf_counter() - t0) / 2
    k = max(2, int(seconds / max(per, 1e-6)))
    t0 = time.perf_counter()
    o.iterate(k)
    dt = time.perf_counter() - t0
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(dt * 1000.0 / k, 3), "unit": "ms/iter", "cores": int(L.lo_threads()),
            "kind": "port",
            "sample": f"{k} LL iterations of M{p} at n=2^{int(math.log2(n))} after 2 warm-up "
                      f"iterations ({dt:.1f} s), oracle/lnt_oracle.c, CPU: {cpu}"}


def batched_leg(seed: int, iters: int, chunk: int = 1000):
    """The batched uneven workload: 64 LL tests, exponents log-uniform in [1M, 20M] from
    splitmix64(seed), 10,000 iterations each; all advance together (one hipGraph launch per
    iteration) versus the same tests one at a time on the plain path.  Both timings include
    each test's set-up (tables) and its final res64 readout."""
    from cuda_amd import Batch, LucasPRP
    from cuda_amd.workloads import batched_exponents
    ps = batched_exponents(seed)
    t0 = time.perf_counter()
    batch = Batch()
    ts = [LucasPRP(p) for p in ps]
    for t in ts:
        batch.add(t)
    done = 0
    while done < iters:
        k = min(chunk, iters - done)
        batch.iterate(k)
        done += k
    batch.sync()
    res_b = [t.res64() for t in ts]
    tb = time.perf_counter() - t0
    batch.close()
    for t in ts:
        t.close()
    log(f"bench: batched {len(ps)} tests x {iters} iterations in {tb:.2f} s")
    t0 = time.perf_counter()
    res_p = []
    for p in ps:
        t = LucasPRP(p)
        t.iterate(iters)
        res_p.append(t.res64())
        t.close()
    tp = time.perf_counter() - t0
    log(f"bench: one at a time {tp:.2f} s")
    gmp_checked = gmp_equal = 0
    try:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from helpers import golden
        for p, r in zip(ps, res_b):
            g = golden(p, "LL") or {}
            if iters in g:
                gmp_checked += 1
                gmp_equal += int(g[iters] == r)
    except Exception as e:  # fixtures are optional for the measurement
        log("bench: no fixture check:", e)
    return {"workload": f"{len(ps)} LL tests, p log-uniform in [1e6, 2e7] (splitmix64 seed {seed}), "
                        f"n = 2^{min(t for t in map(secured, ps))}..2^{max(map(secured, ps))}, "
                        f"{iters} iterations each",
            "tests": len(ps), "iterations": iters,
            "batched_s": round(tb, 3), "plain_s": round(tp, 3),
            "tests_per_hour": round(len(ps) / tb * 3600, 1),
            "plain_tests_per_hour": round(len(ps) / tp * 3600, 1),
            "speedup": round(tp / tb, 3),
            "res64_equal_to_plain": res_b == res_p,
            "res64_checked_against_gmp": gmp_checked, "res64_equal_to_gmp": gmp_equal}


def secured(p: int) -> int:
    from cuda_amd import secured_log2_length
    return secured_log2_length(p)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--profile-iters", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--exponent", type=int, default=0,
                    help="measure this exponent instead (e.g. 140000041 with --prp: the 2^23 PRP test)")
    ap.add_argument("--prp", action="store_true", help="PRP (start 3, no -2) instead of LL")
    ap.add_argument("--batch-iters", type=int, default=10000,
                    help="iterations per test of the batched workload (0: skip that leg)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain copy-like kernel (int64 negate) over MIB MiB, three "
                         "times: the reference point that scales HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the hot path is a single test on one device "
              f"(independent tests shard across GPUs without communication); --gpus {args.gpus} "
              f"is not run.", file=sys.stderr)
        return 2

    import cuda_amd
    from cuda_amd import LucasPRP

    if args.copy_calibration:
        import torch
        words = args.copy_calibration * (1 << 20) // 8
        a = torch.arange(words, dtype=torch.int64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        log(f"bench: copy calibration: 3 x int64 negate of {args.copy_calibration} MiB "
            f"(reads {8 * words} B, writes {8 * words} B per launch)")
        del a, b

    p = args.exponent or next_prime(70_000_000 + args.seed % 100_000)
    t = LucasPRP(p, prp=args.prp)
    n = t.n
    log(f"bench: p={p} n={n} plan={t.describe()}")

    t.iterate(args.warmup)
    t.sync()
    t0 = time.perf_counter()
    t.sync()
    t.iterate(args.steps)
    gpu_ms = t.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    gpu_ms_per_step = gpu_ms / args.steps
    done = args.warmup + args.steps

    # parity spot check against the GMP fixture, when this iteration count has one
    check = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from helpers import golden
        g = golden(p, "PRP" if args.prp else "LL") or {}
        if done in g:
            check = t.res64() == g[done]
    except Exception as e:  # fixtures are optional for the measurement
        log("bench: no fixture check:", e)

    prof = t.profile(args.profile_iters)
    name, kms = max(prof, key=lambda kv: kv[1])
    alg_bytes = 16 * n  # one read and one write of the n-word residue per launch
    achieved = alg_bytes / (kms * 1e-3) / 1e9
    iter_ms = sum(v for _, v in prof)
    # HBM bytes and VALU lane-ops per launch of each kernel, from the rocprofv3 counter runs
    # summarised in profiles/ (FETCH_SIZE / WRITE_SIZE scaled by the copy calibration)
    counters = {}
    tfile = os.path.join(ROOT, "profiles", "traffic_2p22.json")
    if n == 1 << 22 and os.path.exists(tfile):
        try:
            counters = json.load(open(tfile)).get("kernels", {}).get(name.split("<")[0], {})
        except Exception:
            counters = {}
    traffic = counters.get("traffic_bytes")
    valu = None
    if counters.get("valu_lane_ops"):
        va = counters["valu_lane_ops"] / (kms * 1e-3)
        valu = {"achieved": round(va / 1e12, 2), "peak": VALU_PEAK / 1e12, "unit": "T lane-ops/s",
                "frac": round(va / VALU_PEAK, 4), "lane_ops_per_launch": counters["valu_lane_ops"]}
    line = {
        "metric": METRIC,
        "value": round(ms_per_step, 5),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"{'PRP' if args.prp else 'LL'} test of M{p}, n=2^{n.bit_length() - 1} u64 "
                               f"words (GF(2^64-2^32+1) IBDWT), iterations {args.warmup + 1}..{done}",
                   "p": p, "n": n, "plan": t.describe()},
        "gpu_ms_per_step": round(gpu_ms_per_step, 5),
        "res64": f"{t.res64():016X}",
        "res64_matches_gmp": check,
        "kernels_ms": {k: round(v, 5) for k, v in prof},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "valu": valu,
                     "iteration": {"bytes": 32 * n, "ms": round(iter_ms, 5),
                                   "achieved": round(32 * n / (iter_ms * 1e-3) / 1e9, 1),
                                   "frac": round(32 * n / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
    }
    if args.batch_iters:
        line["batched"] = batched_leg(args.seed, args.batch_iters)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p, n, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
