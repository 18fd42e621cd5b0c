#!/usr/bin/env python3
"""Throughput of the per-iteration squaring mod 2^p - 1 on one MI355X.

Workload (the reference's single-test throughput configuration): LL test of M_p with
p = smallest prime >= 70,000,000 + (seed mod 10^5) (p = 70,000,027 for seed 0), NTT length
2^22 u64 words as the reference's length selection assigns (~16.7 bits/word), start
residue 4.  A step is one LL iteration (one squaring and -2) over the whole residue, which
stays resident in HBM.  With the default warm-up and step counts the run ends on
iteration 3,000, whose res64 is checked against the GMP fixture.  Prints one JSON line.

Further legs in the same line (each can be switched off):
  batched  64 LL tests of p log-uniform in [1M, 20M], 10,000 iterations each, as one batch
           and one at a time on the plain path;
  served   the first requests of the served mix (Poisson 4/s) through the serving loop at a
           reduced iteration scale, against the same stream one at a time;
  checked  the PRP test of p = 140,000,041 at 2^23 with and without the Gerbicz check;
  cpu_baseline  the oracle's C restatement on the host cores, a bounded sample.

  python bench.py [--steps K] [--warmup W] [--seed S] [--no-cpu] [--batch-iters 0]
                  [--served-requests 0] [--no-checked]
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
LANES_PER_CU_CLOCK = 64  # one 32-bit VALU instruction per lane per clock and CU (4 SIMD x 16)


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


def golden_res64(p: int, mode: str):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        from helpers import golden  # the GMP fixtures
        return golden(p, mode) or {}
    except Exception as e:  # fixtures are optional for the measurement
        log("bench: no fixture check:", e)
        return {}


def cpu_baseline(p: int, n: int, seconds: float, gpu_check):
    """The oracle's C restatement of the same weighted NTT squaring (OpenMP over rows and
    columns), timed on this host for a bounded sample of iterations after 10 warm-up
    iterations; its res64 is compared with the GPU's after the same number of iterations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import OracleTest, oracle  # test infrastructure: the CPU baseline leg
    L = oracle()
    o = OracleTest(p, n, False, n0=1 << (int(math.log2(n)) // 2))
    warm = 10
    t0 = time.perf_counter()
    o.iterate(warm)
    per = (time.perf_counter() - t0) / warm
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
    r_cpu = o.res64()
    r_gpu = gpu_check(warm + k)
    return {"value": round(dt * 1000.0 / k, 3), "unit": "ms/iter", "cores": int(L.lo_threads()),
            "kind": "port",
            "sample": f"{k} LL iterations of M{p} at n=2^{int(math.log2(n))} after {warm} warm-up "
                      f"iterations ({dt:.1f} s), oracle/lnt_oracle.c, CPU: {cpu}",
            "res64": f"{r_cpu:016X}", "res64_equal_to_gpu": r_cpu == r_gpu}


def batched_leg(seed: int, iters: int, chunk: int = 1000, tune_loops: int = 0):
    """The batched uneven workload: 64 LL tests, exponents log-uniform in [1M, 20M] from
    splitmix64(seed), 10,000 iterations each; all advance together (one hipGraph launch per
    iteration) versus the same tests one at a time on the plain path.  Both timings include
    each test's set-up (tables) and its final res64 readout; with tune_loops each length is
    tuned once beforehand (untimed) and both sides run the tuned plans."""
    from cuda_amd import Batch, LucasPRP, secured_log2_length
    from cuda_amd.workloads import batched_exponents
    ps = batched_exponents(seed)
    if tune_loops:
        for p in sorted({secured_log2_length(q): q for q in ps}.values()):
            u = LucasPRP(p)
            u.tune(tune_loops)
            u.close()
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
    for p, r in zip(ps, res_b):
        g = golden_res64(p, "LL")
        if iters in g:
            gmp_checked += 1
            gmp_equal += int(g[iters] == r)
    return {"workload": f"{len(ps)} LL tests, p log-uniform in [1e6, 2e7] (splitmix64 seed {seed}), "
                        f"n = 2^{min(map(secured, ps))}..2^{max(map(secured, ps))}, "
                        f"{iters} iterations each",
            "tests": len(ps), "iterations": iters,
            "batched_s": round(tb, 3), "plain_s": round(tp, 3),
            "tests_per_hour": round(len(ps) / tb * 3600, 1),
            "plain_tests_per_hour": round(len(ps) / tp * 3600, 1),
            "speedup": round(tp / tb, 3),
            "res64_equal_to_plain": res_b == res_p,
            "res64_checked_against_gmp": gmp_checked, "res64_equal_to_gmp": gmp_equal}


def checked_leg(p: int = 140_000_041, block: int = 1000, iters: int = 4000):
    """PRP at 2^23 with the Gerbicz check: ms/iter unchecked, with the block products only
    (blocks per check larger than the run), and with one check every 4 blocks.  The cost of
    a block product and of a check give the overhead at any check interval L; the smallest
    L with overhead <= 2% is reported.  res64 and the checksum at `iters` are compared with
    the GMP fixture."""
    from cuda_amd import LucasPRP

    def run(mode):
        t = LucasPRP(p, prp=True)  # starts from 3
        if mode != "plain":  # set-up of the checkpoint slots, outside the timed region
            t.gerbicz_begin(block, 10 ** 6 if mode == "blocks" else 4)
        t.sync()
        t0 = time.perf_counter()
        if mode == "plain":
            t.iterate(iters)
            t.sync()
            st = None
        else:
            t.gerbicz_iterate(iters)
            t.sync()
            st = t.gerbicz_stats()
        dt = time.perf_counter() - t0
        r = t.res64()
        t.close()
        return dt, r, st
    t_plain, r_plain, _ = run("plain")
    t_blocks, r_blocks, _ = run("blocks")
    t_check, r_check, st = run("check")
    nb = iters // block
    t_iter = t_plain / iters
    block_cost = max(0.0, (t_blocks - t_plain) / nb)
    check_cost = max(0.0, t_check - t_blocks)
    L = 1
    while (L * block_cost + check_cost) / (L * block * t_iter) > 0.02 and L < 10 ** 4:
        L += 1
    g = golden_res64(p, "PRP")
    dg = {}
    try:
        from helpers import golden_cases
        for c in golden_cases():
            if c["p"] == p and "dres64" in c:
                dg.update({int(k): int(v, 16) for k, v in c["dres64"].items()})
    except Exception:
        pass
    return {"workload": f"PRP test of M{p} (n=2^23), {iters} iterations, Gerbicz block {block}",
            "ms_per_iter_unchecked": round(t_plain * 1e3 / iters, 5),
            "ms_per_iter_block_products": round(t_blocks * 1e3 / iters, 5),
            "ms_per_iter_check_every_4_blocks": round(t_check * 1e3 / iters, 5),
            "block_product_ms": round(block_cost * 1e3, 3), "check_ms": round(check_cost * 1e3, 3),
            "overhead_check_every_4_blocks": round(t_check / t_plain - 1, 4),
            "blocks_per_check_for_2pct": L,
            "overhead_at_that_interval": round((L * block_cost + check_cost) / (L * block * t_iter), 4),
            "checks_passed": st["checks"] if st else None,
            "res64_equal": r_plain == r_blocks == r_check,
            "res64_matches_gmp": (g[iters] == r_check) if iters in g else None,
            "checksum_matches_gmp": (dg[iters] == st["checkpoint_res64"]) if (st and iters in dg) else None}


def secured(p: int) -> int:
    from cuda_amd import secured_log2_length
    return secured_log2_length(p)


def counter_doc(n: int):
    """The rocprofv3 counter summary in profiles/ when it was taken at length n, else {}."""
    path = os.path.join(ROOT, "profiles", "traffic_2p22.json")
    try:
        doc = json.load(open(path))
    except (OSError, ValueError):
        return {}
    return doc if doc.get("n") == n else {}


def counters_for(n: int, name: str):
    """HBM bytes and VALU lane-instructions per launch of kernel `name` at length n, from the
    rocprofv3 counter runs summarised in profiles/ (exact kernel name and length only)."""
    return counter_doc(n).get("kernels", {}).get(name, {})


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--profile-iters", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--exponent", type=int, default=0,
                    help="measure this exponent instead (e.g. 140000041 with --prp: the 2^23 PRP test)")
    ap.add_argument("--prp", action="store_true", help="PRP (start 3, no -2) instead of LL")
    ap.add_argument("--batch-iters", type=int, default=10000,
                    help="iterations per test of the batched workload (0: skip that leg)")
    ap.add_argument("--served-requests", type=int, default=120,
                    help="requests of the served mix for the served leg (0: skip it)")
    ap.add_argument("--served-scale", type=float, default=0.1, help="iteration scale of the served leg")
    ap.add_argument("--no-checked", action="store_true", help="skip the Gerbicz-checked PRP leg")
    ap.add_argument("--tune-loops", type=int, default=20,
                    help="tune each length before use, this many loops per plan (0: default plans)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain copy-like kernel (int64 negate) over MIB MiB, three "
                         "times: the reference point that scales HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the hot path is a single test on one device "
              f"(independent tests shard across GPUs without communication); --gpus {args.gpus} "
              f"is not run.", file=sys.stderr)
        return 2

    from cuda_amd import LucasPRP
    from cuda_amd.lnt import valu_rates

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
    mode = "PRP" if args.prp else "LL"
    t = LucasPRP(p, prp=args.prp)
    n = t.n
    tuned = None
    if args.tune_loops:  # as the reference does at start: time every plan, keep the fastest
        t.tune(args.tune_loops)
        tuned = {"loops": args.tune_loops, "candidates_ms": {k: round(v, 4) for k, v in t.tune_results()}}
    plan = t.describe()
    log(f"bench: p={p} n={n} plan={plan}")

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
    res64 = t.res64()
    g = golden_res64(p, mode)
    check = (g[done] == res64) if done in g else None

    # per-kernel durations: each kernel `profile_iters` times back to back as one captured
    # graph, timed by two HIP events on the test's stream (the residue is saved and restored)
    prof = t.profile(args.profile_iters)
    t.close()
    name, kms = max(prof, key=lambda kv: kv[1])
    alg_bytes = 16 * n  # one read and one write of the n-word residue per launch
    achieved = alg_bytes / (kms * 1e-3) / 1e9
    iter_bytes = 32 * n  # two read+write sweeps of the residue per iteration
    rates = valu_rates()
    ghz = rates.pop("ghz")
    counters = counters_for(n, name)
    valu = None
    # VALU roof: the best issue rate the field-operation probes reach (operations per CU per
    # measured clock x their VALU instructions per operation, counted by SQ_INSTS_VALU), in
    # wave64 instructions per CU per clock; 32-bit operations issue faster than one per
    # quad-cycle per SIMD on gfx950, so this exceeds 1 for the mixed 32/64-bit field code
    probe_instr = counter_doc(n).get("valu_probe_instr_per_op", {})
    issue = {k: rates[k] * probe_instr[k] / LANES_PER_CU_CLOCK
             for k in rates if k in probe_instr and k != "v_mad_u64_u32"}
    if counters.get("valu_lane_ops") and issue:
        roof_op = max(issue, key=issue.get)
        peak = 256 * LANES_PER_CU_CLOCK * issue[roof_op] * ghz * 1e9
        va = counters["valu_lane_ops"] / (kms * 1e-3)
        valu = {"achieved": round(va / 1e12, 2), "peak": round(peak / 1e12, 2), "unit": "T lane-instr/s",
                "frac": round(va / peak, 4), "lane_instr_per_launch": counters["valu_lane_ops"],
                "clock_ghz_measured": round(ghz, 3),
                "peak_basis": f"{roof_op} probe: {issue[roof_op]:.3f} wave64 VALU instructions per CU per clock",
                "wave_instr_per_cu_clock": round(va / (256 * LANES_PER_CU_CLOCK * ghz * 1e9), 3),
                "probe_wave_instr_per_cu_clock": {k: round(v, 3) for k, v in issue.items()}}
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
        "config": {"workload": f"{mode} test of M{p}, n=2^{n.bit_length() - 1} u64 words "
                               f"(GF(2^64-2^32+1) IBDWT), iterations {args.warmup + 1}..{done}",
                   "p": p, "n": n, "plan": plan},
        "gpu_ms_per_step": round(gpu_ms_per_step, 5),
        "res64": f"{res64:016X}",
        "res64_matches_gmp": check,
        "kernels_ms": {k: round(v, 5) for k, v in prof},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": counters.get("traffic_bytes"), "valu": valu,
                     "iteration": {"bytes": iter_bytes, "ms": round(ms_per_step, 5),
                                   "achieved": round(iter_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                                   "frac": round(iter_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "valu_field_ops_per_cu_clock": {k: round(v, 3) for k, v in rates.items()},
        "tuned": tuned,
    }
    if args.batch_iters:
        line["batched"] = batched_leg(args.seed, args.batch_iters, tune_loops=args.tune_loops)
    if args.served_requests:
        from cuda_amd.serve import served_mix
        line["served"] = served_mix(requests=args.served_requests, iter_scale=args.served_scale, verify=6,
                                    tune_loops=args.tune_loops)
    if not args.no_checked:
        line["checked"] = checked_leg()
    if not args.no_cpu:
        def gpu_res64_after(k):
            u = LucasPRP(p, prp=args.prp)
            u.iterate(k)
            r = u.res64()
            u.close()
            return r
        line["cpu_baseline"] = cpu_baseline(p, n, args.cpu_seconds, gpu_res64_after)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
