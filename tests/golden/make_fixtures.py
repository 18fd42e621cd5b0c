#!/usr/bin/env python3
"""Regenerate the golden res64 fixtures with the GMP big-integer oracle (oracle/gmp_ll.c).

Each fixture is computed from the definition of the sequence, independently of any
NTT: LL s_0 = 4, s_k = s_{k-1}^2 - 2 mod M_p; PRP s_0 = 3, s_k = s_{k-1}^2 mod M_p
(lucas_prp.cu:127-134, kernels/m2.cu:23-41).  res64 = low 64 bits of s_k in [0, M_p).

  python tests/golden/make_fixtures.py            # the fast fixtures (~1 minute)
  python tests/golden/make_fixtures.py --long     # also the p = 70M / 140M runs (hours)
  python tests/golden/make_fixtures.py --import-log 70000027 LL 20000 out.txt
        # merge the "k res64" lines of `oracle/_build/gmp_ll 70000027 LL 20000 1000 > out.txt`
        # (a long run made in the background) into the fixtures
  python tests/golden/make_fixtures.py --batched 4 > batched.txt
        # res64 after 10,000 LL iterations of the 64 batched-workload exponents (seed 0),
        # 4 GMP processes (about an hour); prints "p res64" lines as they finish, then merges
  python tests/golden/make_fixtures.py --import-batched batched.txt
  python tests/golden/make_fixtures.py --served 4
        # the served-mix sample (cuda_amd.workloads.served_sample, seed 0): full LL tests
        # (res64 at every quarter and at the end) and PRP chunks with their Gerbicz
        # products (block 1000), 4 GMP processes (about 40 minutes)

Output: tests/golden/res64.json
"""
import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GMP = os.path.join(ROOT, "oracle", "_build", "gmp_ll")
OUT = os.path.join(HERE, "res64.json")

# Known Mersenne prime exponents (LL residue 0 after p-2 iterations, PRP residue 9).
MERSENNE_PRIMES = [521, 607, 1279, 2203, 2281, 3217, 4253, 4423, 9689, 9941, 11213, 19937,
                   21701, 23209, 44497, 86243]


def is_prime(x: int) -> bool:
    if x < 2:
        return False
    i = 2
    while i * i <= x:
        if x % i == 0:
            return False
        i += 1
    return True


def next_prime(x: int) -> int:
    while not is_prime(x):
        x += 1
    return x


# Prime exponents whose Mersenne numbers are composite (the residue is nonzero).
# 131 and 307 are the shortest lengths (n = 8 and 16).
COMPOSITES = [p for p in (131, 307, 523, 1009, 2003, 4001, 8191, 9697, 10007, 20011, 44501, 86249)
              if is_prime(p)]


def gmp(p: int, mode: str, iters: int, every: int) -> dict:
    out = subprocess.run([GMP, str(p), mode, str(iters), str(every)], check=True,
                         capture_output=True, text=True).stdout
    res = {}
    for line in out.strip().splitlines():
        k, r = line.split()
        res[k] = r
    return res


def gmp_gerbicz(p: int, iters: int, block: int):
    """res64 of s_k and of the Gerbicz product D = s_0 s_B s_2B ... at every block end."""
    out = subprocess.run([GMP, str(p), "PRP", str(iters), str(block), "0", str(block)], check=True,
                         capture_output=True, text=True).stdout
    res, dres = {}, {}
    for line in out.strip().splitlines():
        k, r, d = line.split()
        res[k], dres[k] = r, d
    return res, dres


BATCHED_ITERS = 10000


def _batched_one(p: int):
    return p, gmp(p, "LL", BATCHED_ITERS, BATCHED_ITERS)[str(BATCHED_ITERS)]


def batched(workers: int, path) -> int:
    """GMP res64 of the batched workload (64 LL tests, seed 0) after 10,000 iterations."""
    if path:
        rows = [line.split() for line in open(path) if len(line.split()) == 2]
        found = {int(a): b for a, b in rows}
    else:
        sys.path.insert(0, ROOT)
        from multiprocessing import Pool
        from cuda_amd.workloads import batched_exponents
        ps = sorted(set(batched_exponents(0)))
        found = {}
        with Pool(workers) as pool:
            for p, r in pool.imap_unordered(_batched_one, ps):
                print(p, r, flush=True)
                found[p] = r
    doc = json.load(open(OUT))
    have = {(c["p"], c["mode"], c["iters"]) for c in doc["cases"]}
    for p, r in sorted(found.items()):
        if (p, "LL", BATCHED_ITERS) not in have:
            doc["cases"].append({"p": p, "mode": "LL", "iters": BATCHED_ITERS, "batched": True,
                                 "res64": {str(BATCHED_ITERS): r}})
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print(f"merged {len(found)} batched-workload fixtures", file=sys.stderr)
    return 0


def _served_one(req):
    if req.prp:
        res, dres = gmp_gerbicz(req.p, req.iterations, 1000)
        return {"p": req.p, "mode": "PRP", "iters": req.iterations, "block": 1000, "served": True,
                "res64": res, "dres64": dres}
    return {"p": req.p, "mode": "LL", "iters": req.iterations, "served": True,
            "res64": gmp(req.p, "LL", req.iterations, max(1, req.iterations // 4))}


def served(workers: int) -> int:
    """GMP fixtures of the served-mix sample (seed 0)."""
    sys.path.insert(0, ROOT)
    from multiprocessing import Pool
    from cuda_amd.workloads import served_sample
    reqs = served_sample(0)
    found = []
    with Pool(workers) as pool:
        for case in pool.imap_unordered(_served_one, sorted(reqs, key=lambda r: -r.p)):
            print(case["p"], case["mode"], case["iters"], flush=True)
            found.append(case)
    doc = json.load(open(OUT))
    keys = {(c["p"], c["mode"], c["iters"]) for c in found}
    doc["cases"] = [c for c in doc["cases"] if (c["p"], c["mode"], c["iters"]) not in keys or not c.get("served")]
    have = {(c["p"], c["mode"], c["iters"], c.get("block", 0)) for c in doc["cases"]}
    for c in sorted(found, key=lambda c: (c["p"], c["mode"])):
        if (c["p"], c["mode"], c["iters"], c.get("block", 0)) not in have:
            doc["cases"].append(c)
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print(f"merged {len(found)} served-sample fixtures", file=sys.stderr)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--long", action="store_true", help="also run the multi-hour p=70M/140M fixtures")
    ap.add_argument("--import-log", nargs=4, metavar=("P", "MODE", "ITERS", "FILE"))
    ap.add_argument("--block", type=int, default=0, help="with --import-log: the log has Gerbicz columns")
    ap.add_argument("--batched", type=int, default=0, metavar="WORKERS")
    ap.add_argument("--import-batched", metavar="FILE")
    ap.add_argument("--served", type=int, default=0, metavar="WORKERS")
    args = ap.parse_args()
    if args.served:
        return served(args.served)
    if args.batched or args.import_batched:
        return batched(args.batched, args.import_batched)
    if args.import_log:
        p, mode, iters, path = args.import_log
        doc = json.load(open(OUT))
        rows = [line.split() for line in open(path).read().strip().splitlines()]
        res = {r[0]: r[1] for r in rows}
        case = {"p": int(p), "mode": mode, "iters": int(iters), "res64": res}
        if args.block:
            case["block"] = args.block
            case["dres64"] = {r[0]: r[2] for r in rows}
        doc["cases"] = [c for c in doc["cases"] if not (c["p"] == int(p) and c["mode"] == mode
                                                        and c["iters"] == int(iters)
                                                        and c.get("block", 0) == args.block)]
        doc["cases"].append(case)
        with open(OUT, "w") as f:
            json.dump(doc, f, indent=1, sort_keys=True)
        print(f"imported {len(res)} checkpoints for p={p} {mode}")
        return 0
    if not os.path.exists(GMP):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    old = json.load(open(OUT)) if os.path.exists(OUT) else {"cases": []}
    cases = []
    for p in MERSENNE_PRIMES:
        cases.append({"p": p, "mode": "LL", "iters": p - 2, "res64": gmp(p, "LL", p - 2, max(1, (p - 2) // 4))})
        cases.append({"p": p, "mode": "PRP", "iters": p, "res64": gmp(p, "PRP", p, max(1, p // 4))})
    for p in COMPOSITES:
        cases.append({"p": p, "mode": "LL", "iters": p - 2, "res64": gmp(p, "LL", p - 2, max(1, (p - 2) // 4))})
        cases.append({"p": p, "mode": "PRP", "iters": p, "res64": gmp(p, "PRP", p, max(1, p // 4))})
    # the single-test parity workload: LL of M86243, res64 every 10,000 iterations
    cases.append({"p": 86243, "mode": "LL", "iters": 86241, "res64": gmp(86243, "LL", 86241, 1000)})
    # Gerbicz products (PRP): small path n = 4096 and large path n = 2^17
    for p, iters, block in ((86243, 86243, 1000), (2000003, 6000, 200)):
        res, dres = gmp_gerbicz(p, iters, block)
        cases.append({"p": p, "mode": "PRP", "iters": iters, "block": block, "res64": res, "dres64": dres})
    long_cases = [(70000027, "LL", 20000, 1000), (140000041, "PRP", 10000, 1000)]
    for p, mode, iters, every in long_cases:
        if args.long:
            cases.append({"p": p, "mode": mode, "iters": iters, "res64": gmp(p, mode, iters, every)})
        else:
            prev = [c for c in old["cases"] if c["p"] == p and c["mode"] == mode and "block" not in c]
            if prev:
                cases.append(prev[0])
    prev = [c for c in old["cases"] if c["p"] == 140000041 and "block" in c]
    if args.long:
        res, dres = gmp_gerbicz(140000041, 10000, 1000)
        cases.append({"p": 140000041, "mode": "PRP", "iters": 10000, "block": 1000, "res64": res, "dres64": dres})
    elif prev:
        cases.append(prev[0])
    cases += [c for c in old["cases"] if c.get("batched") or c.get("served")]  # --batched, --served
    doc = {
        "generator": "tests/golden/make_fixtures.py (oracle/gmp_ll.c, GMP mpz)",
        "definition": "LL: s0=4, s=s^2-2 mod 2^p-1; PRP: s0=3, s=s^2 mod 2^p-1; res64 = low 64 bits "
                      "of s_k in [0, 2^p-1); keys are iteration numbers k",
        "cases": cases,
    }
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print(f"wrote {OUT}: {len(cases)} cases")
    return 0


if __name__ == "__main__":
    sys.exit(main())
