"""Benchmark: plain single-query matching on the power-law workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one pass of the plain matcher over the 20 queries of the power-law
workload (R-MAT scale 18, edge factor 8, 16 vertex labels Zipf(1.0), 1 edge
label, labelled mode), each restricted to its fixed root window (the last K
root candidates by id; see gfsm_amd/workloads.py), counting embeddings on the
device.  Data graph and queries are generated from fixed seeds and the data
graph is resident in HBM before the timed region.

Prints one JSON line: embeddings/s over the timed steps, the roofline of the
dominant kernel (the expand kernel: algorithmic bytes per launch / mean launch
time, timed with HIP events on the matcher's stream inside the timed region),
and the CPU oracle timed on a bounded sample of the same workload on this
host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="only run an elementwise int32 kernel that reads and writes MIB MiB, 5 times "
                        "(scales the HBM byte counters of a profiler run)")
    return p.parse_args()


def calibrate_copy(mib):
    """A kernel of known HBM traffic: out = in + 1 over mib MiB of int32,
    i.e. mib MiB read and mib MiB written per launch."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.zeros(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.add(src, 1, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "launches": 5, "read_bytes_per_launch": n * 4,
                      "write_bytes_per_launch": n * 4}))


def traffic_from_profiles():
    """HBM bytes per expand launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE
    (profiles/expand_config2_traffic.json), or None when not recorded."""
    path = os.path.join(ROOT, "profiles", "expand_config2_traffic.json")
    try:
        with open(path) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(data, queries, windows, budget_s):
    """The CPU oracle on a bounded sample of one pass: each query's window in
    turn, each stopped after at most 2e7 embeddings, until budget_s of
    matching has been spent (embeddings/s on this host's cores)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    V, s, d, l, vl = data.edges()
    od = O.CCSR(O.Edges.from_arrays(V, s, d, l, vl), O.MODE_VLABEL)
    threads = min(os.cpu_count() or 1, 16)
    emb, spent, done = 0, 0.0, 0
    for q, (lo, hi) in zip(queries, windows):
        qV, qs, qd, ql, qvl = q.edges()
        qc = O.CCSR(O.Edges.from_arrays(qV, qs, qd, ql, qvl), O.MODE_VLABEL)
        p = O.Plan(qc)
        t = time.perf_counter()
        r = O.match_capped(od, qc, p, 20_000_000, root_lo=lo, root_hi=hi, threads=threads)
        spent += time.perf_counter() - t
        emb += r.count
        done += 1
        if spent >= budget_s:
            break
    return {"value": emb / spent if spent else 0.0, "unit": "embeddings/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of the 20 windowed queries, each capped at 2e7 embeddings "
                      f"({emb} embeddings in {spent:.1f} s, OpenMP)"}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU design shards roots, see the design notes)", file=sys.stderr)
        sys.exit(2)
    import gfsm_amd as G
    import gfsm_amd.workloads as W

    data = W.config2_data()
    queries = W.config2_queries(data)
    windows = W.config2_windows(data, queries)
    ctx = G.Context(0, mem_budget=64 << 30)
    ctx.upload(data)

    def step(profile):
        emb, cand, kms, launches, alg = 0, 0, 0.0, 0, 0
        for q, (lo, hi) in zip(queries, windows):
            m = ctx.match(data, q, root_lo=lo, root_hi=hi, profile=profile)
            st = m.stats
            emb += m.count
            cand += sum(st["cand"])
            kms += st["expand_ms"]
            launches += st["expand_launches"]
            alg += st["alg_bytes"] - 4 * st["roots"]  # the root kernel's share is not the expand kernel's
        return emb, cand, kms, launches, alg

    for _ in range(args.warmup):
        step(True)
    t0 = time.perf_counter()
    tot = [0, 0, 0.0, 0, 0]
    for _ in range(args.steps):
        r = step(True)
        tot = [a + b for a, b in zip(tot, r)]
    wall = time.perf_counter() - t0
    emb, cand, kms, launches, alg = tot
    per_launch_bytes = alg / max(launches, 1)
    mean_ms = kms / max(launches, 1)
    achieved = per_launch_bytes / (mean_ms / 1e3) / 1e9 if mean_ms else 0.0
    traffic = traffic_from_profiles()
    line = {
        "metric": "embeddings/s + queries/s on 1 MI355X; p50/p99 latency, served vs serial",
        "value": emb / wall,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded R-MAT graph and random-walk queries)",
        "config": {
            "workload": "power-law plain single query: R-MAT scale 18, edge factor 8, 16 vertex labels "
                        "Zipf(1.0), 1 edge label; 20 random-walk queries of 8 vertices (10 sparse, 10 dense), "
                        "each over its fixed root window; count mode",
            "queries_per_step": len(queries),
            "embeddings_per_step": emb // args.steps,
            "candidates_per_step": cand // args.steps,
        },
        "queries_per_s": len(queries) * args.steps / wall,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "gfsm::k_expand",
            "launches": launches // args.steps,
            "mean_launch_ms": mean_ms,
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(data, queries, windows, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
