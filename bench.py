"""Benchmark: the matcher on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--sections plain,batch,serve] [--serve-queries N]

plain (the headline `value`): a step is one pass of the single-query matcher
over the 20 queries of the power-law workload (R-MAT scale 18, edge factor 8,
16 vertex labels Zipf(1.0), 1 edge label, labelled mode), each restricted to
its fixed root window (the last K root candidates by id; see
gfsm_amd/workloads.py), counting embeddings on the device.  The data graph is
resident in HBM before the timed region.  Also reported: the roofline of the
dominant kernel (the expand kernel: algorithmic bytes per launch / mean launch
time, timed with HIP events on the matcher's stream inside the timed region)
and the CPU oracle on a bounded sample of the same workload on this host.

batch: the 256 queries of the batched workload (R-MAT scale 22) over their
root windows, once one at a time on the single-query path and once in one
batched call (autotuned shapes, captured steps); queries/s of both.

serve: the served stream (R-MAT scale 23, edge factor 16) -- its queries run
one at a time on the single-query path (serial), then through the server
with all queries submitted at once (saturated throughput), then as an
open-loop Poisson stream; queries/s and p50/p99 latency of each, and whether
every uncapped query's count and digest are the same served and serial.
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
    p.add_argument("--sections", default="plain,batch,serve",
                   help="comma-separated sections to run (plain is the headline value)")
    p.add_argument("--serve-queries", type=int, default=10000, help="length of the served stream")
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


def pct(xs, p):
    import numpy as np

    return float(np.percentile(np.asarray(xs, dtype=float), p)) if len(xs) else 0.0


def batch_section():
    """The batched workload: serial single-query pass vs one batched call."""
    import gfsm_amd as G
    import gfsm_amd.workloads as W

    data = W.config4_data()
    qs = W.config4_queries(data)
    wins = W.config4_windows(data, qs)
    caps = W.config4_caps(len(qs))
    mc = [c for c, _ in caps]
    ml = [l for _, l in caps]
    ctx = G.Context(0, mem_budget=64 << 30)
    ctx.upload(data)
    # autotune the batch shapes on the workload's own queries (untimed): smaller samples of these
    # mixed 4-16 vertex queries picked shapes that were slower on the whole batch
    tuned = ctx.autotune(data, qs, windows=wins, caps=mc, depth_caps=ml, which=G.TUNE_BATCH)
    ctx.match_batch(data, qs[:16], windows=wins[:16], caps=mc[:16], depth_caps=ml[:16])  # engine set-up
    for i in range(16):  # single-query arena set-up, likewise outside the timing
        ctx.match(data, qs[i], root_lo=wins[i][0], root_hi=wins[i][1], max_cand=mc[i], max_depth_cand=ml[i],
                  allow_partial=True)
    # each side timed twice, the faster pass kept (both are whole passes over the 256 queries)
    t_ser = t_bat = float("inf")
    for _ in range(2):
        t = time.perf_counter()
        ser = [ctx.match(data, q, root_lo=lo, root_hi=hi, max_cand=c, max_depth_cand=l, allow_partial=True,
                         digest=True) for q, (lo, hi), c, l in zip(qs, wins, mc, ml)]
        t_ser = min(t_ser, time.perf_counter() - t)
        t = time.perf_counter()
        bat = ctx.match_batch(data, qs, windows=wins, caps=mc, depth_caps=ml, digest=True)
        t_bat = min(t_bat, time.perf_counter() - t)
    equal = sum((b.count, b.digest, b.status) == (s_.count, s_.digest, s_.status) for b, s_ in zip(bat, ser))
    emb = sum(m.count for m in bat)
    ctx.close()
    return {
        "workload": "batched: R-MAT scale 22, edge factor 8, 32 vertex labels Zipf(1.2), 4 edge labels; "
                    "256 random-walk queries of 4-16 vertices over fixed root windows; 30 of them capped "
                    "deterministically at 2^22 candidates per depth",
        "queries": len(qs),
        "batched_queries_per_s": len(qs) / t_bat,
        "batched_embeddings_per_s": emb / t_bat,
        "batched_ms": t_bat * 1e3,
        "serial_queries_per_s": len(qs) / t_ser,
        "serial_ms": t_ser * 1e3,
        "speedup_vs_serial": t_ser / t_bat,
        "equal_to_serial": f"{equal}/{len(qs)}",
        "capped": sum(m.status == 5 for m in bat),
        "tuned_shape": {k: tuned[k] for k in ("batch_items", "batch_pool_words", "batch_wgs_per_cu", "batch_slots")},
        "tuned_on": "the same 256 queries, before timing",
    }


def serve_section(n):
    """The served stream: serial single-query pass, saturated server, and an
    open-loop Poisson stream."""
    import numpy as np

    import gfsm_amd as G
    import gfsm_amd.workloads as W

    data = W.config5_data()
    qs = W.config5_queries(data, n)
    wins = W.config5_windows(data, qs)
    cap, last = W.CONFIG5_CAP, W.CONFIG5_DEPTH_CAP
    ctx = G.Context(0, mem_budget=16 << 30)
    ctx.upload(data)
    # batch shapes tuned on the first 256 queries of the stream (the server runs on the batch engine)
    tuned = ctx.autotune(data, qs[:256], windows=wins[:256], caps=[cap] * 256, depth_caps=[last] * 256,
                         which=G.TUNE_BATCH)
    for q, (lo, hi) in list(zip(qs, wins))[:32]:  # warm the single-query path
        ctx.match(data, q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last, allow_partial=True)
    svc, ser = [], []
    t = time.perf_counter()
    for q, (lo, hi) in zip(qs, wins):
        t1 = time.perf_counter()
        ser.append(ctx.match(data, q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last, allow_partial=True,
                             digest=True))
        svc.append(time.perf_counter() - t1)
    t_ser = time.perf_counter() - t
    budget = 200 * 10**9
    srv = G.Server(ctx, data, mem_budget=budget, max_inflight=256, digest=True)
    try:
        t = time.perf_counter()
        tk = [srv.submit(q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last) for q, (lo, hi) in zip(qs, wins)]
        sat = [srv.wait(x) for x in tk]
        t_sat = time.perf_counter() - t
        rate = 0.8 * n / t_sat
        arr = W.poisson_arrivals(n, rate)
        tk = []
        t0 = time.perf_counter()
        for a, q, (lo, hi) in zip(arr, qs, wins):
            d = a - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            tk.append(srv.submit(q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last))
        ol = [srv.wait(x) for x in tk]
        times = [srv.times(x) for x in tk]
    finally:
        srv.stop()
    lat = [t2 - t0_ for t0_, t1, t2 in times]
    # the same arrivals served one at a time (FIFO) with the measured serial times
    done, slat = 0.0, []
    for a, s_ in zip(arr, svc):
        done = max(done, a) + s_
        slat.append(done - a)
    capped = [i for i in range(n) if ser[i].status == 5]
    key = lambda m: (m.count, m.digest, m.status)
    eq_sat = sum(key(a) == key(b) for a, b in zip(sat, ser))
    eq_ol = sum(key(a) == key(b) for a, b in zip(ol, ser))
    ctx.close()
    return {
        "workload": f"served: R-MAT scale 23, edge factor 16, 64 vertex labels Zipf(1.0), 4 edge labels, resident; "
                    f"{n} random-walk queries of 4-24 vertices (4 + geometric, mean 8), each over the highest-id roots "
                    f"with <= {W.CONFIG5_FIRST} depth-1 candidates, capped deterministically at {last} candidates "
                    f"per depth; "
                    f"admission by projected memory against a {budget // 10**9} GB budget, 256 slots",
        "queries": n,
        "served_queries_per_s": n / t_sat,
        "serial_queries_per_s": n / t_ser,
        "served_vs_serial": t_ser / t_sat,
        "open_loop_rate_per_s": rate,
        "open_loop_p50_ms": pct(lat, 50) * 1e3,
        "open_loop_p99_ms": pct(lat, 99) * 1e3,
        "serial_same_arrivals_p50_ms": pct(slat, 50) * 1e3,
        "serial_same_arrivals_p99_ms": pct(slat, 99) * 1e3,
        "serial_service_p50_ms": pct(svc, 50) * 1e3,
        "capped": len(capped),
        "equal_saturated": f"{eq_sat}/{n}",
        "equal_open_loop": f"{eq_ol}/{n}",
        "embeddings": int(sum(m.count for m in sat)),
        "tuned_shape": {k: tuned[k] for k in ("batch_items", "batch_pool_words", "batch_wgs_per_cu")},
    }


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: the matcher runs on one GPU; --gpus {args.gpus} is not supported "
              f"(several GPUs would each take a share of the root candidates; not implemented)", file=sys.stderr)
        sys.exit(2)
    import gfsm_amd as G
    import gfsm_amd.workloads as W

    sections = set(args.sections.split(","))
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
            "kernel": "gfsm::k_expand (inner depths) + gfsm::k_expand_bare (last depth)",
            "launches": launches // args.steps,
            "mean_launch_ms": mean_ms,
        },
    }
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(data, queries, windows, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    if "batch" in sections:
        line["batched"] = batch_section()
    if "serve" in sections:
        line["served"] = serve_section(args.serve_queries)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
