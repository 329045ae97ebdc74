"""Benchmark: the matcher on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--sections plain,grid,batch,serve] [--serve-queries N]

plain (the headline `value`): a step is one pass of the single-query matcher
over the 20 queries of the power-law workload (R-MAT scale 18, edge factor 8,
16 vertex labels Zipf(1.0), 1 edge label, labelled mode), each over three
root windows drawn across the degree range (see workloads.config2_sample):
the lowest-id roots (R-MAT hubs) and the roots from the median on, both under
the deterministic per-depth candidate cap, and the fixed tail window, exact;
embeddings counted on the device.  The data graph is resident in HBM before
the timed region.  Reported beside it: rates per stratum, the share of the
workload's roots and depth-1 candidates the sample covers, the roofline of
the dominant kernel (the expand kernel: algorithmic bytes per launch / mean
launch time, timed with HIP events on the matcher's stream inside the timed
region), and the CPU oracle on complete windows of the same sample on this
host's cores.

grid: the road-like workload (1400 x 1400 damaged grid; C4, C6, P7, 3x3
patch over every root), per-query embeddings/s and the expand roofline.

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
    p.add_argument("--sections", default="plain,grid,batch,serve",
                   help="comma-separated sections to run (plain is the headline value)")
    p.add_argument("--serve-queries", type=int, default=10000, help="length of the served stream")
    p.add_argument("--calibrate", choices=("stream", "gather", "gather2"), default=None,
                   help="only run libgfsm's calibration kernel of known HBM footprint, 5 launches "
                        "(scales the byte counters of a profiler run): stream = 1 GiB read and written, "
                        "gather = one word from each of 2^24 distinct 128-byte lines (2 GiB of lines), "
                        "gather2 = both 64-byte halves of the same lines")
    return p.parse_args()


def calibrate(mode):
    """libgfsm's calibration kernel of known HBM footprint, 5 launches."""
    import gfsm_amd as G

    with G.Context(0) as ctx:
        n = (1 << 28) if mode == "stream" else (1 << 24)
        rd, wr = ctx.selftest_traffic({"stream": 0, "gather": 1, "gather2": 2}[mode], n, 5)
    print(json.dumps({"calibrate": mode, "launches": 5, "read_bytes_per_launch": rd, "write_bytes_per_launch": wr}))


def traffic_from_profiles():
    """HBM bytes per expand launch of the timed sample, measured with rocprofv3
    FETCH_SIZE/WRITE_SIZE and scaled by the gather calibration
    (profiles/expand_config2_traffic.json), and where that figure comes from;
    (None, None) when not recorded.  Counters need a profiler run of their
    own, so the figure is not measured in the run that prints it."""
    path = os.path.join(ROOT, "profiles", "expand_config2_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        src = (f"rocprofv3 FETCH_SIZE x {rec['fetch_factor']} + WRITE_SIZE x {rec['write_factor']} per expand "
               f"launch of this sample on commit {rec['commit']} ({rec['source']}); not measured in this run")
        return float(rec["hbm_bytes_per_launch"]), src
    except (OSError, KeyError, ValueError):
        return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_time(fn, runs):
    ts = []
    out = None
    for _ in range(runs):
        t = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts[len(ts) // 2], out


# Complete tail windows of the timed sample that the CPU oracle finishes in a
# few seconds on all cores: the CPU baseline runs exactly these windows to
# completion (query indices of the power-law workload).
CPU_SAMPLE_QUERIES = (0, 6, 11, 12, 18, 19)


def cpu_baseline(data, queries, sample, gpu_window_s, runs=5):
    """The CPU oracle on this host: the same complete windows as the GPU line
    (CPU_SAMPLE_QUERIES' tail windows), all cores and one thread, median of
    `runs`; plus the reference's host-side costs, text parse and CCSR build
    (HnHelper.cuh:496, main.cu:31-46), timed on the oracle's restatement of
    them and on the product's parser."""
    import tempfile

    import gfsm_amd as G

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = os.cpu_count() or 1
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "data.g")
        data.write(path)
        t_parse, edges = median_time(lambda: O.Edges.load(path), runs)
        t_build, od = median_time(lambda: O.CCSR(edges, O.MODE_VLABEL), runs)
        t_product, _ = median_time(lambda: G.Graph.load(path, mode=G.MODE_VLABEL), runs)
    wins = [(i, lo, hi) for i, st, lo, hi, cap in sample if st == "tail" and i in CPU_SAMPLE_QUERIES]
    plans = []
    for i, lo, hi in wins:
        qV, qs, qd, ql, qvl = queries[i].edges()
        qc = O.CCSR(O.Edges.from_arrays(qV, qs, qd, ql, qvl), O.MODE_VLABEL)
        plans.append((qc, O.Plan(qc), lo, hi))

    def run(threads):
        return sum(O.match(od, qc, p, root_lo=lo, root_hi=hi, threads=threads).count for qc, p, lo, hi in plans)

    t_all, emb = median_time(lambda: run(cores), runs)
    t_one, emb1 = median_time(lambda: run(1), runs)
    assert emb1 == emb
    gpu_s = sum(gpu_window_s[(i, "tail")] for i, _, _ in wins)
    return {
        "value": emb / t_all, "unit": "embeddings/s", "cores": cores, "kind": "port",
        "single_thread_value": emb / t_one,
        "cpu_model": cpu_model(),
        "sample": f"tail windows of queries {list(CPU_SAMPLE_QUERIES)} of the timed sample, each run to completion "
                  f"({emb} embeddings); median of {runs} runs on all {cores} host cores (OpenMP over roots) and "
                  f"on one thread",
        "all_cores_s": t_all, "single_thread_s": t_one,
        "gpu_same_windows_s": gpu_s,
        "gpu_vs_all_cores": t_all / gpu_s if gpu_s else None,
        "host_parse_s": {"oracle_text_parse": t_parse, "oracle_ccsr_build": t_build,
                         "product_parse_and_layout": t_product,
                         "what": "R-MAT scale 18 data graph as t/v/e text, median of 5"},
    }


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
        counters = srv.stats()
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
        "admissions": {k: counters[k] for k in ("admitted", "waited_memory", "waited_slot")},
        "tuned_shape": {k: tuned[k] for k in ("batch_items", "batch_pool_words", "batch_wgs_per_cu")},
    }


def expand_roofline(alg, kms, launches):
    """Algorithmic bytes per expand launch over the mean launch time."""
    per_launch = alg / max(launches, 1)
    mean_ms = kms / max(launches, 1)
    achieved = per_launch / (mean_ms / 1e3) / 1e9 if mean_ms else 0.0
    return achieved, mean_ms, per_launch


def grid_section(steps):
    """The road-like workload: C4, C6, P7 and the 3x3 patch over every root of
    the damaged 1400 x 1400 grid, count mode, single-query path; the levels
    exceed any fixed buffer, so chunked expansion runs throughout."""
    import gfsm_amd as G
    import gfsm_amd.workloads as W

    data = W.config3_data()
    qs = W.config3_queries()
    ctx = G.Context(0, mem_budget=64 << 30)
    ctx.upload(data)
    out = {}
    tot_alg = tot_kms = tot_l = 0
    for name, q in qs.items():
        ctx.match(data, q)  # warm-up (arena, dense tables)
        emb = cand = alg = launches = chunks = 0
        kms = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            m = ctx.match(data, q, profile=True)
            st = m.stats
            emb, cand = m.count, sum(st["cand"])
            alg += st["alg_bytes"] - 4 * st["roots"]
            kms += st["expand_ms"]
            launches += st["expand_launches"]
            chunks = st["chunks"]
        wall = (time.perf_counter() - t0) / steps
        ach, mean_ms, _ = expand_roofline(alg, kms, launches)
        out[name] = {"embeddings": emb, "candidates": cand, "ms": wall * 1e3, "embeddings_per_s": emb / wall,
                     "expand_launches": launches // steps, "chunked_launches": chunks,
                     "expand_achieved_gbs": ach, "expand_frac": ach / HBM_PEAK_GBS}
        tot_alg += alg
        tot_kms += kms
        tot_l += launches
    ctx.close()
    ach, mean_ms, _ = expand_roofline(tot_alg, tot_kms, tot_l)
    return {
        "workload": "road-like: 1400 x 1400 grid, 5% of edges deleted, 2% diagonals added, unlabeled; "
                    "C4, C6, P7, 3x3 patch, every root, count mode (chunked expansion)",
        "queries": out,
        "embeddings_per_s": sum(v["embeddings"] for v in out.values()) / sum(v["ms"] / 1e3 for v in out.values()),
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "launches": tot_l // steps, "mean_launch_ms": mean_ms},
    }


def main():
    args = parse()
    if args.calibrate:
        calibrate(args.calibrate)
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
    sample = W.config2_sample(data, queries)
    ctx = G.Context(0, mem_budget=64 << 30)
    ctx.upload(data)
    strata = W.CONFIG2_STRATA

    def step(profile, acc):
        for i, stratum, lo, hi, cap in sample:
            t = time.perf_counter()
            m = ctx.match(data, queries[i], root_lo=lo, root_hi=hi, max_depth_cand=cap, allow_partial=True,
                          profile=profile)
            dt = time.perf_counter() - t
            st = m.stats
            a = acc[stratum]
            a["emb"] += m.count
            a["cand"] += sum(st["cand"])
            a["roots"] += st["cand"][0]
            a["first"] += st["cand"][1] if len(st["cand"]) > 1 else 0
            a["s"] += dt
            a["kms"] += st["expand_ms"]
            a["launches"] += st["expand_launches"]
            a["alg"] += st["alg_bytes"] - 4 * st["roots"]  # the root kernel's share is not the expand kernel's
            a["capped"] += m.status == 5
            win_s[(i, stratum)] = win_s.get((i, stratum), 0.0) + dt

    def fresh():
        return {s: dict(emb=0, cand=0, roots=0, first=0, s=0.0, kms=0.0, launches=0, alg=0, capped=0) for s in strata}

    win_s = {}
    for _ in range(args.warmup):
        step(True, fresh())
    win_s = {}
    acc = fresh()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True, acc)
    wall = time.perf_counter() - t0
    K = args.steps
    win_s = {k: v / K for k, v in win_s.items()}
    emb = sum(a["emb"] for a in acc.values())
    alg = sum(a["alg"] for a in acc.values())
    kms = sum(a["kms"] for a in acc.values())
    launches = sum(a["launches"] for a in acc.values())
    achieved, mean_ms, _ = expand_roofline(alg, kms, launches)
    # coverage of the whole workload (every root of every query)
    roots_total = first_total = 0
    for q in queries:
        nodes, cand = data.projection(q)
        roots_total += int(cand[0])
        first_total += int(cand[1])
    per = {}
    for sname, a in acc.items():
        ach, mms, _ = expand_roofline(a["alg"], a["kms"], a["launches"])
        per[sname] = {"embeddings_per_step": a["emb"] // K, "candidates_per_step": a["cand"] // K,
                      "ms_per_step": a["s"] * 1e3 / K, "embeddings_per_s": a["emb"] / a["s"] if a["s"] else 0.0,
                      "candidates_per_s": a["cand"] / a["s"] if a["s"] else 0.0,
                      "expand_achieved_gbs": ach, "expand_frac": ach / HBM_PEAK_GBS,
                      "roots_tested": a["roots"] // K, "capped_windows": a["capped"] // K}
    roots_tested = sum(a["roots"] for a in acc.values()) // K
    first_tested = sum(a["first"] for a in acc.values()) // K
    line = {
        "metric": "embeddings/s + queries/s on 1 MI355X; p50/p99 latency, served vs serial",
        "value": emb / wall,
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded R-MAT graph and random-walk queries)",
        "config": {
            "workload": "power-law plain single query: R-MAT scale 18, edge factor 8, 16 vertex labels "
                        "Zipf(1.0), 1 edge label; 20 random-walk queries of 8 vertices (10 sparse, 10 dense), each "
                        f"over a stratified root sample: the {W.CONFIG2_STRATUM_ROOTS} lowest-id roots (hub end) and "
                        f"{W.CONFIG2_STRATUM_ROOTS} roots from the median root on, both under the deterministic "
                        f"per-depth cap of {W.CONFIG2_STRATUM_CAP} candidates, and the fixed tail window (the last "
                        "1-256 roots by id), exact; count mode",
            "queries_per_step": len(queries),
            "windows_per_step": len(sample),
            "embeddings_per_step": emb // K,
            "candidates_per_step": sum(a["cand"] for a in acc.values()) // K,
            "coverage": {"roots_tested": roots_tested, "roots_total": roots_total,
                         "roots_frac": roots_tested / roots_total,
                         "depth1_candidates_tested": first_tested, "depth1_candidates_total": first_total,
                         "depth1_frac": first_tested / first_total,
                         "note": "of all 20 queries over all their roots; hub and middle windows test only a "
                                 "capped prefix of their roots' subtrees"},
            "strata": per,
        },
        "queries_per_s": len(sample) * K / wall,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "gfsm::k_expand (inner depths) + gfsm::k_expand_bare (last depth)",
            "launches": launches // K,
            "mean_launch_ms": mean_ms,
        },
    }
    ctx.close()
    line["roofline"]["traffic"], line["roofline"]["traffic_source"] = traffic_from_profiles()
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(data, queries, sample, win_s)
    if "grid" in sections:
        line["grid"] = grid_section(max(1, min(K, 3)))
    if "batch" in sections:
        line["batched"] = batch_section()
    if "serve" in sections:
        line["served"] = serve_section(args.serve_queries)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
