"""Write the oracle's results for the benchmark workloads (CPU only).

    python tests/golden/make_workload_fixtures.py [config2] [config3] [config4] [config5]

For each named workload it builds the seeded inputs with the product's host
generators (gfsm_amd.workloads), runs the CPU oracle on them, and writes
tests/golden/fixtures_<name>.json: per query (and root window), the embedding
count, the two 64-bit digest sums and whether the per-depth cap stopped it.
The GPU tests compare the HIP path with these numbers.

  config2  the 20 power-law queries over their hub-end, middle and tail-end
           root windows (the timed sample; hub and middle windows under the
           per-depth cap, tails exact)
  config3  C4, C6, P7 and the 3x3 patch on the damaged 1400 x 1400 grid, all
           roots, exact
  config4  the 256 batched queries over their root windows (30 under the
           per-depth cap)
  config5  a sample of the served stream: every CONFIG5_FIXTURE_STRIDE-th
           query over its root window, under the per-depth cap

Uncapped windows use the oracle's parallel enumeration (orc_match); capped
windows use its canonical-order enumeration (orc_match_canon), which is the
only one whose partial result is defined.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import gfsm_amd as G  # noqa: E402
import gfsm_amd.workloads as W  # noqa: E402
import oracle as O  # noqa: E402


def oracle_graph(g, mode):
    V, s, d, l, vl = g.edges()
    return O.CCSR(O.Edges.from_arrays(V, s, d, l, vl), mode)


class Runner:
    def __init__(self, data, mode, canon=True):
        t = time.time()
        self.mode = mode
        self.od = oracle_graph(data, mode)
        self.cn = O.Canon(self.od) if canon else None
        print(f"  oracle graph built in {time.time() - t:.1f} s", flush=True)

    def run(self, query, lo, hi, cap):
        qc = oracle_graph(query, self.mode)
        p = O.Plan(qc)
        t = time.time()
        if cap:
            r, capped, _ = O.match_canon(self.od, self.cn, qc, p, cap, root_lo=lo, root_hi=hi)
        else:
            r, capped = O.match(self.od, qc, p, root_lo=lo, root_hi=hi), False
        return {"root_lo": lo, "root_hi": hi, "max_depth_cand": cap, "count": r.count,
                "digest": [f"{r.digest[0]:016x}", f"{r.digest[1]:016x}"], "capped": capped,
                "oracle_s": round(time.time() - t, 2)}


def write(name, obj):
    path = os.path.join(HERE, f"fixtures_{name}.json")
    with open(path, "w") as f:
        json.dump(obj, f, indent=1)
        f.write("\n")
    print(f"wrote {path}", flush=True)


def config2():
    data = W.config2_data()
    qs = W.config2_queries(data)
    run = Runner(data, G.MODE_VLABEL)
    out = []
    for i, stratum, lo, hi, cap in W.config2_sample(data, qs):
        e = run.run(qs[i], lo, hi, cap)
        e.update(query=i, stratum=stratum)
        out.append(e)
        print(f"  config2 q{i} {stratum}: {e['count']} capped={e['capped']} ({e['oracle_s']} s)", flush=True)
    write("config2", {"workload": "config 2 timed sample (workloads.config2_sample)", "windows": out})


def config3():
    data = W.config3_data()
    run = Runner(data, G.MODE_REF, canon=False)
    out = {}
    for name, q in W.config3_queries().items():
        e = run.run(q, 0, 0xFFFFFFFF, 0)
        out[name] = e
        print(f"  config3 {name}: {e['count']} ({e['oracle_s']} s)", flush=True)
    write("config3", {"workload": "config 3: damaged 1400 x 1400 grid, all roots", "queries": out})


def config4():
    data = W.config4_data()
    qs = W.config4_queries(data)
    wins = W.config4_windows(data, qs)
    caps = W.config4_caps(len(qs))
    run = Runner(data, G.MODE_VLABEL)
    out = []
    for i, (q, (lo, hi), (_, cap)) in enumerate(zip(qs, wins, caps)):
        e = run.run(q, lo, hi, cap)
        e["query"] = i
        out.append(e)
        if i % 16 == 0:
            print(f"  config4 q{i}: {e['count']} capped={e['capped']} ({e['oracle_s']} s)", flush=True)
    write("config4", {"workload": "config 4: 256 queries over workloads.config4_windows / config4_caps",
                      "queries": out})


def config5():
    data = W.config5_data()
    qs = W.config5_queries(data)
    wins = W.config5_windows(data, qs)
    cap = W.CONFIG5_DEPTH_CAP
    run = Runner(data, G.MODE_VLABEL)
    out = []
    for i in range(0, len(qs), W.CONFIG5_FIXTURE_STRIDE):
        lo, hi = wins[i]
        e = run.run(qs[i], lo, hi, cap)
        e["query"] = i
        out.append(e)
        if len(out) % 64 == 0:
            print(f"  config5 q{i}: {e['count']} capped={e['capped']} ({e['oracle_s']} s)", flush=True)
    write("config5", {"workload": "config 5: every CONFIG5_FIXTURE_STRIDE-th query of the served stream",
                      "stride": W.CONFIG5_FIXTURE_STRIDE, "queries": out})


if __name__ == "__main__":
    names = sys.argv[1:] or ["config3", "config2", "config4", "config5"]
    for n in names:
        t = time.time()
        print(f"{n}:", flush=True)
        globals()[n]()
        print(f"{n} done in {time.time() - t:.0f} s", flush=True)
