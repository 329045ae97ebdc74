"""GPU parity on the benchmark workloads against the CPU oracle's results.

The oracle's counts and digests for every timed window of the power-law
sample, the four road-like queries over every root, the 256 batched queries
and a sample of the served stream are committed under tests/golden/
(fixtures_config*.json, written by tests/golden/make_workload_fixtures.py).
Each test runs the HIP path -- single query, batched, served -- on the same
seeded inputs and compares the count, both digest sums and whether the
deterministic per-depth cap stopped the query."""
import json
import os

import numpy as np
import pytest

import gfsm_amd as G
import gfsm_amd.workloads as W
from helpers import assert_same, to_oracle

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CAPPED = 5  # GFSM_ERR_CAP: the per-depth cap stopped the query


def fixture(name):
    path = os.path.join(GOLD, f"fixtures_{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated yet (tests/golden/make_workload_fixtures.py {name})")
    with open(path) as f:
        return json.load(f)


def key(e):
    return e["count"], (int(e["digest"][0], 16), int(e["digest"][1], 16)), bool(e["capped"])


def got(m):
    assert m.status in (0, CAPPED), m.status
    return m.count, m.digest, m.status == CAPPED


# ------------------------------------------------------------------ config 2

@pytest.fixture(scope="module")
def c2():
    data = W.config2_data()
    return data, W.config2_queries(data)


def test_config2_sample_vs_oracle(ctx, c2):
    """All 60 timed windows (hub end and middle under the per-depth cap, tail
    windows exact): count, digest and capped flag equal the oracle's, and the
    lean counting kernel (no digest) gives the same counts."""
    data, qs = c2
    fx = fixture("config2")["windows"]
    sample = W.config2_sample(data, qs)
    assert len(fx) == len(sample) == 60
    for e, (i, stratum, lo, hi, cap) in zip(fx, sample):
        assert (e["query"], e["stratum"], e["root_lo"], e["root_hi"], e["max_depth_cand"]) == (i, stratum, lo, hi, cap)
        m = ctx.match(data, qs[i], root_lo=lo, root_hi=hi, max_depth_cand=cap, allow_partial=True, digest=True)
        assert got(m) == key(e), (i, stratum)
        lean = ctx.match(data, qs[i], root_lo=lo, root_hi=hi, max_depth_cand=cap, allow_partial=True)
        assert lean.count == e["count"], (i, stratum)
    assert sum(e["count"] for e in fx if e["stratum"] == "tail") == 6771680081


def test_config2_capped_windows_do_not_depend_on_chunking(ctx, c2):
    """Hub-end windows under a 64 MiB level budget (many more, smaller
    chunks) give the oracle's capped count and digest."""
    data, qs = c2
    fx = [e for e in fixture("config2")["windows"] if e["stratum"] == "hub"][:6]
    for e in fx:
        m = ctx.match(data, qs[e["query"]], root_lo=e["root_lo"], root_hi=e["root_hi"],
                      max_depth_cand=e["max_depth_cand"], allow_partial=True, digest=True, mem_budget=64 << 20)
        assert got(m) == key(e), e["query"]


# ------------------------------------------------------------------ config 3

@pytest.fixture(scope="module")
def c3():
    return W.config3_data(), W.config3_queries()


def test_config3_vs_oracle(ctx, c3):
    """C4, C6, P7 (1.3e9 embeddings) and the 3x3 patch over every root of the
    damaged 1400 x 1400 grid: count and digest equal the oracle's."""
    data, qs = c3
    fx = fixture("config3")["queries"]
    for name, q in qs.items():
        m = ctx.match(data, q, digest=True)
        assert got(m) == key(fx[name]), name
        assert ctx.match(data, q).count == fx[name]["count"], name


def test_config3_p7_chunked_vs_oracle(ctx, c3):
    """P7 under a 256 MiB level budget: its levels (up to ~1e9 nodes) are
    expanded in many depth-first chunks, which replaces the reference's
    MEMLIMIT -> UPPERMEMLIMIT rerun; the result is still the oracle's."""
    data, qs = c3
    e = fixture("config3")["queries"]["P7"]
    m = ctx.match(data, qs["P7"], digest=True, mem_budget=256 << 20)
    assert m.stats["chunks"] > 100
    assert got(m) == key(e)


def test_config3_c4_rows_vs_oracle(ctx, c3):
    """C4's sorted embedding rows over one root window of the damaged grid
    equal the oracle's rows."""
    data, qs = c3
    lo, hi = 700 * 1400, 700 * 1400 + 20000
    m = ctx.match(data, qs["C4"], root_lo=lo, root_hi=hi, rows=True, sorted=True, digest=True)
    o = O.solve(to_oracle(data), to_oracle(qs["C4"]), mode=G.MODE_REF, rows=True, root_lo=lo, root_hi=hi)
    assert o.count > 1000
    assert_same(m, o)


# ------------------------------------------------------------------ config 4

def test_config4_batched_and_single_vs_oracle(ctx):
    """All 256 queries of the batched workload (30 under the per-depth cap)
    in one batched call, and each alone: the oracle's count, digest and
    capped flag."""
    data = W.config4_data()
    qs = W.config4_queries(data)
    wins = W.config4_windows(data, qs)
    caps = W.config4_caps(len(qs))
    fx = fixture("config4")["queries"]
    assert len(fx) == 256
    ms = ctx.match_batch(data, qs, windows=wins, caps=[c for c, _ in caps], depth_caps=[d for _, d in caps],
                         digest=True)
    for i, (e, m) in enumerate(zip(fx, ms)):
        assert (e["root_lo"], e["root_hi"], e["max_depth_cand"]) == (wins[i][0], wins[i][1], caps[i][1])
        assert got(m) == key(e), i
    for i in range(0, 256, 3):
        lo, hi = wins[i]
        one = ctx.match(data, qs[i], root_lo=lo, root_hi=hi, max_cand=caps[i][0], max_depth_cand=caps[i][1],
                        allow_partial=True, digest=True)
        assert got(one) == key(fx[i]), i
    assert sum(e["capped"] for e in fx) == len(W.CONFIG4_CAPPED)


# ------------------------------------------------------------------ config 5

def test_config5_served_and_single_vs_oracle(ctx):
    """Every CONFIG5_FIXTURE_STRIDE-th query of the served stream through the
    server (200 GB admission budget) and alone: the oracle's count, digest
    and capped flag."""
    data = W.config5_data()
    fx = fixture("config5")
    idx = [e["query"] for e in fx["queries"]]
    assert fx["stride"] == W.CONFIG5_FIXTURE_STRIDE
    allq = W.config5_queries(data, max(idx) + 1)
    qs = [allq[i] for i in idx]
    wins = W.config5_windows(data, qs)
    cap, last = W.CONFIG5_CAP, W.CONFIG5_DEPTH_CAP
    srv = G.Server(ctx, data, mem_budget=200 * 10**9, max_inflight=256, digest=True)
    try:
        tk = [srv.submit(q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last) for q, (lo, hi) in zip(qs, wins)]
        served = [srv.wait(x) for x in tk]
    finally:
        srv.stop()
    for e, q, (lo, hi), m in zip(fx["queries"], qs, wins, served):
        assert (e["root_lo"], e["root_hi"]) == (lo, hi), e["query"]
        assert got(m) == key(e), e["query"]
        one = ctx.match(data, q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last, allow_partial=True,
                        digest=True)
        assert got(one) == key(e), e["query"]
    finished = sum(not e["capped"] for e in fx["queries"])
    assert finished >= 256 and finished > len(idx) // 2
