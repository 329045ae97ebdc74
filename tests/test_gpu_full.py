"""GPU parity at the benchmark workloads' full sizes, through properties that
hold at any size: the oracle on the windows it finishes in seconds, counts and
digests that add over a split of the root window, chunked == unchunked, the
lean counting kernel == the digest kernel, closed forms on the full
1400 x 1400 grid, batched == single query on the scale-22 graph, and served
== alone on the resident scale-23 graph."""
import os

import numpy as np
import pytest

import gfsm_amd as G
import gfsm_amd.workloads as W
from helpers import assert_same, oracle_solve

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def c2(ctx):
    data = W.config2_data()
    qs = W.config2_queries(data)
    return data, qs, W.config2_windows(data, qs)


def test_config2_windows_vs_oracle(ctx, c2):
    """R-MAT scale 18: the windowed queries with at most a few million
    embeddings, count and digest against the oracle; sorted rows for the
    smallest."""
    data, qs, wins = c2
    for i in (0, 18, 19):
        lo, hi = wins[i]
        m = ctx.match(data, qs[i], root_lo=lo, root_hi=hi, digest=True)
        o = oracle_solve(data, qs[i], G.MODE_VLABEL, rows=False, root_lo=lo, root_hi=hi)
        assert m.count > 0 and (m.count, m.digest) == (o.count, o.digest), i
    for i in (18, 19):
        lo, hi = wins[i]
        m = ctx.match(data, qs[i], root_lo=lo, root_hi=hi, rows=True, sorted=True, digest=True)
        assert_same(m, oracle_solve(data, qs[i], G.MODE_VLABEL, rows=True, root_lo=lo, root_hi=hi))


def test_config2_window_splits(ctx, c2):
    """All 20 windowed queries (6.8e9 embeddings): the two halves of each root
    window add up to the whole in count and digest, and the lean counting
    last depth gives the digest run's count."""
    data, qs, wins = c2
    total = 0
    for i, (q, (lo, hi)) in enumerate(zip(qs, wins)):
        whole = ctx.match(data, q, root_lo=lo, root_hi=hi, digest=True)
        mid = (lo + hi) // 2
        a = ctx.match(data, q, root_lo=lo, root_hi=mid, digest=True)
        b = ctx.match(data, q, root_lo=mid, root_hi=hi, digest=True)
        assert a.count + b.count == whole.count, i
        assert ((a.digest[0] + b.digest[0]) & M64, (a.digest[1] + b.digest[1]) & M64) == whole.digest, i
        assert ctx.match(data, q, root_lo=lo, root_hi=hi).count == whole.count, i
        total += whole.count
    assert total == 6771680081  # the headline workload's embeddings per pass


def test_config2_chunked_full_size(ctx, c2):
    """Query 13 of the workload (1.6e8 embeddings, levels up to 3.1e7 nodes)
    under a 256 MiB level budget: many chunks, same count and digest."""
    data, qs, wins = c2
    lo, hi = wins[13]
    full = ctx.match(data, qs[13], root_lo=lo, root_hi=hi, digest=True)
    small = ctx.match(data, qs[13], root_lo=lo, root_hi=hi, digest=True, mem_budget=256 << 20)
    assert small.stats["chunks"] > 0
    assert (small.count, small.digest) == (full.count, full.digest)


def test_config3_full_grid_closed_forms(ctx):
    """The 1400 x 1400 grid without deletions: ordered counts of paths,
    stars and squares in closed form."""
    R = Cc = 1400
    g = G.Graph.grid(R, Cc)
    qdir = os.path.join(os.path.dirname(__file__), "golden", "queries")
    r, c = np.divmod(np.arange(R * Cc), Cc)
    deg = ((r > 0).astype(np.int64) + (r < R - 1) + (c > 0) + (c < Cc - 1))
    expect = {
        "road2.g": 2 * (R * (Cc - 1) + Cc * (R - 1)),
        "road3.g": int((deg * (deg - 1)).sum()),
        "triangle.g": 0,
        "square.g": 8 * (R - 1) * (Cc - 1),
        "roadY.g": int((deg * (deg - 1) * (deg - 2)).sum()),
        "plus.g": 24 * (R - 2) * (Cc - 2),
    }
    for name, want in expect.items():
        q = G.Graph.load(os.path.join(qdir, name))
        assert ctx.match(g, q).count == want, name
        assert ctx.match(g, q, digest=True).count == want, name


def test_config3_window_splits(ctx):
    """The config-3 grid (5% deleted, 2% diagonals): C4 and the 3x3 patch,
    counts and digests add over four root windows."""
    data = W.config3_data()
    qs = W.config3_queries()
    V = 1400 * 1400
    cuts = [0, V // 4, V // 2, 3 * V // 4, V]
    for name in ("C4", "grid3x3"):
        whole = ctx.match(data, qs[name], digest=True)
        parts = [ctx.match(data, qs[name], root_lo=cuts[k], root_hi=cuts[k + 1], digest=True) for k in range(4)]
        assert sum(p.count for p in parts) == whole.count > 0, name
        assert (sum(p.digest[0] for p in parts) & M64, sum(p.digest[1] for p in parts) & M64) == whole.digest, name


def test_config4_batched_equals_single(ctx):
    """R-MAT scale 22, 4 edge labels: the first 32 windowed queries (with
    their deterministic caps) in one batched call equal each query alone."""
    data = W.config4_data()
    qs = W.config4_queries(data)[:32]
    wins = W.config4_windows(data, qs, ks=W.CONFIG4_ROOT_K[:32])
    caps = W.config4_caps()[:32]
    ms = ctx.match_batch(data, qs, windows=wins, caps=[c for c, _ in caps], depth_caps=[d for _, d in caps],
                         digest=True)
    for i, (q, (lo, hi), (cap, dcap), m) in enumerate(zip(qs, wins, caps, ms)):
        one = ctx.match(data, q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=dcap, allow_partial=True,
                        digest=True)
        assert (m.count, m.digest, m.status) == (one.count, one.digest, one.status), i


def test_config5_served_equals_alone(ctx):
    """R-MAT scale 23, edge factor 16, resident: the first 512 queries of the
    served stream (4-24 vertices, deterministic caps) through the server with
    a 200 GB admission budget equal each query run alone."""
    data = W.config5_data()
    qs = W.config5_queries(data, 512)
    wins = W.config5_windows(data, qs)
    cap, last = W.CONFIG5_CAP, W.CONFIG5_DEPTH_CAP
    srv = G.Server(ctx, data, mem_budget=200 * 10**9, max_inflight=256, digest=True)
    try:
        tk = [srv.submit(q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last) for q, (lo, hi) in zip(qs, wins)]
        got = [srv.wait(x) for x in tk]
    finally:
        srv.stop()
    capped = 0
    for i, (q, (lo, hi), m) in enumerate(zip(qs, wins, got)):
        one = ctx.match(data, q, root_lo=lo, root_hi=hi, max_cand=cap, max_depth_cand=last, allow_partial=True,
                        digest=True)
        assert (m.count, m.digest, m.status) == (one.count, one.digest, one.status), i
        capped += one.status != 0
    assert 0 < capped < len(qs)  # both capped and finished queries are compared
