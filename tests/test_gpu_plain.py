"""GPU parity of the plain (single-query) path against the CPU oracle."""
import numpy as np
import pytest

import gfsm_amd as G
import oracle as O
from helpers import assert_same, oracle_solve, to_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 63, 64, 255, 256, 257, 1000, 65537, 1 << 20])
def test_chained_scan_random(ctx, n):
    rng = np.random.default_rng(n)
    vals = rng.integers(0, 5, size=n).astype(np.uint32)
    vals[rng.random(n) < 0.3] = 0
    pre, idx, (nk, tot) = ctx.selftest_scan(vals)
    ref = np.concatenate([[0], np.cumsum(vals.astype(np.uint64))[:-1]]).astype(np.uint64)
    assert np.array_equal(pre, ref)
    assert np.array_equal(idx, np.nonzero(vals)[0].astype(np.uint32))
    assert nk == int((vals != 0).sum()) and tot == int(vals.astype(np.uint64).sum())


@pytest.mark.parametrize("kind", ["zeros", "ones", "big"])
def test_chained_scan_adversarial(ctx, kind):
    n = 300_000
    if kind == "zeros":
        vals = np.zeros(n, np.uint32)
    elif kind == "ones":
        vals = np.ones(n, np.uint32)
    else:  # per-wave sums stay below 2^32, the total exceeds it
        vals = np.full(n, (1 << 25) + 12345, np.uint32)
    pre, idx, (nk, tot) = ctx.selftest_scan(vals)
    ref = np.concatenate([[0], np.cumsum(vals.astype(np.uint64))[:-1]]).astype(np.uint64)
    assert np.array_equal(pre, ref)
    assert nk == int((vals != 0).sum()) and tot == int(vals.astype(np.uint64).sum())
    assert tot < (1 << 32) or kind == "big"


def test_readme_graph(ctx, tmp_path):
    p = tmp_path / "readme.g"
    p.write_text("t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -1")
    g = G.Graph.load(p)
    m = ctx.match(g, g, rows=True, sorted=True)
    assert m.count == 2
    assert m.rows.tolist() == [[0, 1], [1, 0]]


@pytest.mark.parametrize("mode", [G.MODE_REF, G.MODE_VLABEL])
@pytest.mark.parametrize("qsize,kind,seed", [(3, G.WALK_DENSE, 1), (4, G.WALK_DENSE, 2), (5, G.WALK_DENSE, 3),
                                             (5, G.WALK_SPARSE, 4), (6, G.WALK_SPARSE, 5), (7, G.WALK_DENSE, 6),
                                             (2, G.WALK_DENSE, 7)])
def test_er_vs_oracle(ctx, mode, qsize, kind, seed):
    data = G.Graph.er(600, 6.0, nvlabels=3, nelabels=2, seed=100 + seed, mode=mode)
    query = G.Graph.walk_query(data, qsize, kind=kind, seed=seed)
    m = ctx.match(data, query, rows=True, sorted=True, digest=True)
    ref = oracle_solve(data, query, mode)
    assert_same(m, ref)


def test_config1_sorted_rows(ctx):
    data = G.Graph.er(2000, 8.0, nvlabels=4, nelabels=1, seed=1000, mode=G.MODE_VLABEL)
    query = G.Graph.walk_query(data, 5, kind=G.WALK_DENSE, seed=1001)
    m = ctx.match(data, query, rows=True, sorted=True, digest=True)
    ref = oracle_solve(data, query, G.MODE_VLABEL)
    assert_same(m, ref)
    # soundness restated from the reference's checker
    assert O.check_sound(to_oracle(query), to_oracle(data), m.rows) == m.count


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_rmat_count_digest(ctx, seed):
    data = G.Graph.rmat(12, 8, nvlabels=4, zipf_s=1.0, nelabels=2, seed=seed, mode=G.MODE_VLABEL)
    for qs, kind in [(4, G.WALK_DENSE), (5, G.WALK_SPARSE), (6, G.WALK_DENSE)]:
        query = G.Graph.walk_query(data, qs, kind=kind, seed=seed * 10 + qs)
        m = ctx.match(data, query, digest=True)
        ref = oracle_solve(data, query, G.MODE_VLABEL, rows=False)
        assert m.count == ref.count and m.digest == ref.digest, (qs, kind, m.count, ref.count)


def test_grid_closed_forms(ctx):
    R, Cc = 30, 40
    g = G.Graph.grid(R, Cc)
    import os
    qdir = os.path.join(os.path.dirname(__file__), "golden", "queries")
    def q(name):
        return G.Graph.load(os.path.join(qdir, name))
    deg = np.array([(r > 0) + (r < R - 1) + (c > 0) + (c < Cc - 1) for r in range(R) for c in range(Cc)])
    expect = {
        "road2.g": 2 * (R * (Cc - 1) + Cc * (R - 1)),
        "road3.g": int((deg * (deg - 1)).sum()),
        "triangle.g": 0,
        "square.g": 8 * (R - 1) * (Cc - 1),
        "roadY.g": int((deg * (deg - 1) * (deg - 2)).sum()),
        "plus.g": 24 * (R - 2) * (Cc - 2),
    }
    for name, want in expect.items():
        m = ctx.match(g, q(name), digest=True)
        assert m.count == want, (name, m.count, want)


def test_chunked_equals_unchunked(ctx):
    data = G.Graph.grid(60, 60, del_frac=0.05, diag_frac=0.02, seed=3)
    query = G.Graph.walk_query(data, 6, kind=G.WALK_SPARSE, seed=9)
    full = ctx.match(data, query, digest=True)
    small = ctx.match(data, query, digest=True, mem_budget=1 << 20)
    assert full.count == small.count and full.digest == small.digest
    assert small.stats["chunks"] > 0
    rows_full = ctx.match(data, query, rows=True, sorted=True)
    rows_small = ctx.match(data, query, rows=True, sorted=True, mem_budget=1 << 20)
    assert np.array_equal(rows_full.rows, rows_small.rows)


def test_chunked_steps_with_tile_index(ctx):
    """Chunked expansion whose chunks start past candidate 0 and still span
    hundreds of tiles, so the tile index runs on offset ranges: counts and
    digests equal the unchunked match and the oracle (34.7M embeddings)."""
    data = G.Graph.er(20000, 12.0, nvlabels=1, nelabels=1, seed=81, mode=G.MODE_REF)
    query = G.Graph.walk_query(data, 4, kind=G.WALK_SPARSE, seed=8104)
    ref = oracle_solve(data, query, G.MODE_REF, rows=False)
    full = ctx.match(data, query, digest=True)
    assert (full.count, full.digest) == (ref.count, ref.digest)
    assert full.stats["chunks"] == 0
    for budget in (32 << 20, 48 << 20):
        m = ctx.match(data, query, digest=True, mem_budget=budget)
        assert (m.count, m.digest) == (ref.count, ref.digest), budget
        assert m.stats["chunks"] > 0
        assert ctx.match(data, query, mem_budget=budget).count == ref.count  # counting last depth


def test_root_range_partition(ctx):
    data = G.Graph.er(1500, 7.0, nvlabels=2, seed=77, mode=G.MODE_VLABEL)
    query = G.Graph.walk_query(data, 5, kind=G.WALK_SPARSE, seed=78)
    full = ctx.match(data, query, digest=True)
    parts = [ctx.match(data, query, digest=True, root_lo=lo, root_hi=lo + 500) for lo in (0, 500, 1000)]
    assert sum(p.count for p in parts) == full.count
    d0 = sum(p.digest[0] for p in parts) % (1 << 64)
    d1 = sum(p.digest[1] for p in parts) % (1 << 64)
    assert (d0, d1) == full.digest


@pytest.mark.parametrize("mode", [G.MODE_REF, G.MODE_VLABEL])
def test_cost_order_same_embeddings(ctx, mode):
    """GFSM_ORDER_COST plans another order; the embeddings are the same:
    count and digest against the reference order and the oracle, sorted rows,
    the batched path, and root windows rejected."""
    data = G.Graph.er(700, 6.0, nvlabels=3, nelabels=2, seed=41, mode=mode)
    qs = [G.Graph.walk_query(data, s, kind=s % 2, seed=4100 + s) for s in (3, 4, 5, 6, 7)]
    for q in qs:
        a = ctx.match(data, q, digest=True)
        b = ctx.match(data, q, digest=True, order_cost=True)
        o = oracle_solve(data, q, mode, rows=False)
        assert (b.count, b.digest) == (a.count, a.digest) == (o.count, o.digest)
    q = qs[2]
    ra = ctx.match(data, q, rows=True, sorted=True)
    rb = ctx.match(data, q, rows=True, sorted=True, order_cost=True)
    assert np.array_equal(ra.rows, rb.rows)
    ms = ctx.match_batch(data, qs, digest=True, order_cost=True)
    assert [(m.count, m.digest) for m in ms] == [(ctx.match(data, q, digest=True).count,
                                                   ctx.match(data, q, digest=True).digest) for q in qs]
    with pytest.raises(G.GfsmError) as ei:
        ctx.match(data, q, order_cost=True, root_lo=0, root_hi=10)
    assert ei.value.code == 2


def test_lean_last_depth_counts(ctx):
    """Count-only matches whose last depth has no checks run the lean
    last-depth kernel; their counts equal the digest runs (general kernel)
    and the oracle, for every tile size and with tiny level buffers."""
    cases = [(G.Graph.er(900, 7.0, nvlabels=2, nelabels=1, seed=61, mode=G.MODE_VLABEL), G.MODE_VLABEL),
             (G.Graph.rmat(11, 8, nvlabels=3, zipf_s=1.0, nelabels=1, seed=62, mode=G.MODE_VLABEL), G.MODE_VLABEL),
             (G.Graph.grid(50, 50, del_frac=0.05, diag_frac=0.02, seed=63), G.MODE_REF)]
    for data, mode in cases:
        for s in (2, 3, 4, 5, 6):
            q = G.Graph.walk_query(data, s, kind=G.WALK_SPARSE, seed=6300 + s)
            ref = oracle_solve(data, q, mode, rows=False).count
            for tile in (0, 256, 512, 1024, 2048):
                assert ctx.match(data, q, tile=tile).count == ref, (s, tile)
            assert ctx.match(data, q, digest=True).count == ref
            assert ctx.match(data, q, mem_budget=2 << 20).count == ref


def test_directed_self_loops_vs_oracle(ctx):
    """Directed data and queries in REF mode: one-way pairs, two-way pairs
    whose directions carry different labels, and self-loops (every vertex has
    an out-entry).  Queries the reference compiler accepts give the oracle's
    count, digest and sorted rows; the ones it rejects fail with status 1."""
    from test_oracle import directed_entries

    rng = np.random.default_rng(97)
    n = 60
    ent = directed_entries(rng, n, p_edge=0.12, p_loop=0.1)
    s, d, l = (np.array(x, np.uint32) for x in zip(*ent))
    data = G.Graph.from_edges(n, s, d, l, vlabel=[1] * n)
    und = {}
    for a, b, _ in ent:
        if a != b:
            und.setdefault(a, set()).add(b)
            und.setdefault(b, set()).add(a)
    checked = nonzero = rejected = 0
    for trial in range(300):
        q = int(rng.integers(3, 6))
        verts = [int(rng.integers(0, n))]
        while len(verts) < q:  # a connected vertex set grown at random
            nb = sorted(set().union(*(und.get(v, set()) for v in verts)) - set(verts))
            if not nb:
                break
            verts.append(int(rng.choice(nb)))
        if len(verts) < q:
            continue
        idx = {v: i for i, v in enumerate(verts)}
        qent = [(idx[a], idx[b], lab) for a, b, lab in ent if a in idx and b in idx]
        if {a for a, _, _ in qent} != set(range(q)):
            continue  # a query vertex without an out-entry has no row
        qs, qd, ql = (np.array(x, np.uint32) for x in zip(*qent))
        query = G.Graph.from_edges(q, qs, qd, ql, vlabel=[1] * q)
        try:
            ref = oracle_solve(data, query, G.MODE_REF, rows=True)
        except O.OracleError as e:
            assert e.code == 1
            with pytest.raises(G.GfsmError) as ei:
                ctx.match(data, query)
            assert ei.value.code == 1
            rejected += 1
            continue
        m = ctx.match(data, query, rows=True, sorted=True, digest=True)
        assert_same(m, ref)
        checked += 1
        nonzero += ref.count > 0
        if checked >= 40:
            break
    assert checked >= 20 and nonzero >= 10 and rejected > 0, (checked, nonzero, rejected)


def test_rowless_vertices_follow_the_rule(ctx):
    """Directed data where some targets have no out-entries (no row; the
    reference reads past its segment array there): the product keeps them
    distinct with empty rows, and its sorted rows equal the placement rule
    taken literally over all injective maps."""
    from test_oracle import brute_force_rule, directed_entries, edges_obj

    rng = np.random.default_rng(131)
    checked = with_rowless = 0
    for trial in range(1500):
        n = int(rng.integers(5, 8))
        ent = directed_entries(rng, n, p_edge=0.5, p_loop=0.25, all_rows=False, p_both=0.2)
        rowless = set(range(n)) - {a for a, _, _ in ent}
        if not ent:
            continue
        q = int(rng.integers(2, 5))
        verts = [int(v) for v in rng.choice(n, q, replace=False)]
        idx = {v: i for i, v in enumerate(verts)}
        qent = [(idx[a], idx[b], lab) for a, b, lab in ent if a in idx and b in idx and rng.random() < 0.9]
        if {a for a, _, _ in qent} != set(range(q)):
            continue  # every query vertex needs an out-entry
        try:
            order = [int(c) for c in O.Plan(O.CCSR(edges_obj(q, qent), O.MODE_REF)).fields()["col"]]
        except O.OracleError as e:
            assert e.code == 1
            continue
        s, d, lab = (np.array(x, np.uint32) for x in zip(*ent))
        qs, qd, ql = (np.array(x, np.uint32) for x in zip(*qent))
        data = G.Graph.from_edges(n, s, d, lab, vlabel=[1] * n)
        query = G.Graph.from_edges(q, qs, qd, ql, vlabel=[1] * q)
        m = ctx.match(data, query, rows=True, sorted=True)
        got = m.rows.tolist() if m.count else []
        assert got == brute_force_rule(n, ent, q, qent, order), (trial, ent, qent, order)
        checked += 1
        with_rowless += bool(rowless & {b for _, b, _ in ent})
    assert checked >= 50 and with_rowless >= 20, (checked, with_rowless)


def test_traffic_calibration_kernels(ctx):
    """The calibration kernels report their footprint by construction (the
    profiler's byte counters are scaled by them) and reject bad shapes."""
    assert ctx.selftest_traffic(0, 1 << 20, 2) == (4 << 20, 4 << 20)
    assert ctx.selftest_traffic(1, 1 << 16, 2) == (128 << 16, 4 << 16)
    assert ctx.selftest_traffic(2, 1 << 16, 2) == (128 << 16, 8 << 16)
    with pytest.raises(G.GfsmError):
        ctx.selftest_traffic(1, 3 << 16, 1)  # the gather needs a power of two
    with pytest.raises(G.GfsmError):
        ctx.selftest_traffic(3, 1 << 10, 1)


def test_one_vertex_queries_on_the_device(ctx):
    """One-vertex queries run on the device kernel: the vertices with an
    out-entry (the reference's root level lists rows only) carrying every
    self-loop relation, in id order; count, digest and rows against the
    oracle, over the whole graph and over a root window, in both modes, and
    through the allocation-free entry point."""
    n = 300
    rng = np.random.default_rng(33)
    src, dst, lab = [], [], []
    for a in range(n):
        if rng.random() < 0.3:  # self loops with labels 1..2
            src.append(a); dst.append(a); lab.append(int(rng.integers(1, 3)))
    for a in range(n):
        for b in range(a + 1, n):
            if rng.random() < 0.02:
                l = int(rng.integers(1, 3))
                src += [a, b]; dst += [b, a]; lab += [l, l]
    vl = [int(x) for x in rng.integers(1, 4, n)]
    for mode in (G.MODE_REF, G.MODE_VLABEL):
        data = G.Graph.from_edges(n, src, dst, lab, vlabel=vl, mode=mode)
        ctx.upload(data)
        od = O.Edges.from_arrays(n, src, dst, lab, vl)
        queries = [([], [], []), ([0], [0], [1]), ([0], [0], [2])]  # plain vertex, self loop of label 1 or 2
        for qs, qd, ql in queries:
            q = G.Graph.from_edges(1, qs, qd, ql, vlabel=[int(vl[7])], mode=mode)
            oq = O.Edges.from_arrays(1, qs, qd, ql, [int(vl[7])])
            for lo, hi in ((0, 0), (40, 170)):
                if qs:
                    ref = O.solve(od, oq, mode=mode, rows=True, root_lo=lo, root_hi=hi if hi else 0xFFFFFFFF)
                else:
                    # no edge at all: the oracle's compiler rejects a query without a row
                    # (as the reference's does); the rule lists every vertex with an out-entry
                    ids = sorted({a for a in src if lo <= a < (hi or n)})
                    ref = type("R", (), {})()
                    ref.count, ref.width = len(ids), 1
                    ref.rows = np.array(ids, np.uint32).reshape(-1, 1)
                    hs = [O.row_hash([x]) for x in ids]
                    ref.digest = (sum(h[0] for h in hs) % (1 << 64), sum(h[1] for h in hs) % (1 << 64))
                m = ctx.match(data, q, rows=True, sorted=True, digest=True, root_lo=lo, root_hi=hi)
                assert_same(m, ref)
                c = ctx.match(data, q, digest=True, root_lo=lo, root_hi=hi)
                assert (c.count, c.digest) == (ref.count, ref.digest)
                assert ctx.match_noalloc(data, q, digest=True, root_lo=lo, root_hi=hi) == (0, ref.count, ref.digest)
