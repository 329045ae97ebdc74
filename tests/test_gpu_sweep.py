"""Seeded sweep over graph families, relation modes, label counts, query
sizes and query shapes: every device path against the CPU oracle.

Per case: count only (the lean last-depth kernel where the last depth has
no checks), count and digest under each tile size (1, 2, 4 and 8 candidates
per thread), sorted rows under a tight level budget (chunked expansion), the
allocation-free entry point, and the batched engine (256 slots and 3 slots)
and a server (5 in flight) over all queries of the same data graph."""
import numpy as np
import pytest

import gfsm_amd as G
from helpers import assert_same, oracle_solve

pytestmark = pytest.mark.gpu

TILES = (256, 512, 1024, 2048)


def cases():
    """48 data graphs with 3 queries each.  R-MAT graphs (hubs) run in the
    labelled mode with at least 5 vertex labels and queries of at most 5
    vertices, so that counts stay within what the oracle enumerates in
    seconds; ER graphs run in both modes with queries of up to 6 vertices."""
    out = []
    for seed in range(48):
        rmat = seed % 4 == 1
        mode = G.MODE_VLABEL if (seed % 2 or rmat) else G.MODE_REF
        nvl = 5 + seed % 3 if rmat else 1 + seed % 4
        nel = 1 + (seed // 4) % 3
        if rmat:
            data = ("rmat", dict(scale=11, edge_factor=6, nvlabels=nvl, nelabels=nel, seed=700 + seed, mode=mode))
        else:
            data = ("er", dict(n=1500 + 100 * (seed % 16), avg_deg=4.0 + seed % 3, nvlabels=nvl, nelabels=nel,
                               seed=700 + seed, mode=mode))
        qs = [(3 + (seed + k) % (3 if rmat else 4), G.WALK_DENSE if (seed + k) % 2 else G.WALK_SPARSE,
               7000 + 10 * seed + k) for k in range(3)]
        out.append((seed, data, qs))
    return out


def build(kind, kw):
    if kind == "rmat":
        return G.Graph.rmat(kw["scale"], kw["edge_factor"], 0.57, 0.19, 0.19, nvlabels=kw["nvlabels"], zipf_s=1.0,
                            nelabels=kw["nelabels"], seed=kw["seed"], mode=kw["mode"])
    return G.Graph.er(kw["n"], kw["avg_deg"], nvlabels=kw["nvlabels"], nelabels=kw["nelabels"], seed=kw["seed"],
                      mode=kw["mode"])


@pytest.mark.parametrize("seed,data_spec,qspecs", cases(), ids=[f"{c[1][0]}{c[0]}" for c in cases()])
def test_sweep_every_path_vs_oracle(ctx, seed, data_spec, qspecs):
    kind, kw = data_spec
    data = build(kind, kw)
    mode = kw["mode"]
    ctx.upload(data)
    queries, refs = [], []
    for size, shape, qseed in qspecs:
        q = G.Graph.walk_query(data, size, kind=shape, seed=qseed)
        queries.append(q)
        refs.append(oracle_solve(data, q, mode, rows=False))
    nonzero = 0
    for i, (q, ref) in enumerate(zip(queries, refs)):
        nonzero += ref.count > 0
        assert ctx.match(data, q).count == ref.count, (seed, i)
        tile = TILES[(seed + i) % len(TILES)]
        m = ctx.match(data, q, digest=True, tile=tile)
        assert (m.count, m.digest) == (ref.count, ref.digest), (seed, i, tile)
        if ref.count <= 300_000:
            r = ctx.match(data, q, rows=True, sorted=True, digest=True, mem_budget=2 << 20)
            assert_same(r, oracle_solve(data, q, mode, rows=True))
        rc, n, dg = ctx.match_noalloc(data, q, digest=True)
        assert (rc, n, dg) == (0, ref.count, ref.digest), (seed, i)
    want = [(ref.count, ref.digest) for ref in refs]
    batched = ctx.match_batch(data, queries, digest=True)
    assert [(b.count, b.digest) for b in batched] == want, seed
    saved = ctx.tuning()
    try:
        ctx.set_tuning(batch_slots=3)
        assert [(b.count, b.digest) for b in ctx.match_batch(data, queries, digest=True)] == want, seed
    finally:
        ctx.set_tuning(batch_slots=saved["batch_slots"])
    srv = G.Server(ctx, data, max_inflight=5, digest=True)
    try:
        served = [srv.wait(t) for t in [srv.submit(q) for q in queries]]
    finally:
        srv.stop()
    assert [(m.count, m.digest) for m in served] == want, seed
    assert nonzero >= 1, seed
