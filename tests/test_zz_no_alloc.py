"""The allocation-free match (gfsm_match_noalloc) against gfsm_match, through
the C ABI: count and digest bit-exact on two calls in a row, the second with
a smaller input, on a small graph and at a full-size window whose levels far
exceed the 60 MiB level scratch (chunked expansion inside it); capped
results; and the argument checks."""
import pytest

import gfsm_amd as G
import gfsm_amd.workloads as W

pytestmark = pytest.mark.gpu


def same(ctx, data, q, **kw):
    ref = ctx.match(data, q, digest=True, allow_partial=True, **kw)
    rc, n, dg = ctx.match_noalloc(data, q, digest=True, **kw)
    assert rc == ref.status, (rc, ref.status)
    assert (n, dg) == (ref.count, ref.digest)
    assert ctx.match_noalloc(data, q, **kw)[1] == ref.count  # count only (lean last depth)
    return ref


def test_noalloc_equals_match_small_then_smaller():
    with G.Context(0) as ctx:
        data = G.Graph.er(3000, 8.0, nvlabels=3, nelabels=1, seed=9100, mode=G.MODE_VLABEL)
        ctx.upload(data)
        big = G.Graph.walk_query(data, 6, kind=G.WALK_DENSE, seed=9101)
        small = G.Graph.walk_query(data, 4, kind=G.WALK_SPARSE, seed=9102)
        assert same(ctx, data, big).count > 0
        assert same(ctx, data, small).count > 0
        same(ctx, data, big, max_depth_cand=5000)  # capped: the partial count and digest


def test_noalloc_full_size_window_then_smaller():
    """Power-law query 13 over its tail window: 1.6e8 embeddings, levels up to
    3.1e7 nodes (~1 GB) expanded in chunks inside the scratch; then query 18,
    a smaller window, through the same scratch."""
    with G.Context(0) as ctx:
        data = W.config2_data()
        qs = W.config2_queries(data)
        wins = W.config2_windows(data, qs)
        ctx.upload(data)
        for i in (13, 18):
            lo, hi = wins[i]
            assert same(ctx, data, qs[i], root_lo=lo, root_hi=hi).count > 0


def test_noalloc_argument_checks():
    with G.Context(0) as ctx:
        data = G.Graph.er(500, 6.0, nvlabels=2, nelabels=1, seed=9200, mode=G.MODE_VLABEL)
        q = G.Graph.walk_query(data, 4, seed=9201)
        with pytest.raises(G.GfsmError) as ei:  # not resident yet
            ctx.match_noalloc(data, q)
        assert ei.value.code == 2
        ctx.upload(data)
        with pytest.raises(G.GfsmError) as ei:  # rows need host and device allocations
            ctx.match_noalloc(data, q, rows=True)
        assert ei.value.code == 2
        with pytest.raises(G.GfsmError) as ei:  # beyond the level scratch
            ctx.match_noalloc(data, q, mem_budget=G.NOALLOC_LEVEL_BYTES + 1)
        assert ei.value.code == 2
        assert ctx.match_noalloc(data, q, mem_budget=G.NOALLOC_LEVEL_BYTES)[1] == ctx.match(data, q).count
