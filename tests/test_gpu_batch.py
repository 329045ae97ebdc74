"""GPU parity of the batched path (device scheduler + persistent step
kernel, hipGraph replays) and of the serving loop, against the CPU oracle and
the single-query path."""
import random
import time

import numpy as np
import pytest

import gfsm_amd as G
from helpers import oracle_solve

pytestmark = pytest.mark.gpu


def _queries(data, sizes, seed):
    qs = []
    for i, s in enumerate(sizes):
        kind = G.WALK_DENSE if i % 2 == 0 else G.WALK_SPARSE
        qs.append(G.Graph.walk_query(data, s, kind=kind, seed=seed + i))
    return qs


@pytest.fixture(scope="module")
def er_vlabel():
    return G.Graph.er(1500, 7.0, nvlabels=3, nelabels=2, seed=501, mode=G.MODE_VLABEL)


def test_batch_equals_oracle(ctx, er_vlabel):
    data = er_vlabel
    qs = _queries(data, [2, 3, 4, 5, 6, 7, 3, 4, 5, 6, 2, 8], seed=900)
    ms = ctx.match_batch(data, qs, digest=True)
    for q, m in zip(qs, ms):
        ref = oracle_solve(data, q, G.MODE_VLABEL, rows=False)
        assert m.status == 0
        assert (m.count, m.digest) == (ref.count, ref.digest)


def test_batch_ref_mode_and_windows(ctx):
    data = G.Graph.rmat(11, 8, nvlabels=1, zipf_s=1.0, nelabels=3, seed=77, mode=G.MODE_REF)
    qs = _queries(data, [3, 4, 5, 4, 3, 5, 6, 4], seed=70)
    windows = [(0, 0), (0, 500), (100, 1500), (0, 0), (1000, 2048), (7, 9), (0, 300), (2000, 2001)]
    ms = ctx.match_batch(data, qs, windows=windows, digest=True)
    for q, (lo, hi), m in zip(qs, windows, ms):
        ref = ctx.match(data, q, root_lo=lo, root_hi=hi, digest=True)
        assert (m.count, m.digest) == (ref.count, ref.digest), (lo, hi)


def test_batch_few_slots_continuous_admission(ctx, er_vlabel):
    """More queries than slots: finished slots are refilled between replays."""
    data = er_vlabel
    qs = _queries(data, [3, 5, 4, 6, 2, 7, 5, 3, 4, 6, 5, 4, 3, 6, 5, 7], seed=1300)
    ref = [ctx.match(data, q, digest=True) for q in qs]
    saved = ctx.tuning()
    try:
        for slots in (1, 3, 5):
            ctx.set_tuning(batch_slots=slots)
            ms = ctx.match_batch(data, qs, digest=True)
            assert [(m.count, m.digest) for m in ms] == [(r.count, r.digest) for r in ref], slots
    finally:
        ctx.set_tuning(batch_slots=saved["batch_slots"])


def test_batch_scheduler_workgroup_counts(ctx, er_vlabel):
    """The scheduler runs one wave per slot, four slots per workgroup, and the
    last workgroup to finish lays the tiles out: slot counts that leave the
    last workgroup partly empty (257: 65 workgroups, one slot in the last)
    and the maximum (1024: 256 workgroups), batched and served, give the
    single-query results."""
    data = er_vlabel
    qs = _queries(data, [3, 4, 5, 6, 2, 7] * 8, seed=1700)
    ref = [ctx.match(data, q, digest=True) for q in qs]
    want = [(r.count, r.digest, r.status) for r in ref]
    saved = ctx.tuning()
    try:
        for slots in (257, 1024):
            ctx.set_tuning(batch_slots=slots)
            ms = ctx.match_batch(data, qs, digest=True)
            assert [(m.count, m.digest, m.status) for m in ms] == want, slots
    finally:
        ctx.set_tuning(batch_slots=saved["batch_slots"])
    for inflight in (257, 1024):
        srv = G.Server(ctx, data, max_inflight=inflight, digest=True)
        try:
            got = [srv.wait(t) for t in [srv.submit(q) for q in qs]]
        finally:
            srv.stop()
        assert [(m.count, m.digest, m.status) for m in got] == want, inflight


@pytest.mark.parametrize("items,pool", [(1, 1024), (2, 4096), (4, 8192), (8, 14336), (8, 64)])
def test_batch_shapes(ctx, er_vlabel, items, pool):
    """Every tile shape and staging pool (a tiny pool forces unstaged checks)."""
    data = er_vlabel
    qs = _queries(data, [4, 5, 6, 3, 7], seed=2100)
    ref = [ctx.match(data, q, digest=True) for q in qs]
    saved = ctx.tuning()
    try:
        ctx.set_tuning(batch_items=items, batch_pool_words=pool)
        ms = ctx.match_batch(data, qs, digest=True)
        assert [(m.count, m.digest) for m in ms] == [(r.count, r.digest) for r in ref]
    finally:
        ctx.set_tuning(batch_items=saved["batch_items"], batch_pool_words=saved["batch_pool_words"])


def test_batch_device_chunking(ctx):
    """A tight arena: levels overflow their regions and the device scheduler
    halves and re-estimates chunks; counts are unchanged."""
    data = G.Graph.grid(60, 60, del_frac=0.05, diag_frac=0.02, seed=3, mode=G.MODE_REF)
    qs = _queries(data, [6, 7, 8, 6, 7], seed=4400)
    ref = [ctx.match(data, q, digest=True) for q in qs]
    ms = ctx.match_batch(data, qs, digest=True, mem_budget=48 << 20)
    assert [(m.count, m.digest) for m in ms] == [(r.count, r.digest) for r in ref]
    assert any(m.stats["overflows"] > 0 or m.stats["launches"] > 2 * m.stats["q"] for m in ms)


def test_batch_work_cap_and_mixed_flags(ctx, er_vlabel):
    data = er_vlabel
    qs = _queries(data, [5, 6], seed=5100)
    full = [ctx.match(data, q) for q in qs]
    ms = ctx.match_batch(data, qs, max_cand=1000)
    for m, f in zip(ms, full):
        if sum(f.stats["cand"]) > 1000:
            assert m.status == 5 and m.count <= f.count
    # rows requested: served by the single-query path inside the batch call
    ms = ctx.match_batch(data, qs, rows=True, sorted=True)
    for q, m in zip(qs, ms):
        ref = oracle_solve(data, q, G.MODE_VLABEL)
        assert np.array_equal(m.rows, ref.rows)


def test_batch_invalid_query_status(ctx, er_vlabel, tmp_path):
    p = tmp_path / "disc.g"  # two components: rejected like the reference (exit 1)
    p.write_text("t # 0\n4 2 1 1\nv 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 1 0 1\nt # -1")
    bad = G.Graph.load(p, mode=G.MODE_VLABEL)
    good = G.Graph.walk_query(er_vlabel, 4, seed=5)
    ms = ctx.match_batch(er_vlabel, [good, bad, good])
    assert ms[0].status == 0 and ms[2].status == 0 and ms[1].status == 1
    assert ms[0].count == ms[2].count == ctx.match(er_vlabel, good).count


def test_server_alone_vs_mixed(ctx, er_vlabel):
    """Each served result equals the same query run alone, while queries of
    different sizes (some over root windows) arrive during others' work."""
    data = er_vlabel
    rng = random.Random(5)
    qs = _queries(data, [rng.randint(2, 7) for _ in range(40)], seed=6000)
    wins = [(0, 0) if i % 3 else (rng.randint(0, 700), rng.randint(800, 1500)) for i in range(len(qs))]
    alone = [ctx.match(data, q, digest=True, root_lo=lo, root_hi=hi) for q, (lo, hi) in zip(qs, wins)]
    srv = G.Server(ctx, data, max_inflight=6, digest=True)
    try:
        tickets = []
        for q, (lo, hi) in zip(qs, wins):
            tickets.append(srv.submit(q, root_lo=lo, root_hi=hi))
            time.sleep(rng.expovariate(4000.0))
        got = [srv.wait(t) for t in tickets]
        times = [srv.times(t) for t in tickets]
    finally:
        srv.stop()
    assert [(m.count, m.digest) for m in got] == [(a.count, a.digest) for a in alone]
    for t0, t1, t2 in times:
        assert t0 <= t1 <= t2


def test_server_statuses(ctx, er_vlabel, tmp_path):
    p = tmp_path / "disc.g"
    p.write_text("t # 0\n4 2 1 1\nv 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 1 0 1\nt # -1")
    bad = G.Graph.load(p, mode=G.MODE_VLABEL)
    one = G.Graph.walk_query(er_vlabel, 1, seed=3)
    good = G.Graph.walk_query(er_vlabel, 4, seed=5)
    srv = G.Server(ctx, er_vlabel, max_inflight=2)
    try:
        tb, t1, tg = srv.submit(bad), srv.submit(one), srv.submit(good)
        mb, m1, mg = srv.wait(tb), srv.wait(t1), srv.wait(tg)
    finally:
        srv.stop()
    assert mb.status == 1
    assert m1.status == 0 and m1.count == ctx.match(er_vlabel, one).count
    assert mg.status == 0 and mg.count == ctx.match(er_vlabel, good).count


def test_autotune_keeps_results(ctx, er_vlabel):
    data = er_vlabel
    qs = _queries(data, [4, 5, 6], seed=7100)
    ref = [ctx.match(data, q, digest=True) for q in qs]
    saved = ctx.tuning()
    try:
        t = ctx.autotune(data, qs)
        assert t["batch_items"] in (2, 4, 8) and t["items"][0][1] in (2, 4, 8)
        ms = ctx.match_batch(data, qs, digest=True)
        assert [(m.count, m.digest) for m in ms] == [(r.count, r.digest) for r in ref]
        assert [(ctx.match(data, q, digest=True).count) for q in qs] == [r.count for r in ref]
    finally:
        ctx.set_tuning(**{k: v for k, v in saved.items()})


def test_deterministic_last_cap_same_on_every_path(ctx):
    """max_depth_cand cuts the enumeration at the same candidate of every depth
    on every path: single query with large or tiny level buffers (different
    chunking), batched with different slot counts, and served."""
    data = G.Graph.rmat(12, 8, nvlabels=3, zipf_s=1.0, nelabels=2, seed=91, mode=G.MODE_VLABEL)
    qs = _queries(data, [5, 6, 7, 6, 5, 7], seed=9100)
    full = [ctx.match(data, q, digest=True) for q in qs]
    for M in (1, 777, 50_000, 3_000_000):
        a = [ctx.match(data, q, digest=True, max_depth_cand=M, allow_partial=True) for q in qs]
        b = [ctx.match(data, q, digest=True, max_depth_cand=M, allow_partial=True, mem_budget=16 << 20) for q in qs]
        assert [(x.count, x.digest, x.status) for x in a] == [(x.count, x.digest, x.status) for x in b], M
        for x, f in zip(a, full):
            most = max([f.stats["roots"]] + f.stats["cand"][1:f.stats["q"]])
            if most <= M:
                assert (x.status, x.count, x.digest) == (0, f.count, f.digest)
            else:
                assert x.status == 5 and x.count <= f.count
        saved = ctx.tuning()
        try:
            for slots in (2, 6):
                ctx.set_tuning(batch_slots=slots)
                c = ctx.match_batch(data, qs, digest=True, max_depth_cand=M)
                assert [(x.count, x.digest, x.status) for x in c] == [(x.count, x.digest, x.status) for x in a], (M, slots)
        finally:
            ctx.set_tuning(batch_slots=saved["batch_slots"])
        srv = G.Server(ctx, data, max_inflight=3, digest=True)
        try:
            got = [srv.wait(t) for t in [srv.submit(q, max_depth_cand=M) for q in qs]]
        finally:
            srv.stop()
        assert [(x.count, x.digest, x.status) for x in got] == [(x.count, x.digest, x.status) for x in a], M


@pytest.mark.gpu
def test_server_ticket_lifetime_and_graph_pins(er_vlabel):
    """A result is handed out once (a second poll or wait fails with
    GFSM_ERR_ARG); a ticket's record is dropped once its result is taken and
    its timestamps read; a query graph freed while in flight, and the data
    graph freed while the server runs, stay valid until the server lets go;
    closing the context stops its servers after their submitted work."""
    c = G.Context(0)
    data = G.Graph.er(1500, 8.0, nvlabels=3, nelabels=1, seed=77, mode=G.MODE_VLABEL)
    qs = _queries(data, [4, 5, 3, 6], seed=6100)
    alone = [c.match(data, q, digest=True) for q in qs]
    srv = G.Server(c, data, max_inflight=2, digest=True)
    tk = [srv.submit(q) for q in qs]
    for q in qs:  # the server holds its own pins on queries in flight
        q.free()
    data.free()
    first = srv.wait(tk[0])
    assert (first.count, first.digest) == (alone[0].count, alone[0].digest)
    with pytest.raises(G.GfsmError) as ei:
        srv.wait(tk[0])
    assert ei.value.code == 2
    t0, t1, t2 = srv.times(tk[0])
    assert t0 <= t1 <= t2
    with pytest.raises(G.GfsmError) as ei:  # taken and timed: the record is gone
        srv.times(tk[0])
    assert ei.value.code == 2
    c.close()  # stops the server after the submitted work
    rest = [srv.wait(t) for t in tk[1:]]
    assert [(m.count, m.digest) for m in rest] == [(a.count, a.digest) for a in alone[1:]]
    with pytest.raises(G.GfsmError) as ei:
        srv.poll(tk[1])
    assert ei.value.code == 2
    with pytest.raises(G.GfsmError) as ei:
        srv.submit(G.Graph.walk_query(er_vlabel, 3, seed=1))
    assert ei.value.code == 2
    srv.stop()


@pytest.mark.gpu
def test_server_admission_waits_on_projected_memory(ctx):
    """A tight memory budget with many free slots: the server holds queries
    back because their projected levels do not fit (not for want of a slot),
    admits them as earlier queries release memory, and every result equals
    the same query run alone."""
    data = G.Graph.er(3000, 8.0, nvlabels=2, nelabels=1, seed=7700, mode=G.MODE_VLABEL)
    qs = _queries(data, [5, 6, 5, 6, 7, 5, 6, 7] * 5, seed=7701)
    alone = [ctx.match(data, q, digest=True, allow_partial=True, max_depth_cand=1 << 22) for q in qs]
    budget = 256 << 20  # regions of up to budget / 16 each: at most ~16 queries fit at once
    srv = G.Server(ctx, data, mem_budget=budget, max_inflight=128, digest=True)
    try:
        tk = [srv.submit(q, max_depth_cand=1 << 22) for q in qs]
        got = [srv.wait(t) for t in tk]
        st = srv.stats()
    finally:
        srv.stop()
    assert [(m.count, m.digest, m.status) for m in got] == [(a.count, a.digest, a.status) for a in alone]
    assert st["submitted"] == st["completed"] == len(qs)
    assert st["admitted"] == len(qs)
    assert st["waited_memory"] > 0 and st["waited_slot"] == 0
