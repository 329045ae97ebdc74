"""Pinning the CPU oracle (CPU only).

Every oracle function is checked by something other than its own formula:
  * orc_load_tve / orc_ccsr_build / orc_ccsr_packed / orc_offset_to_id:
    the README graph's known CCSR words (KAT-1, README.md:9-16 with the layout
    rules of HnHelper.cuh:383-478), error codes of malformed files, and
    round-trip invariants (every entry points at a row header, rows sorted,
    hnWrite's id mapping = the inverse permutation);
  * orc_plan_build: the triangle's requirement table (KAT-2,
    build/query/_3_triangle.g with ccsr.cpp:98-204) and, for many queries,
    structural invariants (every query edge covered exactly once, firstQ =
    distance to the nearest placed neighbour);
  * orc_match: closed-form counts (grids, complete graphs, paths, cycles,
    complete bipartite graphs), networkx VF2 golden vectors
    (tests/golden/nx_fixtures.json) and a pure-Python brute force, plus the
    per-depth level sizes = counts of the order-prefix sub-queries;
  * orc_check_sound: a planted wrong row is found;
  * orc_row_hash: an independent Python SplitMix64 restatement;
  * orc_fs_write / orc_fs_read: byte layout of output.fs.
"""
import itertools
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def tve_text(n, entries, vlabels=None):
    lines = ["t # 0", f"{n} {len(entries)} 1 1"]
    lines += [f"v {i} {vlabels[i] if vlabels else 1}" for i in range(n)]
    lines += [f"e {a} {b} {l}" for a, b, l in entries]
    lines.append("t # -1")
    return "\n".join(lines)


def undirected(edges, label=1):
    out = []
    for a, b in edges:
        out += [(a, b, label), (b, a, label)]
    return out


def edges_obj(n, entries, vlabels=None):
    s = [e[0] for e in entries]
    d = [e[1] for e in entries]
    l = [e[2] for e in entries]
    return O.Edges.from_arrays(n, s, d, l, vlabels)


def grid_entries(R, C):
    e = []
    for r in range(R):
        for c in range(C):
            v = r * C + c
            if c + 1 < C:
                e.append((v, v + 1))
            if r + 1 < R:
                e.append((v, v + C))
    return undirected(e)


def query(name):
    return O.Edges.load(os.path.join(GOLD, "queries", name + ".g"))


# ---------------------------------------------------------------- parsing / CCSR

def test_kat1_readme_ccsr(tmp_path):
    p = tmp_path / "readme.g"
    p.write_text("t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -1")
    g = O.Edges.load(p)
    c = O.CCSR(g)
    f = c.fields()
    assert [int(x) for x in c.packed()] == [1, 0x01000002, 1, 0x01000000]
    assert f["verticeLocs"].tolist() == [0, 2]
    assert f["invVerticeLocs"].tolist() == [0, 1]
    assert f["staggers"].tolist() == [4]
    assert f["stagOffset"].tolist() == [0] and f["stagNorm"].tolist() == [0]
    r = O.solve(g, g, rows=True)
    assert r.count == 2 and r.rows.tolist() == [[0, 1], [1, 0]]


def test_parse_error_codes(tmp_path):
    cases = {
        "missing": (None, 232),
        "badhead": ("t # 1\n2 0 1 1\nv 0 1\nv 1 1\nt # -1", 233),
        "fewv": ("t # 0\n3 0 1 1\nv 0 1\nv 1 1\nt # -1", 234),
        "fewe": ("t # 0\n2 3 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -1", 234),
        "badline": ("t # 0\n2 2 1 1\nv 0 1\nv 1 1\nx 0 1 1\nt # -1", 233),
        "noend": ("t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\n", 233),
    }
    for name, (text, code) in cases.items():
        p = tmp_path / f"{name}.g"
        if text is not None:
            p.write_text(text)
        with pytest.raises(O.OracleError) as ei:
            O.Edges.load(p)
        assert ei.value.code == code, name


def test_ccsr_invariants_random():
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(2, 40))
        m = int(rng.integers(1, 120))
        a = rng.integers(0, n, m)
        b = rng.integers(0, n, m)
        keep = a != b
        ent = undirected(list(set(zip(a[keep].tolist(), b[keep].tolist()))), label=int(rng.integers(1, 4)))
        if not ent:
            continue
        g = edges_obj(n, ent)
        c = O.CCSR(g)
        f = c.fields()
        tgt, rel, W = f["tgt"], f["rel"], f["size"]
        heads, i, degs = set(), 0, []
        while i < W:
            heads.add(i)
            d = int(tgt[i])
            degs.append(d)
            row = tgt[i + 1:i + 1 + d]
            assert np.all(row[:-1] <= row[1:]), "rows sorted by target"
            i += d + 1
        assert i == W and sum(d + 1 for d in degs) == W
        assert degs == sorted(degs, reverse=True), "rows in degree order"
        for k in range(W):
            if k not in heads:
                assert int(tgt[k]) in heads and rel[k] >= 1
        # hnWrite's id recovery is the inverse of verticeLocs
        for v in range(n):
            off = int(f["verticeLocs"][v])
            if off in heads:
                assert c.offset_to_id(off) == v
        # the entries as a multiset equal the input
        got = sorted((c.offset_to_id(h), c.offset_to_id(int(tgt[h + 1 + k])), int(rel[h + 1 + k]))
                     for h in heads for k in range(int(tgt[h])))
        assert got == sorted(ent)


# ---------------------------------------------------------------- query compiler

def test_kat2_triangle_requirements():
    c = O.CCSR(query("triangle"))
    assert [int(x) for x in c.packed()] == [2, 0x01000003, 0x01000006, 2, 0x01000000, 0x01000006, 2,
                                            0x01000000, 0x01000003]
    p = O.Plan(c).fields()
    # The reference's requirement words: {distance, relation} pairs where the
    # relation is taken from the packed CCSR word, (rel << 24) & relMask
    # (ccsr.cpp:148).  The oracle keeps the relation value itself (its CCSR
    # splits the relation out of the word, so relations need not fit 8 bits);
    # packed the reference's way, its table is KAT-2 word for word.
    R = 0x01000000
    req = p["requirements"].tolist()
    packed = [v << 24 if i % 2 else v for i, v in enumerate(req)]
    assert packed == [0, 0, 0, 0, 1, R, 0, 0, 1, R, 1, R]
    assert p["header"].tolist() == [0, 0, 0, 1, 4, 1, 2, 8, 1]


def plan_invariants(qedges):
    """Every query edge is covered once; firstQ = nearest placed neighbour."""
    c = O.CCSR(qedges)
    p = O.Plan(c).fields()
    q, mv = p["q"], p["maxValency"]
    col = p["col"].tolist()
    depth_of = {v: d for d, v in enumerate(col)}
    s, d_, _, _ = qedges.arrays()
    und = {(min(a, b), max(a, b)) for a, b in zip(s.tolist(), d_.tolist())}
    covered = []
    for d in range(q):
        num, off, fq = (int(x) for x in p["header"][3 * d:3 * d + 3])
        dist, dists = 0, []
        for j in range(num):
            dist += int(p["requirements"][off + 2 * j])
            dists.append(dist)
            u = col[d - dist]
            covered.append((min(u, col[d]), max(u, col[d])))
        if d:
            nb = [d - depth_of[u] for (a, b) in und for u in (a, b)
                  if col[d] in (a, b) and u != col[d] and depth_of[u] < d]
            assert fq == min(nb)
    assert sorted(covered) == sorted(und)


def test_plan_invariants_fixture_queries():
    for name in ("road2", "road3", "triangle", "road4", "roadY", "square", "plus", "road5"):
        plan_invariants(query(name))


def test_plan_disconnected_query_is_error():
    g = edges_obj(4, undirected([(0, 1), (2, 3)]))
    with pytest.raises(O.OracleError) as ei:
        O.Plan(O.CCSR(g))
    assert ei.value.code == 1


# ---------------------------------------------------------------- closed forms

@pytest.mark.parametrize("R,C", [(3, 4), (5, 7), (8, 6)])
def test_kat3_grid_closed_forms(R, C):
    g = edges_obj(R * C, grid_entries(R, C))
    deg = np.array([(r > 0) + (r < R - 1) + (c > 0) + (c < C - 1) for r in range(R) for c in range(C)])
    expect = {
        "road2": 2 * (R * (C - 1) + C * (R - 1)),
        "road3": int((deg * (deg - 1)).sum()),
        "triangle": 0,
        "square": 8 * (R - 1) * (C - 1),
        "roadY": int((deg * (deg - 1) * (deg - 2)).sum()),
        "plus": 24 * (R - 2) * (C - 2),
    }
    for name, want in expect.items():
        assert O.solve(g, query(name)).count == want, name


def path_edges(k):
    return undirected([(i, i + 1) for i in range(k - 1)])


def cycle_edges(k):
    return undirected([(i, (i + 1) % k) for i in range(k)])


def test_closed_forms_complete_path_cycle_bipartite():
    for n in (5, 7):
        kn = edges_obj(n, undirected([(a, b) for a in range(n) for b in range(a + 1, n)]))
        for name, q in (("road3", 3), ("triangle", 3), ("square", 4), ("plus", 5)):
            assert O.solve(kn, query(name)).count == math.perm(n, q), (n, name)
    for n, k in ((9, 3), (9, 5), (12, 7)):
        assert O.solve(edges_obj(n, path_edges(n)), edges_obj(k, path_edges(k))).count == 2 * (n - k + 1)
        assert O.solve(edges_obj(n, cycle_edges(n)), edges_obj(k, path_edges(k))).count == 2 * n
    for a, b in ((3, 4), (4, 5)):
        kab = edges_obj(a + b, undirected([(i, a + j) for i in range(a) for j in range(b)]))
        assert O.solve(kab, query("square")).count == 8 * math.comb(a, 2) * math.comb(b, 2)


# ---------------------------------------------------------------- golden vectors

def load_fixtures():
    with open(os.path.join(GOLD, "nx_fixtures.json")) as f:
        return json.load(f)["fixtures"]


def fixture_edges(spec):
    return O.Edges.from_arrays(spec["V"], spec["src"], spec["dst"], spec["elabel"], spec["vlabel"])


@pytest.mark.parametrize("fx", load_fixtures(), ids=lambda f: f["name"])
def test_networkx_golden(fx):
    r = O.solve(fixture_edges(fx["data"]), fixture_edges(fx["query"]), mode=fx["mode"], rows=True)
    want = np.array(fx["rows"], dtype=np.uint32).reshape(-1, fx["query"]["V"])
    assert r.count == len(fx["rows"])
    assert np.array_equal(r.rows, want)


def brute_force(n, data_entries, q, query_entries, vl=None, qvl=None):
    """All injective maps keeping every query entry (with label); vertex labels
    must match when given."""
    dset = set(data_entries)
    rows = []
    for f in itertools.permutations(range(n), q):
        if vl is not None and any(vl[f[i]] != qvl[i] for i in range(q)):
            continue
        if all((f[a], f[b], l) in dset for a, b, l in query_entries):
            rows.append(list(f))
    return sorted(rows)


def test_brute_force_random_small():
    rng = np.random.default_rng(11)
    for trial in range(25):
        n = int(rng.integers(4, 8))
        edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < 0.5]
        if not edges:
            continue
        labels = {e: int(rng.integers(1, 3)) for e in edges}
        ent = [(a, b, labels[(a, b)]) for a, b in edges] + [(b, a, labels[(a, b)]) for a, b in edges]
        vl = [int(x) for x in rng.integers(1, 3, n)]
        q = int(rng.integers(2, 5))
        # query: a connected random labelled tree plus maybe one edge
        qe = [(i, int(rng.integers(0, i)), int(rng.integers(1, 3))) for i in range(1, q)]
        qent = [(a, b, l) for a, b, l in qe] + [(b, a, l) for a, b, l in qe]
        qvl = [int(x) for x in rng.integers(1, 3, q)]
        for mode in (O.MODE_REF, O.MODE_VLABEL):
            r = O.solve(edges_obj(n, ent, vl), edges_obj(q, qent, qvl), mode=mode, rows=True)
            want = brute_force(n, ent, q, qent, vl if mode else None, qvl if mode else None)
            assert r.rows.tolist() == want, (trial, mode)


def test_level_sizes_are_prefix_counts():
    """S_k (survivors after placing k+1 vertices) = embeddings of the
    sub-query induced on the first k+1 placed vertices."""
    n = 7
    ent = undirected([(a, b) for a in range(n) for b in range(a + 1, n) if (a * 3 + b) % 4 != 0])
    data = edges_obj(n, ent)
    qe = [(0, 1), (1, 2), (2, 3), (3, 0), (2, 4)]
    qent = undirected(qe)
    qg = edges_obj(5, qent)
    c = O.CCSR(qg)
    p = O.Plan(c)
    col = p.fields()["col"].tolist()
    r = O.match(O.CCSR(data), c, p)
    for k in range(5):
        placed = col[:k + 1]
        sub = [(placed.index(a), placed.index(b), 1) for a, b, _ in qent if a in placed and b in placed]
        assert r.level[k] == len(brute_force(n, ent, k + 1, sub)), k


# ---------------------------------------------------------------- checker / hash / output.fs

def test_soundness_checker_finds_planted_error():
    g = edges_obj(5 * 5, grid_entries(5, 5))
    q = query("square")
    r = O.solve(g, q, rows=True)
    assert O.check_sound(q, g, r.rows) == r.count
    bad = r.rows.copy()
    bad[17, 2] = (bad[17, 2] + 7) % 25
    assert O.check_sound(q, g, bad) == 17


def mix64(z):
    M = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def py_row_hash(row):
    a = b = 0
    for c, v in enumerate(row):
        x = int(v) | (c << 32)
        a = (a + mix64(0x9E3779B97F4A7C15 ^ x)) % (1 << 64)
        b = (b + mix64(0xD1B54A32D192ED03 ^ x)) % (1 << 64)
    return mix64(a), mix64(b)


def test_row_hash_and_digest():
    rng = np.random.default_rng(3)
    for w in (1, 2, 5, 9):
        row = rng.integers(0, 1 << 32, w, dtype=np.uint64).astype(np.uint32)
        assert O.row_hash(row) == py_row_hash(row)
    g = edges_obj(16, grid_entries(4, 4))
    r = O.solve(g, query("road3"), rows=True)
    d0 = sum(py_row_hash(x)[0] for x in r.rows) % (1 << 64)
    d1 = sum(py_row_hash(x)[1] for x in r.rows) % (1 << 64)
    assert r.digest == (d0, d1)


def test_output_fs_layout(tmp_path):
    rows = np.array([[3, 1, 4], [1, 5, 9]], dtype=np.uint32)
    p = tmp_path / "output.fs"
    O.fs_write(p, rows)
    raw = p.read_bytes()
    assert raw[:8] == struct.pack("<II", 2, 3)
    assert struct.unpack("<6I", raw[8:]) == (3, 1, 4, 1, 5, 9)
    assert np.array_equal(O.fs_read(p), rows)


def brute_force_rule(n, data_entries, q, query_entries, order):
    """The reference's placement rule taken literally, over all injective maps
    (ccsr.cpp:142-167, HnGPU.cu:76-109, 243-257): in the placement order, each
    query entry a->b whose target b is placed no later than a (b == a is a
    self-loop) requires the data entry f(b)->f(a) with the label of b->a, or
    of a->b when the query has no b->a entry; the root's image needs an entry
    of its own.  Entries are unique per ordered pair."""
    dset = set(data_entries)
    outdeg = [0] * n
    for a, _, _ in data_entries:
        outdeg[a] += 1
    qlab = {(a, b): l for a, b, l in query_entries}
    pos = {v: i for i, v in enumerate(order)}
    reqs = [(a, b, qlab.get((b, a), l)) for (a, b), l in qlab.items() if pos[b] <= pos[a]]
    rows = []
    for f in itertools.permutations(range(n), q):
        if outdeg[f[order[0]]] and all((f[b], f[a], l) in dset for a, b, l in reqs):
            rows.append(list(f))
    return sorted(rows)


def directed_entries(rng, n, p_edge=0.6, p_loop=0.25, nlabels=2, all_rows=True, p_both=0.7):
    """Random directed entries, unique per ordered pair, with self-loops; with
    all_rows every vertex gets an out-entry (a vertex without one has no row)."""
    ent = []
    for a in range(n):
        if rng.random() < p_loop:
            ent.append((a, a, int(rng.integers(1, nlabels + 1))))
        for b in range(a + 1, n):
            if rng.random() < p_edge:
                way = rng.random()  # both ways (p_both, labels drawn apart), else one way
                one = (1.0 - p_both) / 2
                if way < p_both + one:
                    ent.append((a, b, int(rng.integers(1, nlabels + 1))))
                if way < p_both or way >= p_both + one:
                    ent.append((b, a, int(rng.integers(1, nlabels + 1))))
    has_out = {a for a, _, _ in ent}
    if all_rows:
        ent += [(a, (a + 1) % n, 1) for a in range(n) if a not in has_out]
    return ent


def test_directed_self_loops_against_rule():
    """Asymmetric entries, self-loops and two labels (REF mode), on graphs
    whose vertices all have out-entries: the oracle equals the literal rule on
    every query the compiler accepts, and rejects the queries whose later
    vertices have no entry back to placed ones."""
    rng = np.random.default_rng(23)
    checked = nonzero = rejected = 0
    for trial in range(400):
        n = int(rng.integers(5, 8))
        ent = directed_entries(rng, n)
        q = int(rng.integers(2, 5))
        verts = [int(v) for v in rng.choice(n, q, replace=False)]
        idx = {v: i for i, v in enumerate(verts)}
        qent = [(idx[a], idx[b], l) for a, b, l in ent if a in idx and b in idx and rng.random() < 0.9]
        if {a for a, _, _ in qent} != set(range(q)):
            continue  # a query vertex without an out-entry has no row
        und = {(a, b) for a, b, _ in qent} | {(b, a) for a, b, _ in qent}
        seen, todo = {0}, [0]
        while todo:
            x = todo.pop()
            for y in range(q):
                if (x, y) in und and y not in seen:
                    seen.add(y)
                    todo.append(y)
        if len(seen) < q:
            continue  # disconnected: not a query
        try:
            qc = O.CCSR(edges_obj(q, qent), O.MODE_REF)
            order = [int(c) for c in O.Plan(qc).fields()["col"]]
        except O.OracleError as e:
            assert e.code == 1
            rejected += 1
            continue
        r = O.solve(edges_obj(n, ent), edges_obj(q, qent), mode=O.MODE_REF, rows=True)
        want = brute_force_rule(n, ent, q, qent, order)
        got = sorted(r.rows.tolist()) if r.count else []
        assert got == want, (trial, ent, qent, order)
        checked += 1
        nonzero += bool(want)
    assert checked >= 20 and nonzero >= 5, (checked, nonzero, rejected)


# ---------------------------------------------------------------- canonical order and per-depth cap

def capped_by_levels(n, data_entries, q, query_entries, order, cap, root_lo=0, root_hi=None):
    """The per-depth cap restated breadth first, independently of the oracle's
    depth-first search: build every level of candidates in canonical order
    (roots by id; children of each survivor, in level order, from its
    generator's entries of the generator relation by ascending target id),
    number the candidates of each level, and count the leaves whose own and
    ancestors' numbers are all below the cap.  The match is capped when some
    level holds more candidates than the cap.  Symmetric simple data only."""
    adj = {}
    for a, b, l in data_entries:
        adj.setdefault(a, {}).setdefault(l, []).append(b)
    for a in adj:
        for l in adj[a]:
            adj[a][l].sort()
    qlab = {(a, b): l for a, b, l in query_entries}
    depth = {v: i for i, v in enumerate(order)}
    reqs = []  # per depth: [(depth of u, relation)]; distance 0 = self loop
    for d, nq in enumerate(order):
        r = [(depth[u], qlab.get((u, nq), l)) for (a, u), l in qlab.items() if a == nq and depth[u] <= d]
        reqs.append(sorted(r, key=lambda t: (d - t[0], t[1])))
    gen = [None] + [min((t for t in reqs[d] if t[0] < d), key=lambda t: (d - t[0], t[1])) for d in range(1, q)]
    hi = n if root_hi is None else min(root_hi, n)
    level = []  # (path, ok) with the candidate's numbers along the path
    r1 = gen[1][1] if q > 1 else None
    for x in range(root_lo, hi):
        if x not in adj or (r1 is not None and not adj[x].get(r1)):
            continue
        ok = all(x in adj[x].get(l, []) for du, l in reqs[0])
        level.append(([x], ok, [len(level)]))
    capped = len(level) > cap
    for d in range(1, q):
        nxt = []
        for path, ok, nums in level:
            if not ok:
                continue
            gd, gl = gen[d]
            for t in adj.get(path[gd], {}).get(gl, []):
                good = t not in path and all(
                    t in adj.get(t if du == d else path[du], {}).get(l, []) for du, l in reqs[d])
                nxt.append((path + [t], good, nums + [len(nxt)]))
        capped |= len(nxt) > cap
        level = nxt
    rows = [p for p, ok, nums in level if ok and max(nums) < cap]
    out = []
    for p in rows:
        row = [0] * q
        for d, v in enumerate(p):
            row[order[d]] = v
        out.append(row)
    return out, capped


def test_canonical_cap_against_breadth_first_restatement():
    """orc_match_canon: without a cap it finds orc_match's embeddings; under a
    cap its partial count and digest and its capped flag equal the
    breadth-first restatement's, for caps from 1 up to past every level."""
    rng = np.random.default_rng(2024)
    cases = 0
    for trial in range(40):
        n = int(rng.integers(5, 11))
        edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < 0.55]
        if not edges:
            continue
        labels = {e: int(rng.integers(1, 3)) for e in edges}
        ent = [(a, b, labels[(a, b)]) for a, b in edges] + [(b, a, labels[(a, b)]) for a, b in edges]
        q = int(rng.integers(2, 6))
        qe = [(i, int(rng.integers(0, i)), int(rng.integers(1, 3))) for i in range(1, q)]
        if q > 2 and rng.random() < 0.5:
            a, b = int(rng.integers(0, q)), int(rng.integers(0, q))
            if a != b and all({a, b} != {x, y} for x, y, _ in qe):
                qe.append((a, b, int(rng.integers(1, 3))))
        qent = [(a, b, l) for a, b, l in qe] + [(b, a, l) for a, b, l in qe]
        dg = O.CCSR(edges_obj(n, ent), O.MODE_REF)
        qg = O.CCSR(edges_obj(q, qent), O.MODE_REF)
        p = O.Plan(qg)
        cn = O.Canon(dg)
        order = [int(x) for x in p.fields()["col"]]
        full = O.match(dg, qg, p)
        r, capped, cand = O.match_canon(dg, cn, qg, p, 0)
        assert (r.count, r.digest, capped) == (full.count, full.digest, False)
        rows_all, _ = capped_by_levels(n, ent, q, qent, order, 1 << 40)
        assert len(rows_all) == full.count
        lo = int(rng.integers(0, 3))
        for cap in sorted({1, 2, 3, 5, 8, 13, max(1, max(cand)), max(cand) + 1, 1000}):
            r, capped, _ = O.match_canon(dg, cn, qg, p, cap, root_lo=lo)
            want, wcap = capped_by_levels(n, ent, q, qent, order, cap, root_lo=lo)
            dg0 = dg1 = 0
            for row in want:
                h = O.row_hash(row)
                dg0, dg1 = (dg0 + h[0]) % (1 << 64), (dg1 + h[1]) % (1 << 64)
            assert (r.count, capped) == (len(want), wcap), (trial, cap, r.count, len(want), capped, wcap)
            assert r.digest == (dg0, dg1)
            cases += 1
    assert cases > 150
