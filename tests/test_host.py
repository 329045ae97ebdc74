"""Host-side logic of libgfsm on the CPU: the C ABI surface, the t/v/e parser,
the relation-partitioned layout, the query planner (against the oracle's
restatement of the reference compiler) and the workload generators.  No
device calls."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import gfsm_amd as G
import oracle as O
from helpers import to_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "gfsm.h")).read()
    declared = set(re.findall(r"\b(gfsm_\w+)\s*\(", header))
    assert len(declared) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", G.library_path()], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (gfsm_\w+)", out))
    assert declared <= exported, declared - exported
    G.lib()  # loads; every declared function has a ctypes prototype
    assert G.lib().gfsm_version().startswith(b"gfsm")


def test_parser_matches_oracle_and_error_codes(tmp_path):
    for name in os.listdir(os.path.join(GOLD, "queries")):
        p = os.path.join(GOLD, "queries", name)
        g = G.Graph.load(p)
        o = O.Edges.load(p)
        V, s, d, l, vl = g.edges()
        os_, od, ol, ovl = o.arrays()
        assert V == o.V and np.array_equal(s, os_) and np.array_equal(d, od) and np.array_equal(l, ol)
    cases = {
        "missing": (None, 232), "badhead": ("t # 1\n1 0 1 1\nv 0 1\nt # -1", 233),
        "fewv": ("t # 0\n3 0 1 1\nv 0 1\nt # -1", 234),
        "fewe": ("t # 0\n2 3 1 1\nv 0 1\nv 1 1\ne 0 1 1\nt # -1", 234),
        "badline": ("t # 0\n2 1 1 1\nv 0 1\nv 1 1\nq 0 1 1\nt # -1", 233),
    }
    for name, (text, code) in cases.items():
        p = tmp_path / f"{name}.g"
        if text is not None:
            p.write_text(text)
        with pytest.raises(G.GfsmError) as ei:
            G.Graph.load(p)
        assert ei.value.code == code, name


def test_rejects_out_of_domain_entries():
    with pytest.raises(G.GfsmError):
        G.Graph.from_edges(3, [0], [5], [1])          # vertex id >= V
    with pytest.raises(G.GfsmError):
        G.Graph.from_edges(3, [0], [1], [0])          # label 0 (never remapped by the reference)
    with pytest.raises(G.GfsmError):
        G.Graph.from_edges(3, [0], [1], [1], vlabel=[5000, 1, 1], mode=G.MODE_VLABEL)


def check_layout(g):
    V, s, d, l, vl = g.edges()
    info = g.info()
    dir_ptr, dirs, nbr = g.layout()
    assert dir_ptr[0] == 0 and np.all(np.diff(dir_ptr.astype(np.int64)) >= 0)
    assert dirs[-1, 1] == len(nbr) == len(s)
    got = []
    for x in range(V):
        rels = dirs[dir_ptr[x]:dir_ptr[x + 1], 0]
        assert np.all(np.diff(rels.astype(np.int64)) > 0), "relations ascending, one segment each"
        for j in range(dir_ptr[x], dir_ptr[x + 1]):
            seg = nbr[dirs[j, 1]:dirs[j + 1, 1]]
            assert len(seg) > 0 and np.all(np.diff(seg.astype(np.int64)) >= 0)
            got += [(x, int(y), int(dirs[j, 0])) for y in seg]
    if info["mode"] == G.MODE_REF:
        rel = l
    else:
        rel = (vl[s].astype(np.uint64) << 21) | (vl[d].astype(np.uint64) << 10) | l
    want = sorted(zip(s.tolist(), d.tolist(), [int(r) for r in rel]))
    assert sorted(got) == want


@pytest.mark.parametrize("mode", [G.MODE_REF, G.MODE_VLABEL])
def test_layout_invariants(mode):
    check_layout(G.Graph.er(300, 5.0, nvlabels=3, nelabels=3, seed=4, mode=mode))
    check_layout(G.Graph.rmat(9, 6, nvlabels=4, nelabels=2, seed=5, mode=mode))


def test_planner_matches_oracle_restatement():
    qs = [G.Graph.load(os.path.join(GOLD, "queries", n)) for n in sorted(os.listdir(os.path.join(GOLD, "queries")))]
    data = G.Graph.rmat(10, 8, nvlabels=3, nelabels=2, seed=9, mode=G.MODE_VLABEL)
    for i in range(40):
        qs.append(G.Graph.walk_query(data, 2 + i % 12, kind=i % 2, seed=100 + i))
    with open(os.path.join(GOLD, "nx_fixtures.json")) as f:
        for fx in json.load(f)["fixtures"]:
            q = fx["query"]
            qs.append(G.Graph.from_edges(q["V"], q["src"], q["dst"], q["elabel"], q["vlabel"], mode=fx["mode"]))
    for q in qs:
        col, hdr, req = q.plan()
        mode = q.info()["mode"]
        p = O.Plan(O.CCSR(to_oracle(q), mode)).fields()
        assert col.tolist() == p["col"].tolist()
        assert hdr.tolist() == p["header"].tolist()
        assert req.tolist() == p["requirements"].tolist()


def test_planner_rejects_disconnected_query():
    q = G.Graph.from_edges(4, [0, 1, 2, 3], [1, 0, 3, 2], [1, 1, 1, 1])
    with pytest.raises(G.GfsmError) as ei:
        q.plan()
    assert ei.value.code == 1


def test_generators_deterministic_simple_symmetric():
    a = G.Graph.rmat(11, 8, nvlabels=8, zipf_s=1.2, nelabels=4, seed=77, mode=G.MODE_VLABEL)
    b = G.Graph.rmat(11, 8, nvlabels=8, zipf_s=1.2, nelabels=4, seed=77, mode=G.MODE_VLABEL)
    for x, y in zip(a.edges()[1:], b.edges()[1:]):
        assert np.array_equal(x, y)
    V, s, d, l, vl = a.edges()
    assert np.all(s != d), "no self loops"
    pairs = set(zip(s.tolist(), d.tolist()))
    assert len(pairs) == len(s), "no parallel entries"
    lab = dict(zip(zip(s.tolist(), d.tolist()), l.tolist()))
    assert all((y, x) in pairs and lab[(y, x)] == lab[(x, y)] for x, y in pairs), "symmetric, same label"
    assert vl.min() >= 1 and vl.max() <= 8 and l.min() >= 1 and l.max() <= 4
    # Zipf: label 1 is the most frequent
    counts = np.bincount(vl, minlength=9)[1:]
    assert counts[0] == counts.max()


def test_grid_generator_counts():
    R = C = 40
    g = G.Graph.grid(R, C, del_frac=0.05, diag_frac=0.02, seed=3)
    e0 = R * (C - 1) + C * (R - 1)
    want = e0 - round(0.05 * e0) + round(0.02 * e0)
    assert g.info()["entries"] == 2 * want
    assert g.info()["max_degree"] <= 8


def test_walk_queries_are_induced_connected_subgraphs():
    data = G.Graph.er(500, 6.0, nvlabels=3, nelabels=2, seed=21, mode=G.MODE_VLABEL)
    _, ds, dd, dl, dvl = data.edges()
    dset = set(zip(ds.tolist(), dd.tolist(), dl.tolist()))
    for kind in (G.WALK_DENSE, G.WALK_SPARSE):
        for seed in range(10):
            q = G.Graph.walk_query(data, 6, kind=kind, seed=seed)
            V, s, d, l, vl = q.edges()
            assert V == 6
            und = {(min(a, b), max(a, b)) for a, b in zip(s.tolist(), d.tolist())}
            if kind == G.WALK_SPARSE:
                assert len(und) <= V  # walk tree (V-1 edges) plus at most one extra
            col, hdr, req = q.plan()  # connected: plans without error
            assert sorted(col.tolist()) == list(range(6))


def test_cli_input_errors_need_no_device(tmp_path):
    cli = os.path.join(ROOT, "gfsm_amd", "GFSM")
    q = os.path.join(GOLD, "queries", "square.g")
    assert subprocess.run([cli, q, str(tmp_path / "nope.g")], capture_output=True).returncode == 232
    bad = tmp_path / "bad.g"
    bad.write_text("t # 0\n3 0 1 1\nv 0 1\nt # -1")
    assert subprocess.run([cli, q, str(bad)], capture_output=True).returncode == 234
    disc = tmp_path / "disc.g"
    disc.write_text("t # 0\n4 4 1 1\nv 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 1 0 1\ne 2 3 1\ne 3 2 1\nt # -1")
    assert subprocess.run([cli, str(disc), q], capture_output=True).returncode == 1


@pytest.mark.parametrize("mode", [G.MODE_REF, G.MODE_VLABEL])
def test_binary_image_round_trip(tmp_path, mode):
    """A saved image loads back to the same input arrays, layout, relation
    index (seen through root sets and windows) and oracle result."""
    for g in (G.Graph.er(400, 6.0, nvlabels=3, nelabels=2, seed=8, mode=mode),
              G.Graph.rmat(10, 5, nvlabels=4, nelabels=3, seed=9, mode=mode),
              G.Graph.from_edges(5, [], [], [], mode=mode)):
        p = tmp_path / "g.bin"
        g.save_bin(p)
        h = G.Graph.load_bin(p)
        assert h.info() == g.info()
        for a, b in zip(g.edges()[1:], h.edges()[1:]):
            assert np.array_equal(a, b)
        for a, b in zip(g.layout(), h.layout()):
            assert np.array_equal(a, b)
        if g.info()["entries"] == 0:
            continue
        check_layout(h)
        for seed in range(4):
            q = G.Graph.walk_query(g, 4, kind=seed % 2, seed=seed)
            assert np.array_equal(g.roots(q), h.roots(q))
            assert g.window(q, 300) == h.window(q, 300)
        q = G.Graph.walk_query(g, 3, seed=11)
        a = O.solve(to_oracle(g), to_oracle(q), mode=mode, rows=False)
        b = O.solve(to_oracle(h), to_oracle(q), mode=mode, rows=False)
        assert (a.count, a.digest) == (b.count, b.digest)


def test_binary_image_rejects_damage(tmp_path):
    g = G.Graph.er(200, 4.0, nvlabels=2, nelabels=2, seed=3, mode=G.MODE_VLABEL)
    p = tmp_path / "g.bin"
    g.save_bin(p)
    raw = bytearray(p.read_bytes())
    header = 8 + 4 * 4 + 6 * 8
    cases = {
        "magic": raw[:0] + b"XFSMBIN1" + raw[8:],
        "version": raw[:8] + (7).to_bytes(4, "little") + raw[12:],
        "mode": raw[:12] + (9).to_bytes(4, "little") + raw[16:],
        "truncated": raw[:-5],
        "extended": raw + b"\0" * 8,
        "payload_bit": raw[:header + 40] + bytes([raw[header + 40] ^ 1]) + raw[header + 41:],
        "last_byte": raw[:-1] + bytes([raw[-1] ^ 0x80]),
        "short_header": raw[:20],
        # a degree beyond the kernels' 2^26 bound, and counts whose byte sizes
        # would wrap a 64-bit sum: rejected before anything is allocated
        "max_degree": raw[:20] + (1 << 26).to_bytes(4, "little") + raw[24:],
        "huge_counts": raw[:24] + (1 << 62).to_bytes(8, "little") + raw[32:40] + ((1 << 64) - (1 << 62)).to_bytes(8, "little") + raw[48:],
    }
    for name, blob in cases.items():
        q = tmp_path / f"{name}.bin"
        q.write_bytes(bytes(blob))
        with pytest.raises(G.GfsmError) as ei:
            G.Graph.load_bin(q)
        assert ei.value.code == 233, name
    with pytest.raises(G.GfsmError) as ei:
        G.Graph.load_bin(tmp_path / "missing.bin")
    assert ei.value.code == 232
    with pytest.raises(G.GfsmError) as ei:  # a text graph is not an image
        G.Graph.load_bin(os.path.join(GOLD, "queries", "square.g"))
    assert ei.value.code == 233


def test_cli_cache_uses_image_of_same_mode(tmp_path):
    """--cache writes the image on the first run; a later run reads it (a
    damaged image would then fail, so it must be rebuilt when the data file
    is newer or the mode differs)."""
    cli = os.path.join(ROOT, "gfsm_amd", "GFSM")
    q = os.path.join(GOLD, "queries", "square.g")
    data = tmp_path / "bad.g"
    data.write_text("t # 0\n3 0 1 1\nv 0 1\nt # -1")  # missing data: fails before any device use
    cache = tmp_path / "d.bin"
    r = subprocess.run([cli, q, str(data), "--cache", str(cache)], capture_output=True)
    assert r.returncode == 234 and not cache.exists()
    # an image of a valid graph in vlabel mode, newer than the (bad) text
    G.Graph.er(50, 3.0, seed=1, mode=G.MODE_VLABEL).save_bin(cache)
    os.utime(data, (1, 1))
    r = subprocess.run([cli, q, str(data), "--cache", str(cache)], capture_output=True)
    assert r.returncode == 234  # ref mode asked: image of another mode ignored, text parsed
    r = subprocess.run([cli, q, str(data), "--cache", str(cache), "--mode", "vlabel", "--device", "99"],
                       capture_output=True)
    assert r.returncode != 234  # image used: parsing skipped, so the (absent) device is what fails
    later = os.stat(cache).st_mtime + 10
    os.utime(data, (later, later))  # text newer than the image: parsed again
    r = subprocess.run([cli, q, str(data), "--cache", str(cache), "--mode", "vlabel"], capture_output=True)
    assert r.returncode == 234


def test_cost_order_is_a_connected_permutation():
    """The statistics-driven order is a permutation in which every later
    vertex has an out-entry to an earlier one (so it has a generator), and the
    default order is the reference planner's."""
    data = G.Graph.rmat(11, 8, nvlabels=4, zipf_s=1.0, nelabels=2, seed=21, mode=G.MODE_VLABEL)
    differ = 0
    for seed in range(16):
        q = G.Graph.walk_query(data, 3 + seed % 8, kind=seed % 2, seed=600 + seed)
        V, s, d, l, vl = q.edges()
        ref, cost = data.order(q), data.order(q, cost=True)
        assert np.array_equal(ref, q.plan()[0])
        for col in (ref, cost):
            assert sorted(col.tolist()) == list(range(V))
            pos = {int(v): i for i, v in enumerate(col)}
            for i in range(1, V):
                v = int(col[i])
                assert any(pos[int(t)] < i for u, t in zip(s, d) if int(u) == v), (seed, col)
        differ += not np.array_equal(ref, cost)
    assert differ > 0


def test_cost_order_roots_at_the_rarest_relation():
    """A labelled path whose last relation has one source vertex in the data
    graph: the statistics-driven order starts at an end of that relation, the
    reference order at the lowest-id vertex of highest out-degree."""
    # data: labels 1,1,...; vertex 0 has label 2 and one neighbour; all others label 1
    nv = 64
    src, dst = [], []
    for i in range(1, nv - 1):
        src += [i, i + 1]
        dst += [i + 1, i]
    src += [0, 1]
    dst += [1, 0]
    vl = [2] + [1] * (nv - 1)
    data = G.Graph.from_edges(nv, src, dst, [1] * len(src), vlabel=vl, mode=G.MODE_VLABEL)
    # query path: 0(label 1) - 1(label 1) - 2(label 1) - 3(label 2)
    q = G.Graph.from_edges(4, [0, 1, 1, 2, 2, 3], [1, 0, 2, 1, 3, 2], [1] * 6, vlabel=[1, 1, 1, 2],
                           mode=G.MODE_VLABEL)
    assert int(data.order(q, cost=True)[0]) in (2, 3)
    assert int(data.order(q)[0]) == 1


def test_read_fs_reads_the_oracle_layout_and_rejects_damage(tmp_path):
    rows = np.array([[3, 1, 4], [1, 5, 9], [2, 6, 5]], dtype=np.uint32)
    p = tmp_path / "output.fs"
    O.fs_write(p, rows)
    assert np.array_equal(G.read_fs(p), rows)
    raw = p.read_bytes()
    (tmp_path / "short.fs").write_bytes(raw[:-4])
    (tmp_path / "tiny.fs").write_bytes(raw[:4])
    for name in ("short.fs", "tiny.fs"):
        with pytest.raises(G.GfsmError) as ei:
            G.read_fs(tmp_path / name)
        assert ei.value.code == 233
    empty = tmp_path / "empty.fs"
    O.fs_write(empty, np.zeros((0, 4), np.uint32))
    assert G.read_fs(empty).shape == (0, 4)
    assert G.format_rows(rows, limit=2) == "0->3 1->1 2->4\n0->1 1->5 2->9\n... 1 more"


@pytest.mark.parametrize("mode,nvl", [(G.MODE_REF, 1), (G.MODE_VLABEL, 3)])
def test_level_projection_tracks_the_oracle_on_random_graphs(mode, nvl):
    """The projector's expected level sizes (relation statistics, checks as
    independent events) against the oracle's exact level sizes on ER graphs:
    depth 0 is the root candidate count, depth 1 is exact without checks, and
    deeper levels stay within 1.6x (counting vertices without a segment of a
    relation as length 0 matters: without it labelled levels drift 2-5x)."""
    data = G.Graph.er(3000, 8.0, nvlabels=nvl, nelabels=2, seed=5, mode=mode)
    for s in range(3, 7):
        for kind in (G.WALK_DENSE, G.WALK_SPARSE):
            q = G.Graph.walk_query(data, s, kind=kind, seed=100 + s * 3 + kind)
            nodes, cand = data.projection(q)
            assert nodes[0] == len(data.roots(q))
            lv = np.array(oracle_levels(data, q, mode), float)
            for k in range(1, s):
                if lv[k] >= 200:
                    assert 1 / 1.6 <= nodes[k] / lv[k] <= 1.6, (s, kind, k, nodes.tolist(), lv.tolist())
    q = G.Graph.walk_query(data, 4, seed=7)
    roots = data.roots(q)
    lo, hi = int(roots[10]), int(roots[20])
    nodes, _ = data.projection(q, root_lo=lo, root_hi=hi)
    assert nodes[0] == 10


def oracle_levels(data, q, mode):
    return O.solve(to_oracle(data), to_oracle(q), mode=mode, rows=False).level


def test_host_code_under_sanitizers(tmp_path):
    """The parser, image loader, layout builder, generators and planner built
    with AddressSanitizer and UBSan (tests/host_asan): valid inputs, malformed
    text, damaged and crafted images and generated query plans run without a
    sanitizer report and with the expected statuses."""
    import shutil
    import subprocess

    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_asan")
    subprocess.run(["make", "-s", "-C", here], check=True, capture_output=True)
    r = subprocess.run([os.path.join(here, "_build", "asan_main"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
