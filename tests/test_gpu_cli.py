"""The GFSM command line as a drop-in for the reference's process boundary
(main.cu:56-63): same arguments, CSV row, exit status and output.fs layout,
on the GPU, checked against the CPU oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

import gfsm_amd as G
import oracle as O
from helpers import assert_same, oracle_solve, to_oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gfsm_amd", "GFSM")


@pytest.mark.parametrize("mode,flag", [(G.MODE_REF, "ref"), (G.MODE_VLABEL, "vlabel")])
def test_cli_matches_oracle(tmp_path, mode, flag):
    data = G.Graph.er(800, 6.0, nvlabels=3, nelabels=2, seed=321, mode=mode)
    query = G.Graph.walk_query(data, 4, kind=G.WALK_DENSE, seed=322)
    dp, qp, out = tmp_path / "data.g", tmp_path / "query.g", tmp_path / "output.fs"
    data.write(dp)
    query.write(qp)
    r = subprocess.run([CLI, str(qp), str(dp), "--mode", flag, "--out", str(out), "--sorted"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = oracle_solve(data, query, mode)
    fields = r.stdout.strip().split(",")
    assert fields[0] == "query.g" and fields[1] == "data.g"
    assert str(ref.count) in fields, r.stdout
    raw = out.read_bytes()
    n, width = struct.unpack("<II", raw[:8])
    rows = np.frombuffer(raw[8:], dtype="<u4").reshape(n, width)
    assert n == ref.count and width == ref.width
    assert np.array_equal(rows, ref.rows)


def test_streamed_output_fs_equals_in_memory(ctx, tmp_path):
    """gfsm_match_write_fs streams each chunk's rows to output.fs: the file
    holds the oracle's rows (as a set; sorted when asked), also with tiny
    level buffers (many chunks) and for a one-vertex query."""
    data = G.Graph.grid(40, 40, del_frac=0.05, diag_frac=0.02, seed=5, mode=G.MODE_REF)
    query = G.Graph.walk_query(data, 5, kind=G.WALK_SPARSE, seed=17)
    ref = oracle_solve(data, query, G.MODE_REF)
    for budget in (0, 1 << 20):
        p = tmp_path / f"s{budget}.fs"
        m = ctx.match_to_fs(data, query, p, mem_budget=budget, digest=True)
        got = G.read_fs(p)
        assert m.count == ref.count == len(got) and m.digest == ref.digest
        assert np.array_equal(got[np.lexsort(got.T[::-1])], ref.rows)
    p = tmp_path / "sorted.fs"
    ctx.match_to_fs(data, query, p, sorted=True)
    assert np.array_equal(G.read_fs(p), ref.rows)
    one = G.Graph.walk_query(data, 1, seed=3)
    p = tmp_path / "one.fs"
    m1 = ctx.match_to_fs(data, one, p)
    full = ctx.match(data, one, rows=True, sorted=True)  # one-vertex queries: the oracle has none
    assert np.array_equal(np.sort(G.read_fs(p)[:, 0]), full.rows[:, 0]) and m1.count == full.count


def test_cli_stream_and_print(tmp_path):
    data = G.Graph.er(500, 5.0, nvlabels=2, seed=77, mode=G.MODE_VLABEL)
    query = G.Graph.walk_query(data, 3, seed=78)
    dp, qp = tmp_path / "data.g", tmp_path / "query.g"
    data.write(dp)
    query.write(qp)
    a, b = tmp_path / "a.fs", tmp_path / "b.fs"
    r1 = subprocess.run([CLI, str(qp), str(dp), "--mode", "vlabel", "--out", str(a)], capture_output=True,
                        text=True, timeout=120)
    r2 = subprocess.run([CLI, str(qp), str(dp), "--mode", "vlabel", "--out", str(b), "--stream", "--print", "3"],
                        capture_output=True, text=True, timeout=120)
    assert r1.returncode == 0 and r2.returncode == 0, (r1.stderr, r2.stderr)
    ra, rb = G.read_fs(a), G.read_fs(b)
    key = lambda x: x[np.lexsort(x.T[::-1])]
    assert len(ra) == len(rb) > 3 and np.array_equal(key(ra), key(rb))
    lines = [l for l in r2.stderr.splitlines() if "->" in l]
    assert lines == [G.format_rows(rb[i:i + 1]) for i in range(3)]


def test_binary_image_matches_like_the_text_graph(ctx, tmp_path):
    """A data graph loaded from its binary image matches exactly like the
    graph parsed from text, through the library and through GFSM --cache."""
    data = G.Graph.rmat(11, 6, nvlabels=3, zipf_s=1.0, nelabels=2, seed=88, mode=G.MODE_VLABEL)
    dp, qp, img = tmp_path / "data.g", tmp_path / "query.g", tmp_path / "data.bin"
    data.write(dp)
    text = G.Graph.load(dp, mode=G.MODE_VLABEL)
    text.save_bin(img)
    image = G.Graph.load_bin(img)
    od = O.Edges.load(dp)  # the oracle parses the same text file
    for s in (3, 4, 5, 6):
        q = G.Graph.walk_query(text, s, seed=880 + s)
        a, b = ctx.match(text, q, digest=True), ctx.match(image, q, digest=True)
        assert (a.count, a.digest) == (b.count, b.digest)
        # the image-loaded graph against the oracle: count, digest and sorted rows
        ref = O.solve(od, to_oracle(q), mode=G.MODE_VLABEL, rows=True)
        assert_same(ctx.match(image, q, rows=True, sorted=True, digest=True), ref)
        assert (b.count, b.digest) == (ref.count, ref.digest)
    q = G.Graph.walk_query(text, 4, seed=900)
    q.write(qp)
    cache = tmp_path / "cache.bin"
    runs = [subprocess.run([CLI, str(qp), str(dp), "--mode", "vlabel", "--cache", str(cache)], capture_output=True,
                           text=True, timeout=120) for _ in range(2)]
    assert all(r.returncode == 0 for r in runs), [r.stderr for r in runs]
    assert cache.exists()
    counts = [r.stdout.strip().split(",")[8] for r in runs]  # "Function Count" column
    assert counts[0] == counts[1] == str(ctx.match(text, q).count)
