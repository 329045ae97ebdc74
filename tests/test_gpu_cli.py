"""The GFSM command line as a drop-in for the reference's process boundary
(main.cu:56-63): same arguments, CSV row, exit status and output.fs layout,
on the GPU, checked against the CPU oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

import gfsm_amd as G
from helpers import oracle_solve

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
