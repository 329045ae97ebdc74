"""Shared helpers of the test suite: convert product graphs to oracle inputs
and compare a GPU match with the oracle."""
import numpy as np

import gfsm_amd as G
import oracle as O


def to_oracle(g: "G.Graph"):
    V, s, d, l, vl = g.edges()
    return O.Edges.from_arrays(V, s, d, l, vl)


def oracle_solve(data, query, mode, rows=True, **kw):
    return O.solve(to_oracle(data), to_oracle(query), mode=mode, rows=rows, **kw)


def assert_same(m, ref, rows=True):
    assert m.count == ref.count, (m.count, ref.count)
    if m.digest != (0, 0) or ref.count == 0:
        assert m.digest == ref.digest, (m.digest, ref.digest)
    if rows:
        got = m.rows if m.rows is not None else np.zeros((0, ref.width), np.uint32)
        assert got.shape == ref.rows.shape, (got.shape, ref.rows.shape)
        assert np.array_equal(got, ref.rows)
