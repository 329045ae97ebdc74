"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only the test suite,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import it;
the product (``gfsm_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MODE_REF = 0
MODE_VLABEL = 1


class _Edges(C.Structure):
    _fields_ = [("V", C.c_uint32), ("E", C.c_uint64), ("vlabel", C.POINTER(C.c_uint32)),
                ("src", C.POINTER(C.c_uint32)), ("dst", C.POINTER(C.c_uint32)),
                ("elabel", C.POINTER(C.c_uint32))]


class _CCSR(C.Structure):
    _fields_ = [("count", C.c_uint32), ("size", C.c_uint64), ("tgt", C.POINTER(C.c_uint32)),
                ("rel", C.POINTER(C.c_uint32)), ("verticeLocs", C.POINTER(C.c_uint32)),
                ("invVerticeLocs", C.POINTER(C.c_uint32)), ("maxLen", C.c_uint32),
                ("staggers", C.POINTER(C.c_uint64)), ("stagOffset", C.POINTER(C.c_uint64)),
                ("stagNorm", C.POINTER(C.c_uint64)), ("rtgt", C.POINTER(C.c_uint32)),
                ("rrel", C.POINTER(C.c_uint32)), ("dup", C.POINTER(C.c_uint8))]


class _Plan(C.Structure):
    _fields_ = [("q", C.c_uint32), ("maxValency", C.c_uint32),
                ("requirements", C.POINTER(C.c_uint32)), ("header", C.POINTER(C.c_uint32)),
                ("order", C.POINTER(C.c_uint32)), ("col", C.POINTER(C.c_uint32))]


class _Result(C.Structure):
    _fields_ = [("count", C.c_uint64), ("digest", C.c_uint64 * 2), ("width", C.c_uint32),
                ("nrows", C.c_uint64), ("rows", C.POINTER(C.c_uint32)),
                ("level", C.POINTER(C.c_uint64))]


class _Canon(C.Structure):
    _fields_ = [("size", C.c_uint64), ("cid", C.POINTER(C.c_uint32)), ("coff", C.POINTER(C.c_uint32)),
                ("crel", C.POINTER(C.c_uint32))]


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: oracle error {code}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_load_tve.argtypes = [C.c_char_p, P(P(_Edges))]
        L.orc_from_arrays.argtypes = [C.c_uint32, P(C.c_uint32), C.c_uint64, P(C.c_uint32),
                                      P(C.c_uint32), P(C.c_uint32), P(P(_Edges))]
        L.orc_edges_free.argtypes = [P(_Edges)]
        L.orc_write_tve.argtypes = [P(_Edges), C.c_char_p]
        L.orc_ccsr_build.argtypes = [P(_Edges), C.c_uint32, P(P(_CCSR))]
        L.orc_ccsr_free.argtypes = [P(_CCSR)]
        L.orc_ccsr_packed.argtypes = [P(_CCSR), P(C.c_uint32)]
        L.orc_offset_to_id.argtypes = [P(_CCSR), C.c_uint32]
        L.orc_offset_to_id.restype = C.c_uint32
        L.orc_plan_build.argtypes = [P(_CCSR), P(P(_Plan))]
        L.orc_plan_free.argtypes = [P(_Plan)]
        L.orc_match.argtypes = [P(_CCSR), P(_CCSR), P(_Plan), C.c_uint32, C.c_uint32, C.c_int,
                                C.c_uint64, P(P(_Result))]
        L.orc_match_capped.argtypes = [P(_CCSR), P(_CCSR), P(_Plan), C.c_uint32, C.c_uint32, C.c_int,
                                       C.c_uint64, C.c_uint64, P(P(_Result))]
        L.orc_result_free.argtypes = [P(_Result)]
        L.orc_canon_build.argtypes = [P(_CCSR), P(P(_Canon))]
        L.orc_canon_free.argtypes = [P(_Canon)]
        L.orc_match_canon.argtypes = [P(_CCSR), P(_Canon), P(_CCSR), P(_Plan), C.c_uint32, C.c_uint32,
                                      C.c_uint64, P(P(_Result)), P(C.c_int), P(C.c_uint64)]
        L.orc_solve.argtypes = [P(_Edges), P(_Edges), C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                C.c_uint64, P(P(_Result))]
        L.orc_check_sound.argtypes = [P(_Edges), P(_Edges), P(C.c_uint32), C.c_uint64, C.c_uint32]
        L.orc_check_sound.restype = C.c_uint64
        L.orc_row_hash.argtypes = [P(C.c_uint32), C.c_uint32, P(C.c_uint64)]
        L.orc_fs_write.argtypes = [C.c_char_p, P(C.c_uint32), C.c_uint64, C.c_uint32]
        L.orc_fs_read.argtypes = [C.c_char_p, P(P(C.c_uint32)), P(C.c_uint64), P(C.c_uint32)]
        L.orc_free.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def _u32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Edges:
    """A graph as its t/v/e file states it (directed entries in file order)."""

    def __init__(self, ptr):
        self._p = ptr

    @classmethod
    def load(cls, path):
        p = C.POINTER(_Edges)()
        rc = lib().orc_load_tve(str(path).encode(), C.byref(p))
        if rc:
            raise OracleError(rc, f"load {path}")
        return cls(p)

    @classmethod
    def from_arrays(cls, V, src, dst, elabel, vlabel=None):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        elabel = np.ascontiguousarray(elabel, dtype=np.uint32)
        vl = None if vlabel is None else np.ascontiguousarray(vlabel, dtype=np.uint32)
        p = C.POINTER(_Edges)()
        rc = lib().orc_from_arrays(int(V), None if vl is None else _u32p(vl), len(src), _u32p(src),
                                   _u32p(dst), _u32p(elabel), C.byref(p))
        if rc:
            raise OracleError(rc, "from_arrays")
        return cls(p)

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_edges_free(self._p)
            self._p = None

    @property
    def V(self):
        return self._p.contents.V

    @property
    def E(self):
        return self._p.contents.E

    def arrays(self):
        c = self._p.contents
        return (_arr(c.src, c.E, np.uint32), _arr(c.dst, c.E, np.uint32),
                _arr(c.elabel, c.E, np.uint32), _arr(c.vlabel, c.V, np.uint32))

    def write(self, path):
        rc = lib().orc_write_tve(self._p, str(path).encode())
        if rc:
            raise OracleError(rc, f"write {path}")


class CCSR:
    def __init__(self, edges: Edges, mode=MODE_REF):
        self._edges = edges
        p = C.POINTER(_CCSR)()
        rc = lib().orc_ccsr_build(edges._p, mode, C.byref(p))
        if rc:
            raise OracleError(rc, "ccsr_build")
        self._p = p

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_ccsr_free(self._p)
            self._p = None

    def fields(self):
        c = self._p.contents
        ml = c.maxLen
        return {
            "count": c.count, "size": c.size, "maxLen": ml,
            "tgt": _arr(c.tgt, c.size, np.uint32), "rel": _arr(c.rel, c.size, np.uint32),
            "verticeLocs": _arr(c.verticeLocs, c.count, np.uint32),
            "invVerticeLocs": _arr(c.invVerticeLocs, c.count, np.uint32),
            "staggers": _arr(c.staggers, ml, np.uint64),
            "stagOffset": _arr(c.stagOffset, ml, np.uint64),
            "stagNorm": _arr(c.stagNorm, ml, np.uint64),
        }

    def packed(self):
        out = np.zeros(max(1, self._p.contents.size), dtype=np.uint32)
        rc = lib().orc_ccsr_packed(self._p, _u32p(out))
        if rc:
            raise OracleError(rc, "packed")
        return out[: self._p.contents.size]

    def offset_to_id(self, off):
        return lib().orc_offset_to_id(self._p, int(off))


class Plan:
    def __init__(self, qccsr: CCSR):
        p = C.POINTER(_Plan)()
        rc = lib().orc_plan_build(qccsr._p, C.byref(p))
        if rc:
            raise OracleError(rc, "plan_build")
        self._p = p
        self._q = qccsr

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_plan_free(self._p)
            self._p = None

    def fields(self):
        c = self._p.contents
        q, mv = c.q, c.maxValency
        return {"q": q, "maxValency": mv,
                "requirements": _arr(c.requirements, 2 * mv * q, np.uint32),
                "header": _arr(c.header, 3 * q, np.uint32),
                "order": _arr(c.order, q, np.uint32), "col": _arr(c.col, q, np.uint32)}


class Result:
    def __init__(self, count, digest, width, rows, level):
        self.count = count
        self.digest = digest
        self.width = width
        self.rows = rows
        self.level = level


def _take_result(p):
    c = p.contents
    rows = None
    if c.rows:
        rows = _arr(c.rows, c.nrows * c.width, np.uint32).reshape(-1, c.width) if c.nrows else \
            np.zeros((0, c.width), dtype=np.uint32)
    r = Result(int(c.count), (int(c.digest[0]), int(c.digest[1])), int(c.width), rows,
               _arr(c.level, c.width, np.uint64))
    lib().orc_result_free(p)
    return r


def solve(data: Edges, query: Edges, mode=MODE_REF, rows=False, root_lo=0, root_hi=0xFFFFFFFF,
          threads=0, max_rows=1 << 31):
    """Count (and optionally list, sorted) the embeddings of query in data."""
    p = C.POINTER(_Result)()
    rc = lib().orc_solve(data._p, query._p, mode, root_lo, root_hi, threads,
                         max_rows if rows else 0, C.byref(p))
    if rc:
        raise OracleError(rc, "solve")
    return _take_result(p)


def match(data: CCSR, query: CCSR, plan: Plan, rows=False, root_lo=0, root_hi=0xFFFFFFFF,
          threads=0, max_rows=1 << 31):
    p = C.POINTER(_Result)()
    rc = lib().orc_match(data._p, query._p, plan._p, root_lo, root_hi, threads,
                         max_rows if rows else 0, C.byref(p))
    if rc:
        raise OracleError(rc, "match")
    return _take_result(p)


def match_capped(data: CCSR, query: CCSR, plan: Plan, max_leaves, root_lo=0, root_hi=0xFFFFFFFF, threads=0):
    """Enumerate until about max_leaves embeddings are found (timing samples)."""
    p = C.POINTER(_Result)()
    rc = lib().orc_match_capped(data._p, query._p, plan._p, root_lo, root_hi, threads, 0, max_leaves,
                                C.byref(p))
    if rc:
        raise OracleError(rc, "match_capped")
    return _take_result(p)


class Canon:
    """Each data row's entries ordered by (relation, original target id): the
    canonical enumeration order of match_canon, built once per data graph."""

    def __init__(self, data: CCSR):
        p = C.POINTER(_Canon)()
        rc = lib().orc_canon_build(data._p, C.byref(p))
        if rc:
            raise OracleError(rc, "canon_build")
        self._p = p
        self._data = data

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_canon_free(self._p)
            self._p = None


def match_canon(data: CCSR, canon: Canon, query: CCSR, plan: Plan, max_depth_cand=0, root_lo=0,
                root_hi=0xFFFFFFFF):
    """Sequential enumeration in canonical order under the deterministic
    per-depth cap (0 = none).  Returns (Result, capped, cand per depth)."""
    p = C.POINTER(_Result)()
    capped = C.c_int()
    q = plan._p.contents.q
    cand = (C.c_uint64 * max(q, 1))()
    rc = lib().orc_match_canon(data._p, canon._p, query._p, plan._p, root_lo, root_hi, max_depth_cand,
                               C.byref(p), C.byref(capped), cand)
    if rc:
        raise OracleError(rc, "match_canon")
    return _take_result(p), bool(capped.value), [int(cand[i]) for i in range(q)]


def check_sound(query: Edges, data: Edges, rows: np.ndarray):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    n, w = rows.shape
    return int(lib().orc_check_sound(query._p, data._p, _u32p(rows), n, w))


def row_hash(row):
    row = np.ascontiguousarray(row, dtype=np.uint32)
    h = (C.c_uint64 * 2)()
    lib().orc_row_hash(_u32p(row), len(row), h)
    return int(h[0]), int(h[1])


def fs_write(path, rows):
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    rc = lib().orc_fs_write(str(path).encode(), _u32p(rows), rows.shape[0], rows.shape[1])
    if rc:
        raise OracleError(rc, "fs_write")


def fs_read(path):
    p = C.POINTER(C.c_uint32)()
    n = C.c_uint64()
    w = C.c_uint32()
    rc = lib().orc_fs_read(str(path).encode(), C.byref(p), C.byref(n), C.byref(w))
    if rc:
        raise OracleError(rc, "fs_read")
    rows = _arr(p, n.value * w.value, np.uint32).reshape(n.value, w.value)
    lib().orc_free(p)
    return rows
