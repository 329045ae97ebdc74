"""gfsm_amd -- MI355X-native forest-based subgraph matching (GFSM semantics).

Python binding of ``libgfsm.so`` over its C ABI (``include/gfsm.h``).  The
compute path is the HIP library; this module only marshals arrays and handles.
Loading fails loudly when the library has not been built.

    ctx = gfsm_amd.Context()
    data = gfsm_amd.Graph.rmat(18, 8, nvlabels=16, mode=gfsm_amd.MODE_VLABEL)
    query = gfsm_amd.Graph.walk_query(data, 8, seed=1)
    m = ctx.match(data, query, digest=True)
    print(m.count, m.digest)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "GfsmError", "Graph", "Context", "Match", "Server", "lib", "library_path", "row_hash",
    "MODE_REF", "MODE_VLABEL", "ROWS", "SORTED", "DIGEST", "PROFILE", "WALK_DENSE", "WALK_SPARSE",
]

_HERE = os.path.dirname(os.path.abspath(__file__))

MODE_REF = 0
MODE_VLABEL = 1
ROWS = 1
SORTED = 2
DIGEST = 4
TUNE_PLAIN = 1
TUNE_BATCH = 2
PROFILE = 8
ORDER_COST = 16  # matching order from the data graph's relation statistics
WALK_DENSE = 0
WALK_SPARSE = 1
ERR_PENDING = 6
MAX_Q = 32
NOALLOC_SCRATCH_BYTES = 64 << 20  # GFSM_NOALLOC_SCRATCH_BYTES
NOALLOC_LEVEL_BYTES = 60 << 20    # its level part (the rest holds the scan workspace)


class GfsmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gfsm error {code}: {msg}")
        self.code = code


def read_fs(path):
    """Rows of an output.fs (little-endian u32 count, u32 width, then count
    rows of width u32 -- the words the reference's Java checker reads,
    HnFunction.java:21-38) as a (count, width) uint32 array."""
    raw = np.fromfile(str(path), dtype="<u4")
    if raw.size < 2:
        raise GfsmError(233, f"{path}: shorter than the output.fs header")
    n, w = int(raw[0]), int(raw[1])
    if raw.size != 2 + n * w:
        raise GfsmError(233, f"{path}: header says {n} rows of {w} words, file holds {raw.size - 2} words")
    return raw[2:].reshape(n, w) if w else np.zeros((n, 0), np.uint32)


def format_rows(rows, limit=10):
    """Embeddings as text, one per line: "q0->f(q0) q1->f(q1) ..." (query
    vertex -> data vertex), at most `limit` lines plus a count of the rest."""
    rows = np.asarray(rows)
    out = [" ".join(f"{c}->{int(v)}" for c, v in enumerate(r)) for r in rows[:limit]]
    if len(rows) > limit:
        out.append(f"... {len(rows) - limit} more")
    return "\n".join(out)


class MatchOpts(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("root_lo", C.c_uint32), ("root_hi", C.c_uint32),
                ("tile", C.c_uint32), ("max_rows", C.c_uint64), ("mem_budget", C.c_uint64),
                ("max_cand", C.c_uint64), ("max_depth_cand", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("q", C.c_uint32), ("launches", C.c_uint32), ("chunks", C.c_uint32),
                ("overflows", C.c_uint32), ("roots", C.c_uint64),
                ("level", C.c_uint64 * MAX_Q), ("cand", C.c_uint64 * MAX_Q),
                ("checks", C.c_uint32 * MAX_Q), ("alg_bytes", C.c_uint64),
                ("device_ms", C.c_double), ("kernel_ms", C.c_double), ("expand_ms", C.c_double),
                ("expand_launches", C.c_uint32), ("pad", C.c_uint32), ("host_ms", C.c_double)]

    def as_dict(self):
        q = self.q
        return {
            "q": q, "launches": self.launches, "chunks": self.chunks, "overflows": self.overflows,
            "roots": self.roots, "level": list(self.level[:q]), "cand": list(self.cand[:q]),
            "checks": list(self.checks[:q]), "alg_bytes": self.alg_bytes,
            "device_ms": self.device_ms, "kernel_ms": self.kernel_ms, "expand_ms": self.expand_ms,
            "expand_launches": self.expand_launches, "host_ms": self.host_ms,
        }


class Tuning(C.Structure):
    _fields_ = [("items", (C.c_uint32 * 32) * 3), ("pool_words", C.c_uint32), ("wgs_per_cu", C.c_uint32),
                ("batch_items", C.c_uint32), ("batch_pool_words", C.c_uint32), ("batch_wgs_per_cu", C.c_uint32),
                ("batch_slots", C.c_uint32), ("batch_steps", C.c_uint32), ("serve_steps", C.c_uint32),
                ("batch_step_budget", C.c_uint64)]

    def as_dict(self):
        return {"items": [list(self.items[k]) for k in range(3)], "pool_words": self.pool_words,
                "wgs_per_cu": self.wgs_per_cu, "batch_items": self.batch_items,
                "batch_pool_words": self.batch_pool_words, "batch_wgs_per_cu": self.batch_wgs_per_cu,
                "batch_slots": self.batch_slots, "batch_steps": self.batch_steps, "serve_steps": self.serve_steps,
                "batch_step_budget": self.batch_step_budget}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("waited_memory", C.c_uint64),
                ("waited_slot", C.c_uint64), ("completed", C.c_uint64)]


class ServeOpts(C.Structure):
    _fields_ = [("mem_budget", C.c_uint64), ("max_inflight", C.c_uint32), ("flags", C.c_uint32)]


_LIB = None


def library_path():
    return os.path.join(_HERE, "libgfsm.so")


def lib():
    """The loaded libgfsm.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() (make -C gfsm_amd)")
    L = C.CDLL(path)
    P, u32, u64, vp = C.POINTER, C.c_uint32, C.c_uint64, C.c_void_p
    pu32 = P(u32)
    sigs = {
        "gfsm_open": [C.c_int, u64, P(vp)],
        "gfsm_close": [vp],
        "gfsm_graph_load_tve": [C.c_char_p, u32, P(vp)],
        "gfsm_graph_from_edges": [u32, pu32, u64, pu32, pu32, pu32, u32, P(vp)],
        "gfsm_graph_write_tve": [vp, C.c_char_p],
        "gfsm_graph_save_bin": [vp, C.c_char_p],
        "gfsm_graph_load_bin": [C.c_char_p, P(vp)],
        "gfsm_graph_edges": [vp, P(u32), P(u64), P(pu32), P(pu32), P(pu32), P(pu32)],
        "gfsm_graph_layout": [vp, P(u32), P(u64), P(u64), P(pu32), P(pu32), P(pu32)],
        "gfsm_graph_info": [vp, P(u32), P(u64), P(u32), P(u32)],
        "gfsm_graph_upload": [vp, vp],
        "gfsm_graph_free": [vp],
        "gfsm_plan_query": [vp, P(u32), P(u32), pu32, pu32, pu32],
        "gfsm_query_roots": [vp, vp, pu32, P(u64)],
        "gfsm_plan_order": [vp, vp, u32, pu32],
        "gfsm_query_projection": [vp, vp, P(MatchOpts), P(C.c_double), P(C.c_double)],
        "gfsm_query_window": [vp, vp, u64, pu32, pu32, P(u64)],
        "gfsm_match": [vp, vp, vp, P(MatchOpts), P(vp)],
        "gfsm_match_noalloc": [vp, vp, vp, P(MatchOpts), P(u64), P(u64)],
        "gfsm_result_count": [vp, P(u64)],
        "gfsm_result_rows": [vp, P(pu32), P(u32), P(u64)],
        "gfsm_result_digest": [vp, P(u64)],
        "gfsm_result_stats": [vp, P(Stats)],
        "gfsm_result_status": [vp],
        "gfsm_result_write_fs": [vp, C.c_char_p],
        "gfsm_match_write_fs": [vp, vp, vp, P(MatchOpts), C.c_char_p, P(vp)],
        "gfsm_result_free": [vp],
        "gfsm_row_hash": [pu32, u32, P(u64)],
        "gfsm_match_batch": [vp, vp, P(vp), u32, P(MatchOpts), u32, P(vp)],
        "gfsm_tuning_get": [vp, P(Tuning)],
        "gfsm_tuning_set": [vp, P(Tuning)],
        "gfsm_autotune": [vp, vp, P(vp), u32, P(MatchOpts), u32, u32, P(Tuning)],
        "gfsm_server_start": [vp, vp, P(ServeOpts), P(vp)],
        "gfsm_server_submit": [vp, vp, P(MatchOpts), P(u64)],
        "gfsm_server_poll": [vp, u64, P(vp)],
        "gfsm_server_wait": [vp, u64, P(vp)],
        "gfsm_server_times": [vp, u64, P(C.c_double)],
        "gfsm_server_stop": [vp],
        "gfsm_server_stats": [vp, P(ServerStats)],
        "gfsm_selftest_scan": [vp, pu32, u64, P(u64), pu32, P(u64)],
        "gfsm_selftest_traffic": [vp, u32, u64, u32, P(u64)],
        "gfsm_gen_er": [u32, C.c_double, u32, u32, u64, u32, P(vp)],
        "gfsm_gen_rmat": [u32, u32, C.c_double, C.c_double, C.c_double, u32, C.c_double, u32, u64, u32, P(vp)],
        "gfsm_gen_grid": [u32, u32, C.c_double, C.c_double, u64, u32, P(vp)],
        "gfsm_gen_query_walk": [vp, u32, u32, u64, P(vp)],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int32
    for name in ("gfsm_close", "gfsm_graph_free", "gfsm_result_free", "gfsm_server_stop", "gfsm_row_hash"):
        getattr(L, name).restype = None
    L.gfsm_last_error.restype = C.c_char_p
    L.gfsm_last_error.argtypes = []
    L.gfsm_version.restype = C.c_char_p
    _LIB = L
    return L


def _check(rc, what=""):
    if rc != 0:
        msg = lib().gfsm_last_error().decode(errors="replace")
        raise GfsmError(rc, f"{what}: {msg}" if what else msg)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def _view(ptr, n, dtype=np.uint32):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def row_hash(row):
    a, p = _u32(row)
    h = (C.c_uint64 * 2)()
    lib().gfsm_row_hash(p, len(a), h)
    return int(h[0]), int(h[1])


class Graph:
    """A host graph in the relation-partitioned layout (uploaded on first use)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gfsm_graph_free(self._h)
            self._h = C.c_void_p()

    @classmethod
    def _make(cls, fn, *args, what=""):
        h = C.c_void_p()
        _check(fn(*args, C.byref(h)), what)
        return cls(h)

    @classmethod
    def load(cls, path, mode=MODE_REF):
        return cls._make(lib().gfsm_graph_load_tve, str(path).encode(), mode, what=f"load {path}")

    @classmethod
    def from_edges(cls, nv, src, dst, elabel, vlabel=None, mode=MODE_REF):
        s, ps = _u32(src)
        d, pd = _u32(dst)
        e, pe = _u32(elabel)
        if vlabel is None:
            pv = None
        else:
            v, pv = _u32(vlabel)
        return cls._make(lib().gfsm_graph_from_edges, int(nv), pv, len(s), ps, pd, pe, mode, what="from_edges")

    @classmethod
    def er(cls, n, avg_degree, nvlabels=1, nelabels=1, seed=0, mode=MODE_REF):
        return cls._make(lib().gfsm_gen_er, n, float(avg_degree), nvlabels, nelabels, seed, mode, what="gen_er")

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, nvlabels=1, zipf_s=1.0, nelabels=1, seed=0,
             mode=MODE_REF):
        return cls._make(lib().gfsm_gen_rmat, scale, edge_factor, a, b, c, nvlabels, float(zipf_s), nelabels,
                         seed, mode, what="gen_rmat")

    @classmethod
    def grid(cls, rows, cols, del_frac=0.0, diag_frac=0.0, seed=0, mode=MODE_REF):
        return cls._make(lib().gfsm_gen_grid, rows, cols, float(del_frac), float(diag_frac), seed, mode,
                         what="gen_grid")

    @classmethod
    def walk_query(cls, data, qsize, kind=WALK_DENSE, seed=0):
        return cls._make(lib().gfsm_gen_query_walk, data._h, qsize, kind, seed, what="gen_query_walk")

    def write(self, path):
        _check(lib().gfsm_graph_write_tve(self._h, str(path).encode()), "write")

    def save_bin(self, path):
        """Binary image of the built graph (input arrays + layout)."""
        _check(lib().gfsm_graph_save_bin(self._h, str(path).encode()), "save_bin")

    @classmethod
    def load_bin(cls, path):
        return cls._make(lib().gfsm_graph_load_bin, str(path).encode(), what=f"load_bin {path}")

    def info(self):
        nv, ne, md, mode = C.c_uint32(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        _check(lib().gfsm_graph_info(self._h, C.byref(nv), C.byref(ne), C.byref(md), C.byref(mode)))
        return {"V": nv.value, "entries": ne.value, "max_degree": md.value, "mode": mode.value}

    @property
    def V(self):
        return self.info()["V"]

    def edges(self):
        """(V, src, dst, elabel, vlabel) copies of the input entries."""
        nv, ne = C.c_uint32(), C.c_uint64()
        ps = [C.POINTER(C.c_uint32)() for _ in range(4)]
        _check(lib().gfsm_graph_edges(self._h, C.byref(nv), C.byref(ne), C.byref(ps[0]), C.byref(ps[1]),
                                      C.byref(ps[2]), C.byref(ps[3])))
        n = ne.value
        return (nv.value, _view(ps[1], n), _view(ps[2], n), _view(ps[3], n), _view(ps[0], nv.value))

    def layout(self):
        """(dir_ptr, dir as (ndir+1, 2) {relation, begin}, nbr) copies of the layout."""
        nv, nd, nn = C.c_uint32(), C.c_uint64(), C.c_uint64()
        ps = [C.POINTER(C.c_uint32)() for _ in range(3)]
        _check(lib().gfsm_graph_layout(self._h, C.byref(nv), C.byref(nd), C.byref(nn), C.byref(ps[0]),
                                       C.byref(ps[1]), C.byref(ps[2])))
        return (_view(ps[0], nv.value + 1), _view(ps[1], 2 * (nd.value + 1)).reshape(-1, 2),
                _view(ps[2], nn.value))

    def roots(self, query):
        """Ascending ids of the data vertices that can be f(root) for `query`."""
        n = C.c_uint64()
        _check(lib().gfsm_query_roots(self._h, query._h, None, C.byref(n)), "roots")
        ids = np.zeros(max(1, n.value), np.uint32)
        _check(lib().gfsm_query_roots(self._h, query._h, ids.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      C.byref(n)), "roots")
        return ids[: n.value]

    def projection(self, query, **kw):
        """(nodes, candidates) expected per depth of `query` on this data graph
        (options as for match: root_lo/root_hi, order_cost)."""
        o = _opts(**kw)
        nodes, cand = np.zeros(query.V), np.zeros(query.V)
        _check(lib().gfsm_query_projection(self._h, query._h, C.byref(o),
                                           nodes.ctypes.data_as(C.POINTER(C.c_double)),
                                           cand.ctypes.data_as(C.POINTER(C.c_double))), "projection")
        return nodes, cand

    def order(self, query, cost=False):
        """Matching order (original query ids by depth) of `query` on this data
        graph: the reference's, or with cost=True the statistics-driven one."""
        col = np.zeros(query.V, np.uint32)
        _check(lib().gfsm_plan_order(self._h, query._h, ORDER_COST if cost else 0,
                                     col.ctypes.data_as(C.POINTER(C.c_uint32))), "plan_order")
        return col

    def window(self, query, max_first):
        """(root_lo, root_hi, first): the highest-id roots of `query` whose
        depth-1 candidates total at most max_first (at least one root)."""
        lo, hi, f = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(lib().gfsm_query_window(self._h, query._h, max_first, C.byref(lo), C.byref(hi), C.byref(f)),
               "window")
        return lo.value, hi.value, f.value

    def plan(self):
        """Reference-layout plan: (col, header[3q], requirements[2*maxval*q])."""
        q, mv = C.c_uint32(), C.c_uint32()
        _check(lib().gfsm_plan_query(self._h, C.byref(q), C.byref(mv), None, None, None), "plan")
        col = np.zeros(q.value, np.uint32)
        hdr = np.zeros(3 * q.value, np.uint32)
        req = np.zeros(max(1, 2 * mv.value * q.value), np.uint32)
        P = C.POINTER(C.c_uint32)
        _check(lib().gfsm_plan_query(self._h, C.byref(q), C.byref(mv), col.ctypes.data_as(P),
                                     hdr.ctypes.data_as(P), req.ctypes.data_as(P)), "plan")
        return col, hdr, req[: 2 * mv.value * q.value]


class Match:
    """Result of one match: count, digest, optional rows and kernel statistics."""

    def __init__(self, handle):
        self._h = handle
        L = lib()
        n = C.c_uint64()
        _check(L.gfsm_result_count(handle, C.byref(n)))
        self.count = n.value
        dg = (C.c_uint64 * 2)()
        _check(L.gfsm_result_digest(handle, dg))
        self.digest = (int(dg[0]), int(dg[1]))
        self._stats = None  # read from the result on first use
        self.status = L.gfsm_result_status(handle)
        p = C.POINTER(C.c_uint32)()
        w, nr = C.c_uint32(), C.c_uint64()
        _check(L.gfsm_result_rows(handle, C.byref(p), C.byref(w), C.byref(nr)))
        self.width = w.value
        self.rows = _view(p, nr.value * w.value).reshape(nr.value, w.value) if nr.value else None

    @property
    def stats(self):
        """Kernel statistics of the match (a dict), read on first use."""
        if self._stats is None:
            st = Stats()
            _check(lib().gfsm_result_stats(self._h, C.byref(st)))
            self._stats = st.as_dict()
        return self._stats

    def write_fs(self, path):
        _check(lib().gfsm_result_write_fs(self._h, str(path).encode()), "write_fs")

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gfsm_result_free(self._h)
            self._h = C.c_void_p()


def _opts(rows=False, sorted=False, digest=False, profile=False, root_lo=0, root_hi=0, max_rows=0,
          mem_budget=0, max_cand=0, tile=0, max_depth_cand=0, order_cost=False):
    flags = (ROWS if rows else 0) | (SORTED if sorted else 0) | (DIGEST if digest else 0) | \
        (PROFILE if profile else 0) | (ORDER_COST if order_cost else 0)
    return MatchOpts(flags, root_lo, root_hi, tile, max_rows, mem_budget, max_cand, max_depth_cand)


class Context:
    """A device context: one HIP stream, its scan workspace and a level arena."""

    def __init__(self, device=0, mem_budget=0):
        self._h = C.c_void_p()
        _check(lib().gfsm_open(device, mem_budget, C.byref(self._h)), "open")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gfsm_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def selftest_traffic(self, mode, n, launches=5):
        """Calibration launches of known HBM footprint (mode 0 stream, 1
        gather); returns (bytes read, bytes written) per launch."""
        b = (C.c_uint64 * 2)()
        _check(lib().gfsm_selftest_traffic(self._h, mode, n, launches, b), "selftest_traffic")
        return int(b[0]), int(b[1])

    def selftest_scan(self, vals):
        """Chained wave64 scan on the device: (exclusive prefix u64, compacted indices, (nonzero, sum))."""
        v, pv = _u32(vals)
        n = len(v)
        pre = np.zeros(max(n, 1), np.uint64)
        idx = np.zeros(max(n, 1), np.uint32)
        tot = (C.c_uint64 * 2)()
        _check(lib().gfsm_selftest_scan(self._h, pv, n, pre.ctypes.data_as(C.POINTER(C.c_uint64)),
                                        idx.ctypes.data_as(C.POINTER(C.c_uint32)), tot), "selftest_scan")
        return pre[:n], idx[: int(tot[0])], (int(tot[0]), int(tot[1]))

    def upload(self, g: Graph):
        _check(lib().gfsm_graph_upload(self._h, g._h), "upload")

    def match_to_fs(self, data: Graph, query: Graph, path, allow_partial=False, **kw) -> Match:
        """Match and write output.fs at `path` while matching (rows streamed
        chunk by chunk unless sorted=True); the Match carries no rows."""
        o = _opts(**kw)
        r = C.c_void_p()
        rc = lib().gfsm_match_write_fs(self._h, data._h, query._h, C.byref(o), str(path).encode(), C.byref(r))
        if rc == 5 and r.value and allow_partial:
            return Match(r)
        _check(rc, "match_write_fs")
        return Match(r)

    def match(self, data: Graph, query: Graph, allow_partial=False, **kw) -> Match:
        """Match query in data.  With a work cap (max_cand) and allow_partial,
        a capped run returns its partial Match (status == 5) instead of raising."""
        o = _opts(**kw)
        r = C.c_void_p()
        rc = lib().gfsm_match(self._h, data._h, query._h, C.byref(o), C.byref(r))
        if rc == 5 and r.value:
            m = Match(r)
            if allow_partial:
                return m
            del m
        _check(rc, "match")
        return Match(r)

    def match_noalloc(self, data: Graph, query: Graph, **kw):
        """gfsm_match_noalloc: count (and digest) without allocating device
        memory; data must have been uploaded.  Returns (status, count,
        digest): status 0, or 5 with the partial values under a cap."""
        o = _opts(**kw)
        n = C.c_uint64()
        dg = (C.c_uint64 * 2)()
        rc = lib().gfsm_match_noalloc(self._h, data._h, query._h, C.byref(o), C.byref(n), dg)
        if rc not in (0, 5):
            _check(rc, "match_noalloc")
        return rc, n.value, (int(dg[0]), int(dg[1]))

    def match_batch(self, data: Graph, queries, windows=None, caps=None, depth_caps=None, **kw):
        """Batched matching: every query of the list in flight together.
        windows: optional per-query (root_lo, root_hi); caps / depth_caps:
        optional per-query max_cand / max_depth_cand.  Returns one Match per query;
        check each Match's status (5 = capped, partial count)."""
        n = len(queries)
        if windows is None and caps is None and depth_caps is None:
            os_ = (MatchOpts * 1)(_opts(**kw))
        else:
            wins = windows if windows is not None else [(0, 0)] * n
            cps = caps if caps is not None else [kw.pop("max_cand", 0)] * n
            lcs = depth_caps if depth_caps is not None else [kw.pop("max_depth_cand", 0)] * n
            os_ = (MatchOpts * n)(*[_opts(root_lo=lo, root_hi=hi, max_cand=c, max_depth_cand=l, **kw)
                                    for (lo, hi), c, l in zip(wins, cps, lcs)])
        hs = (C.c_void_p * n)(*[q._h.value for q in queries])
        rs = (C.c_void_p * n)()
        _check(lib().gfsm_match_batch(self._h, data._h, hs, n, os_, len(os_), rs), "match_batch")
        return [Match(C.c_void_p(rs[i])) for i in range(n)]

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().gfsm_tuning_get(self._h, C.byref(t)), "tuning_get")
        return t.as_dict()

    def set_tuning(self, **kw):
        """Change launch shapes: items (int for all kinds/depths, or 3x32 lists),
        pool_words, wgs_per_cu, batch_items, batch_pool_words, batch_wgs_per_cu,
        batch_slots."""
        t = Tuning()
        _check(lib().gfsm_tuning_get(self._h, C.byref(t)), "tuning_get")
        for k, v in kw.items():
            if k == "items":
                for a in range(3):
                    for d in range(32):
                        t.items[a][d] = v if isinstance(v, int) else v[a][d]
            else:
                setattr(t, k, v)
        _check(lib().gfsm_tuning_set(self._h, C.byref(t)), "tuning_set")

    def autotune(self, data: Graph, queries, windows=None, caps=None, depth_caps=None, which=3) -> dict:
        """Time the sample (per-query windows and caps optional) under each
        candidate launch shape; keep the fastest."""
        n = len(queries)
        hs = (C.c_void_p * n)(*[q._h.value for q in queries])
        wins = windows if windows is not None else [(0, 0)] * n
        cps = caps if caps is not None else [0] * n
        dcs = depth_caps if depth_caps is not None else [0] * n
        os_ = (MatchOpts * n)(*[_opts(root_lo=lo, root_hi=hi, max_cand=c, max_depth_cand=d)
                                for (lo, hi), c, d in zip(wins, cps, dcs)])
        t = Tuning()
        _check(lib().gfsm_autotune(self._h, data._h, hs, n, os_, n, which, C.byref(t)), "autotune")
        return t.as_dict()


class Server:
    """Serving loop over a resident data graph (submit / poll / wait)."""

    def __init__(self, ctx: Context, data: Graph, mem_budget=0, max_inflight=0, digest=False):
        self._h = C.c_void_p()
        self._ctx, self._data = ctx, data
        o = ServeOpts(mem_budget, max_inflight, DIGEST if digest else 0)
        _check(lib().gfsm_server_start(ctx._h, data._h, C.byref(o), C.byref(self._h)), "server_start")
        self._queries = {}

    def submit(self, query: Graph, **kw) -> int:
        """Queue a query; kw: root_lo, root_hi, max_cand, max_depth_cand, digest (see Context.match)."""
        t = C.c_uint64()
        o = _opts(**kw)
        _check(lib().gfsm_server_submit(self._h, query._h, C.byref(o), C.byref(t)), "submit")
        self._queries[t.value] = query  # keep the query alive while in flight
        return t.value

    def poll(self, ticket):
        r = C.c_void_p()
        rc = lib().gfsm_server_poll(self._h, ticket, C.byref(r))
        if rc == ERR_PENDING:
            return None
        _check(rc, "poll")
        self._queries.pop(ticket, None)
        return Match(r)

    def wait(self, ticket) -> Match:
        r = C.c_void_p()
        _check(lib().gfsm_server_wait(self._h, ticket, C.byref(r)), "wait")
        self._queries.pop(ticket, None)
        return Match(r)

    def times(self, ticket):
        t = (C.c_double * 3)()
        _check(lib().gfsm_server_times(self._h, ticket, t), "times")
        return tuple(t)

    def stats(self):
        """Admission counters: submitted, admitted, waited_memory (held back
        for projected memory while a slot was free), waited_slot, completed."""
        st = ServerStats()
        _check(lib().gfsm_server_stats(self._h, C.byref(st)), "server_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def stop(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gfsm_server_stop(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.stop()
