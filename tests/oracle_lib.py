"""ctypes access to the CPU oracle (oracle/liboracle.so) and, when built, the
reference's own CPU validators (oracle/_ref/libcgpref.so).  Test-only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libcgpref.so")
INF = 0xFFFFFFFF

_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def _load_oracle():
    lib = C.CDLL(ORACLE)
    lib.orc_bfs.argtypes = [C.c_int32, _i32p, _i32p, C.c_int32, _u32p]
    lib.orc_sssp.argtypes = [C.c_int32, _i32p, _i32p, _u32p, C.c_int32, _u32p]
    lib.orc_wcc_lp.argtypes = [C.c_int32, _i32p, _i32p, _u32p, C.c_int32]
    lib.orc_wcc_lp.restype = C.c_int32
    lib.orc_wcc_uf.argtypes = [C.c_int32, _i32p, _i32p, _u32p]
    lib.orc_bellman_ford.argtypes = [C.c_int32, _i32p, _i32p, _u32p, C.c_int32, _u32p]
    lib.orc_reached_edges.argtypes = [C.c_int32, _i32p, _u32p]
    lib.orc_reached_edges.restype = C.c_uint64
    lib.orc_run_queries.argtypes = [C.c_int32, _i32p, _i32p, _u32p, _i32p, _i32p, C.c_int32,
                                    C.c_int32, _u32p, _u64p]
    lib.orc_run_queries.restype = C.c_double
    lib.orc_check_sssp.argtypes = [C.c_int32, _i32p, _i32p, _u32p, C.c_int32, _u32p]
    lib.orc_check_sssp.restype = C.c_int
    lib.orc_check_wcc.argtypes = [C.c_int32, _i32p, _i32p, _u32p]
    lib.orc_check_wcc.restype = C.c_int
    lib.orc_check_rows.argtypes = [C.c_int32, _i32p, _i32p, _u32p, _i32p, _i32p, C.c_int32, _u32p,
                                   C.c_int64, C.c_int32, _i32p]
    lib.orc_check_rows.restype = C.c_int32
    return lib


orc = _load_oracle()


def _load_ref():
    if not os.path.exists(REF):
        return None
    lib = C.CDLL(REF)
    lib.ref_puer_bfs.argtypes = [C.c_int32, _i32p, _i32p, C.c_int32, _u32p]
    lib.ref_puer_sssp.argtypes = [C.c_int32, _i32p, _i32p, _u32p, C.c_int32, _u32p]
    lib.ref_puer_wcc.argtypes = [C.c_int32, _i32p, _i32p, C.c_int32, _u32p]
    lib.ref_cgp_bfs.argtypes = [C.c_int32, _i32p, _i32p, C.c_int32, _u32p]
    lib.ref_cgp_sssp.argtypes = [C.c_int32, _i32p, _i32p, _u32p, C.c_int32, _u32p]
    if hasattr(lib, "ref_run_queries"):  # absent from a build older than its source
        lib.ref_run_queries.argtypes = [C.c_int32, _i32p, _i32p, _u32p, _i32p, _i32p, C.c_int32, C.c_int32,
                                        _u64p]
        lib.ref_run_queries.restype = C.c_double
    return lib


ref = _load_ref()


class G:
    """Plain CSR arrays as contiguous numpy (int32 / uint32)."""

    def __init__(self, offsets, indices, weights=None):
        self.off = np.ascontiguousarray(offsets, dtype=np.int32)
        self.idx = np.ascontiguousarray(indices, dtype=np.int32)
        self.wt = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        self.V = len(self.off) - 1
        self.E = len(self.idx)

    @classmethod
    def of(cls, h):
        return cls(h.offsets, h.indices, h.weights)

    def p(self):
        wt = self.wt.ctypes.data_as(_u32p) if self.wt is not None else None
        return self.off.ctypes.data_as(_i32p), self.idx.ctypes.data_as(_i32p), wt


def _out(V):
    o = np.empty(V, dtype=np.uint32)
    return o, o.ctypes.data_as(_u32p)


def bfs(g, s):
    o, op = _out(g.V)
    off, idx, _ = g.p()
    orc.orc_bfs(g.V, off, idx, int(s), op)
    return o


def sssp(g, s, unit=False):
    o, op = _out(g.V)
    off, idx, wt = g.p()
    orc.orc_sssp(g.V, off, idx, None if unit else wt, int(s), op)
    return o


def wcc_uf(g):
    o, op = _out(g.V)
    off, idx, _ = g.p()
    orc.orc_wcc_uf(g.V, off, idx, op)
    return o


def wcc_lp(g, max_iterations=1 << 30):
    o, op = _out(g.V)
    off, idx, _ = g.p()
    orc.orc_wcc_lp(g.V, off, idx, op, int(max_iterations))
    return o


def bellman_ford(g, s, unit=False):
    o, op = _out(g.V)
    off, idx, wt = g.p()
    orc.orc_bellman_ford(g.V, off, idx, None if unit else wt, int(s), op)
    return o


def reached_edges(g, vals):
    v = np.ascontiguousarray(vals, dtype=np.uint32)
    off, _, _ = g.p()
    return int(orc.orc_reached_edges(g.V, off, v.ctypes.data_as(_u32p)))


def expected(g, kind, source):
    """Oracle answer for one query: kind 0 BFS, 1 SSSP, 2 WCC (min-id labels)."""
    if kind == 0:
        return bfs(g, source)
    if kind == 1:
        return sssp(g, source)
    return wcc_uf(g)


def run_queries(g, kinds, sources, threads, keep=False):
    """CPU baseline runner: returns (seconds, values or None, m_q)."""
    n = len(kinds)
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    s = np.ascontiguousarray(sources, dtype=np.int32)
    out = np.empty((n, g.V), dtype=np.uint32) if keep else None
    m = np.zeros(n, dtype=np.uint64)
    off, idx, wt = g.p()
    secs = orc.orc_run_queries(g.V, off, idx, wt, k.ctypes.data_as(_i32p), s.ctypes.data_as(_i32p), n,
                               int(threads), out.ctypes.data_as(_u32p) if keep else None,
                               m.ctypes.data_as(_u64p))
    return secs, out, m


def ref_run_queries(g, kinds, sources, threads):
    """Times the reference's own CPU validators (oracle/_ref) on a batch:
    returns (seconds, checksum = count of reached (query, vertex) pairs)."""
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    s = np.ascontiguousarray(sources, dtype=np.int32)
    ck = np.zeros(1, dtype=np.uint64)
    off, idx, wt = g.p()
    secs = ref.ref_run_queries(g.V, off, idx, wt, k.ctypes.data_as(_i32p), s.ctypes.data_as(_i32p), len(k),
                               int(threads), ck.ctypes.data_as(_u64p))
    return secs, int(ck[0])


def ref_call(name, g, *args):
    o, op = _out(g.V)
    off, idx, wt = g.p()
    fn = getattr(ref, name)
    if name in ("ref_puer_sssp", "ref_cgp_sssp"):
        fn(g.V, off, idx, wt, int(args[0]), op)
    elif name == "ref_puer_wcc":
        fn(g.V, off, idx, int(args[0]), op)
    else:
        fn(g.V, off, idx, int(args[0]), op)
    return o


# Error codes of the C checkers (0 = exact answer).
CHECK_ERRORS = {0: None, 1: "source not 0", 2: "edge not relaxed", 3: "finite vertex without tight in-edge",
                4: "zero distance off the source", 5: "edge crosses labels", 6: "label not the class minimum",
                7: "label class not connected", 8: "unknown kind"}


def host_threads():
    """Worker threads for the checkers: the host's cores, at most 16 (a GPU
    box's CPU share)."""
    return max(1, min(16, os.cpu_count() or 1))


def check_sssp(g, source, d, unit=False):
    """Error text, or None when d is the exact distance (unit: depth) vector."""
    v = np.ascontiguousarray(d, dtype=np.uint32)
    off, idx, wt = g.p()
    return CHECK_ERRORS[orc.orc_check_sssp(g.V, off, idx, None if unit else wt, int(source),
                                            v.ctypes.data_as(_u32p))]


def check_wcc(g, labels):
    v = np.ascontiguousarray(labels, dtype=np.uint32)
    off, idx, _ = g.p()
    return CHECK_ERRORS[orc.orc_check_wcc(g.V, off, idx, v.ctypes.data_as(_u32p))]


def check_rows(g, kinds, sources, rows, threads=None):
    """Checks every row of a [n, V] uint32 result (kinds 0 BFS / 1 SSSP / 2 WCC)
    on host threads; returns the list of (row, error text) that failed."""
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    n = len(kinds)
    assert rows.shape == (n, g.V)
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    s = np.ascontiguousarray(sources, dtype=np.int32)
    st = np.zeros(n, dtype=np.int32)
    off, idx, wt = g.p()
    orc.orc_check_rows(g.V, off, idx, wt, k.ctypes.data_as(_i32p), s.ctypes.data_as(_i32p), n,
                       rows.ctypes.data_as(_u32p), g.V, int(threads or host_threads()),
                       st.ctypes.data_as(_i32p))
    return [(i, CHECK_ERRORS[int(st[i])]) for i in range(n) if st[i] != 0]
