"""cgp_amd — MI355X-native concurrent multi-query graph engine (Python host mirror).

A thin layer over the C-ABI of ``include/cgp/cgp.h``; every computation runs
in ``libcgp_amd.so`` (HIP kernels for gfx950).  The names mirror the
reference's operator API (paths relative to the reference root):

* ``HostCSR``  — graph load to CSR: generators plus the ``.gr`` / Matrix
  Market / CSR-directory loaders (cgp/examples/algorithms/bfs/cgp_bfs.cu:214-275,
  puercgp/examples/mtx_loader.hxx:99-260).
* ``Graph``    — device-resident CSR (puercgp/include/puercgp/backend/csr_graph.hxx:9-75).
* ``Engine``   — per-graph batch engine; ``Engine.run_batch`` replaces
  ``run_heterogeneous`` (puercgp/include/puercgp/engine/hybrid_engine.hxx:1144-1150),
  ``Engine.bfs`` / ``Engine.sssp`` replace ``gunrock::cgp_bfs::run`` /
  ``cgp_sssp::run`` (cgp/include/gunrock/algorithms/cgp_bfs.hxx:1050-1056,
  cgp_sssp.hxx:973-979) and ``Engine.wcc`` the WCC policy
  (puercgp/include/puercgp/algorithms/wcc.hxx:20-42).
* ``Engine.serve_*`` — continuous batching (replenish_engine.hxx:559-573).

Results are uint32 arrays shaped ``[n_queries, V]`` (query-major, as cgp's
``result_t::distances``), ``INF = 0xFFFFFFFF`` for unreachable.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BFS, CGP_EBUSY, CGP_ECAPACITY, CGP_EDEVICE, CGP_EINVAL, CGP_EIO, CGP_ENOMEM,
                   CGP_OK, HYBRID, INF, PULL, PUSH, SSSP, WCC, CgpError, cgp_completion,
                   cgp_engine_opts, cgp_iteration_record, cgp_query, cgp_run_stats, cgp_serve_opts,
                   check, lib)

__all__ = [
    "BFS", "SSSP", "WCC", "PUSH", "PULL", "HYBRID", "INF", "CgpError", "HostCSR", "Graph",
    "Engine", "DeviceBuffer", "device_count", "device_info", "make_queries", "poisson_arrivals",
    "CGP_OK", "CGP_EINVAL", "CGP_ENOMEM", "CGP_EDEVICE", "CGP_EBUSY", "CGP_ECAPACITY", "CGP_EIO",
]


def device_count():
    n = C.c_int32(0)
    check(lib.cgp_device_count(C.byref(n)))
    return n.value


def device_info():
    buf = C.create_string_buffer(256)
    cus = C.c_int32(0)
    mem = C.c_uint64(0)
    check(lib.cgp_device_info(buf, 256, C.byref(cus), C.byref(mem)))
    return {"name": buf.value.decode(), "compute_units": cus.value, "hbm_bytes": mem.value}


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _u32(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def make_queries(kinds, sources):
    """Array of cgp_query from parallel kind / source sequences."""
    kinds = list(kinds)
    sources = list(sources)
    if len(kinds) != len(sources):
        raise ValueError("kinds and sources differ in length")
    arr = (cgp_query * len(kinds))()
    for i, (k, s) in enumerate(zip(kinds, sources)):
        arr[i].kind = int(k)
        arr[i].source = int(s)
    return arr


def poisson_arrivals(n, rate_per_s, seed):
    out = np.zeros(n, dtype=np.float64)
    check(lib.cgp_poisson_arrivals(int(n), float(rate_per_s), int(seed),
                                   out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


class HostCSR:
    """Host CSR graph owned by the library (generators and loaders)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        V = C.c_int32()
        E = C.c_int64()
        off = C.POINTER(C.c_int32)()
        idx = C.POINTER(C.c_int32)()
        wt = C.POINTER(C.c_uint32)()
        check(lib.cgp_host_csr_view(self._h, C.byref(V), C.byref(E), C.byref(off), C.byref(idx),
                                    C.byref(wt)))
        self.V, self.E = V.value, E.value
        self.offsets = np.ctypeslib.as_array(off, shape=(self.V + 1,)) if self.V >= 0 else None
        self.indices = (np.ctypeslib.as_array(idx, shape=(self.E,)) if self.E > 0
                        else np.zeros(0, dtype=np.int32))
        self.weights = (np.ctypeslib.as_array(wt, shape=(self.E,))
                        if (wt and self.E > 0) else None)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.cgp_host_csr_destroy(h)
            self._h = None

    @staticmethod
    def _make(fn, *args):
        out = C.c_void_p()
        check(fn(*args, C.byref(out)))
        return HostCSR(out.value)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, wmin=0, wmax=0):
        return cls._make(lib.cgp_gen_rmat, int(scale), int(edge_factor), float(a), float(b),
                         float(c), int(seed), int(wmin), int(wmax))

    @classmethod
    def grid(cls, rows, cols, p_extra=0.0, seed=1, wmin=0, wmax=0):
        return cls._make(lib.cgp_gen_grid, int(rows), int(cols), float(p_extra), int(seed),
                         int(wmin), int(wmax))

    @classmethod
    def from_coo(cls, n_vertices, src, dst, weights=None):
        s, sp = _i32(src)
        d, dp = _i32(dst)
        w, wp = _u32(weights)
        if len(s) != len(d) or (w is not None and len(w) != len(s)):
            raise ValueError("COO arrays differ in length")
        return cls._make(lib.cgp_host_csr_from_coo, int(n_vertices), len(s), sp, dp, wp)

    @classmethod
    def load_gr(cls, path, unit_weights=False):
        """Galois .gr v1; unit_weights ignores stored weights (the reference's
        SSSP driver reads .gr graphs this way)."""
        return cls._make(lib.cgp_load_graph, str(path).encode(), _lib.FORMAT_GR,
                         _lib.LOAD_UNIT_WEIGHTS if unit_weights else 0)

    @classmethod
    def load_matrix_market(cls, path, unit_weights=False, round_weights=False):
        """Matrix Market coordinate file; values must be integers in [0, 2^32-1]
        unless round_weights rounds them; unit_weights ignores them."""
        flags = (_lib.LOAD_UNIT_WEIGHTS if unit_weights else 0) | (_lib.LOAD_ROUND_WEIGHTS if round_weights else 0)
        return cls._make(lib.cgp_load_graph, str(path).encode(), _lib.FORMAT_MATRIX_MARKET, flags)

    @classmethod
    def load_csr_directory(cls, path, unit_weights=False):
        return cls._make(lib.cgp_load_graph, str(path).encode(), _lib.FORMAT_CSR_DIRECTORY,
                         _lib.LOAD_UNIT_WEIGHTS if unit_weights else 0)

    def degrees(self):
        return np.diff(self.offsets.astype(np.int64))

    def sample_sources(self, n, seed):
        out = np.zeros(n, dtype=np.int32)
        check(lib.cgp_sample_sources(self._h, int(n), int(seed),
                                     out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out


class Graph:
    """Device-resident CSR (HBM)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.V = lib.cgp_graph_num_vertices(self._h)
        self.E = lib.cgp_graph_num_edges(self._h)
        self.symmetric = bool(lib.cgp_graph_is_symmetric(self._h))

    @classmethod
    def from_host(cls, h, renumber=None):
        """Device graph of a host CSR.  renumber: None (the library decides:
        large skewed graphs are stored in descending-degree order), True or
        False; results are always in the caller's numbering."""
        out = C.c_void_p()
        check(lib.cgp_graph_from_host_ex(h._h, _renumber(renumber), C.byref(out)))
        return cls(out.value)

    @classmethod
    def from_csr(cls, offsets, indices, weights=None, renumber=None):
        o, op = _i32(offsets)
        i, ip = _i32(indices)
        w, wp = _u32(weights)
        out = C.c_void_p()
        check(lib.cgp_graph_create_csr_ex(len(o) - 1, len(i), op, ip, wp, _renumber(renumber), C.byref(out)))
        return cls(out.value)

    @classmethod
    def from_device(cls, n_vertices, n_edges, offsets, indices, weights=None):
        """Non-owning graph over device arrays (DeviceBuffer or view objects
        with a ``ptr``); they must outlive the graph."""
        out = C.c_void_p()
        check(lib.cgp_graph_create_csr_device(int(n_vertices), int(n_edges), offsets.ptr, indices.ptr,
                                              weights.ptr if weights is not None else None, C.byref(out)))
        g = cls(out.value)
        g._borrowed = (offsets, indices, weights)  # keep the arrays alive
        return g

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.cgp_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _renumber(r):
    return _lib.RENUMBER_AUTO if r is None else (_lib.RENUMBER_ON if r else _lib.RENUMBER_OFF)


class DeviceBuffer:
    """Plain device allocation (for device-resident results)."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(lib.cgp_device_alloc(self.nbytes, C.byref(self.ptr)))

    @classmethod
    def from_array(cls, a):
        """A device copy of a contiguous numpy array."""
        a = np.ascontiguousarray(a)
        b = cls(max(1, a.nbytes))
        if a.nbytes:
            check(lib.cgp_copy_to_device(b.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return b

    def to_host(self, dtype=np.uint32, shape=None):
        a = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        check(lib.cgp_copy_to_host(a.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return a.reshape(shape) if shape is not None else a

    def copy_from(self, src, nbytes=None):
        """Device-to-device copy of nbytes (default: all of src) from another
        DeviceBuffer."""
        n = src.nbytes if nbytes is None else int(nbytes)
        if n > self.nbytes or n > src.nbytes:
            raise ValueError("copy larger than a buffer")
        check(lib.cgp_copy_device(self.ptr, src.ptr, n))

    def view(self, offset, nbytes):
        """A borrowed window [offset, offset+nbytes) of this buffer."""
        if offset < 0 or offset + nbytes > self.nbytes:
            raise ValueError("view out of range")
        return _View(self, offset, nbytes)

    def close(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            lib.cgp_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


class _View:
    def __init__(self, base, offset, nbytes):
        self.base = base
        self.ptr = C.c_void_p(base.ptr.value + offset)
        self.nbytes = nbytes


def default_opts(**kw):
    o = cgp_engine_opts()
    lib.cgp_engine_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError("unknown engine option %r" % k)
        setattr(o, k, v)
    return o


class Engine:
    """Batch engine over one device graph (one HIP stream)."""

    def __init__(self, graph, **opts):
        self.graph = graph
        o = default_opts(**opts)
        self._h = C.c_void_p()
        check(lib.cgp_engine_create(graph._h, C.byref(o), C.byref(self._h)))

    @property
    def opts(self):
        o = cgp_engine_opts()
        check(lib.cgp_engine_get_opts(self._h, C.byref(o)))
        return {name: getattr(o, name) for name, _ in o._fields_}

    def set_opts(self, **kw):
        """Options for the following runs (cgp_engine_set_opts): fields not
        named keep the engine's current values."""
        o = cgp_engine_opts()
        check(lib.cgp_engine_get_opts(self._h, C.byref(o)))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise AttributeError("unknown engine option %r" % k)
            setattr(o, k, v)
        check(lib.cgp_engine_set_opts(self._h, C.byref(o)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.cgp_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run_batch(self, kinds, sources, out=None):
        """Runs a mixed batch.  Returns (values[n, V] on the host, stats dict), or
        (None, stats) when ``out`` is a DeviceBuffer that receives the results."""
        q = make_queries(kinds, sources)
        n = len(q)
        st = cgp_run_stats()
        if out is None:
            res = np.empty((n, self.graph.V), dtype=np.uint32)
            check(lib.cgp_run_batch(self._h, q, n, res.ctypes.data_as(C.c_void_p), 0, C.byref(st)))
            return res, st.as_dict()
        if out.nbytes < n * self.graph.V * 4:
            raise ValueError("device output buffer too small")
        check(lib.cgp_run_batch(self._h, q, n, out.ptr, 1, C.byref(st)))
        return None, st.as_dict()

    def run_batch_static(self, kinds, sources, out):
        """Allocation-free run (cgp_run_batch_static): every per-run device
        buffer comes from the library's fixed device array; ``out`` is a
        DeviceBuffer (or view) that receives the n x V results."""
        q = make_queries(kinds, sources)
        n = len(q)
        if out.nbytes < n * self.graph.V * 4:
            raise ValueError("device output buffer too small")
        st = cgp_run_stats()
        check(lib.cgp_run_batch_static(self._h, q, n, out.ptr, C.byref(st)))
        return st.as_dict()

    def last_run_report_text(self, fmt="json", dataset=""):
        """The last batch run's report (cgp_last_run_report_text) as a string:
        fmt "json" or "csv"."""
        f = {"json": _lib.REPORT_JSON, "csv": _lib.REPORT_CSV}[fmt]
        n = C.c_uint64(0)
        check(lib.cgp_last_run_report_text(self._h, f, dataset.encode(), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib.cgp_last_run_report_text(self._h, f, dataset.encode(), buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def last_run_report(self):
        """Per-iteration records of the last run_batch (list of dicts)."""
        n = C.c_int32(0)
        check(lib.cgp_last_run_report(self._h, None, 0, C.byref(n)))
        if n.value == 0:
            return []
        recs = (cgp_iteration_record * n.value)()
        check(lib.cgp_last_run_report(self._h, recs, n.value, C.byref(n)))
        return [recs[i].as_dict() for i in range(n.value)]

    def bfs(self, sources):
        return self.run_batch([BFS] * len(sources), sources)[0]

    def sssp(self, sources):
        return self.run_batch([SSSP] * len(sources), sources)[0]

    def wcc(self, n=1):
        return self.run_batch([WCC] * n, [0] * n)[0]

    def autotune(self, kinds, sources):
        q = make_queries(kinds, sources)
        o = cgp_engine_opts()
        check(lib.cgp_engine_autotune(self._h, q, len(q), C.byref(o)))
        return {name: getattr(o, name) for name, _ in o._fields_}

    # ---- continuous batching
    def serve_start(self, max_slots=0, keep_results=1):
        so = cgp_serve_opts(int(max_slots), int(keep_results))
        check(lib.cgp_serve_start(self._h, C.byref(so)))

    def serve_submit(self, kind, source, ticket, arrival_s):
        q = cgp_query(int(source), int(kind))
        check(lib.cgp_serve_submit(self._h, C.byref(q), int(ticket), float(arrival_s)))

    def serve_step(self, max_iterations=0, cap=1024, fetch=True):
        """Runs up to max_iterations iterations (0: until something completes)
        and returns the completions.  With fetch=False the host result
        buffers are released without being copied ("result" stays None)."""
        done = (cgp_completion * cap)()
        nd = C.c_int32(0)
        check(lib.cgp_serve_step(self._h, int(max_iterations), done, cap, C.byref(nd)))
        out = []
        V = self.graph.V
        for i in range(nd.value):
            c = done[i]
            rec = {"ticket": c.ticket, "kind": c.kind, "source": c.source, "arrival_s": c.arrival_s,
                   "admit_s": c.admit_s, "done_s": c.done_s, "iterations": c.iterations, "result": None}
            if c.result:
                if fetch:
                    rec["result"] = np.ctypeslib.as_array(c.result, shape=(V,)).copy()
                lib.cgp_serve_release_result(c.result)
            out.append(rec)
        return out

    def serve_clock(self):
        return lib.cgp_serve_clock(self._h)

    def serve_pending(self):
        a = C.c_int32()
        b = C.c_int32()
        check(lib.cgp_serve_pending(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def serve_capacity(self):
        """(slots, bytes of value rows) of the running session."""
        n = C.c_int32(0)
        b = C.c_uint64(0)
        check(lib.cgp_serve_capacity(self._h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def serve_stop(self):
        check(lib.cgp_serve_stop(self._h))


def reached_edges(graph, out_buffer, n):
    m = np.zeros(n, dtype=np.uint64)
    check(lib.cgp_reached_edges(graph._h, out_buffer.ptr, int(n),
                                m.ctypes.data_as(C.POINTER(C.c_uint64))))
    return m
