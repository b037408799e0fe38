"""ctypes binding of the C-ABI in include/cgp/cgp.h (libcgp_amd.so).

The shared library is built in-tree (``cgp_amd/libcgp_amd.so``) by
``__graft_entry__.build()`` / ``make -C cgp_amd/csrc``.  Importing this module
fails loudly when the library is missing: there is no fallback path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcgp_amd.so")

CGP_OK, CGP_EINVAL, CGP_ENOMEM, CGP_EDEVICE, CGP_EBUSY, CGP_ECAPACITY, CGP_EIO = range(7)
STATUS_NAMES = {
    CGP_OK: "CGP_OK",
    CGP_EINVAL: "CGP_EINVAL",
    CGP_ENOMEM: "CGP_ENOMEM",
    CGP_EDEVICE: "CGP_EDEVICE",
    CGP_EBUSY: "CGP_EBUSY",
    CGP_ECAPACITY: "CGP_ECAPACITY",
    CGP_EIO: "CGP_EIO",
}
BFS, SSSP, WCC = 0, 1, 2
PUSH, PULL, HYBRID = 0, 1, 2
INF = 0xFFFFFFFF
# Loader formats and flags, report formats (cgp.h).
FORMAT_GR, FORMAT_MATRIX_MARKET, FORMAT_CSR_DIRECTORY = 0, 1, 2
LOAD_UNIT_WEIGHTS, LOAD_ROUND_WEIGHTS = 1, 2
REPORT_JSON, REPORT_CSV = 0, 1
RENUMBER_AUTO, RENUMBER_ON, RENUMBER_OFF = 0, 1, 2


class cgp_query(C.Structure):
    _fields_ = [("source", C.c_int32), ("kind", C.c_int32)]


class cgp_engine_opts(C.Structure):
    _fields_ = [
        ("traversal_mode", C.c_int32),
        ("pull_frontier_ratio", C.c_double),
        ("pull_edge_ratio", C.c_double),
        ("max_iterations", C.c_int32),
        ("max_words", C.c_int32),
        ("block_size", C.c_int32),
        ("bin_small", C.c_int32),
        ("bin_large", C.c_int32),
        ("use_hip_graph", C.c_int32),
        ("graph_iterations", C.c_int32),
        ("profile", C.c_int32),
        ("hbm_budget_bytes", C.c_uint64),
        ("dense_threshold", C.c_int32),
        ("frontier_dense_div", C.c_int32),
        ("dense_unroll", C.c_int32),
        ("dense_narrow", C.c_int32),
        ("concurrent_advance", C.c_int32),
        ("pull_max_queries", C.c_int32),
        ("advance_blocks", C.c_int32),
        ("sssp_delta", C.c_int32),
        ("dense_wide_addr", C.c_int32),
    ]


class cgp_run_stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("push_iterations", C.c_int32),
        ("pull_iterations", C.c_int32),
        ("passes", C.c_int32),
        ("wall_ms", C.c_double),
        ("init_ms", C.c_double),
        ("advance_ms", C.c_double),
        ("filter_ms", C.c_double),
        ("wcc_ms", C.c_double),
        ("output_ms", C.c_double),
        ("advance_launches", C.c_int64),
        ("advance_bytes", C.c_double),
        ("frontier_vertices", C.c_int64),
        ("edges_scanned", C.c_int64),
        ("dense_ms", C.c_double),
        ("dense_launches", C.c_int64),
        ("dense_bytes", C.c_double),
        ("work", C.c_uint64 * 24),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["work"] = [int(x) for x in self.work]
        return d


class cgp_iteration_record(C.Structure):
    _fields_ = [
        ("pass_", C.c_int32),
        ("iteration", C.c_int32),
        ("pull", C.c_int32),
        ("dense", C.c_int32),
        ("frontier_items", C.c_int64),
        ("touched", C.c_int64),
        ("next_items", C.c_int64),
        ("next_edges", C.c_int64),
        ("next_bfs_pairs", C.c_int64),
        ("advance_ms", C.c_double),
        ("filter_ms", C.c_double),
        ("next_pairs", C.c_int64),
        ("next_vertices", C.c_int64),
        ("next_query_edges", C.c_int64),
        ("edges_scanned", C.c_int64),
        ("dense_ms", C.c_double),
        ("end_ms", C.c_double),
        ("frontier_union", C.c_uint64 * 8),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["pass"] = d.pop("pass_")
        d["frontier_union"] = [int(x) for x in self.frontier_union]
        return d


class cgp_serve_opts(C.Structure):
    _fields_ = [("max_slots", C.c_int32), ("keep_results", C.c_int32)]


class cgp_completion(C.Structure):
    _fields_ = [
        ("ticket", C.c_int64),
        ("kind", C.c_int32),
        ("source", C.c_int32),
        ("arrival_s", C.c_double),
        ("admit_s", C.c_double),
        ("done_s", C.c_double),
        ("iterations", C.c_int32),
        ("result", C.POINTER(C.c_uint32)),
    ]


_p = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)
_st = C.c_int

# name -> (restype, argtypes); every function the header declares.
SIGNATURES = {
    "cgp_last_error": (C.c_char_p, []),
    "cgp_abi_version": (C.c_int32, []),
    "cgp_device_count": (_st, [_i32p]),
    "cgp_device_info": (_st, [C.c_char_p, C.c_int32, _i32p, _u64p]),
    "cgp_device_alloc": (_st, [C.c_uint64, _pp]),
    "cgp_device_free": (_st, [_p]),
    "cgp_copy_to_host": (_st, [_p, _p, C.c_uint64]),
    "cgp_copy_to_device": (_st, [_p, _p, C.c_uint64]),
    "cgp_copy_device": (_st, [_p, _p, C.c_uint64]),
    "cgp_gen_rmat": (_st, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                           C.c_uint32, C.c_uint32, _pp]),
    "cgp_gen_grid": (_st, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, C.c_uint32, C.c_uint32, _pp]),
    "cgp_host_csr_from_coo": (_st, [C.c_int32, C.c_int64, _i32p, _i32p, _u32p, _pp]),
    "cgp_load_gr": (_st, [C.c_char_p, _pp]),
    "cgp_load_matrix_market": (_st, [C.c_char_p, _pp]),
    "cgp_load_csr_directory": (_st, [C.c_char_p, _pp]),
    "cgp_load_graph": (_st, [C.c_char_p, C.c_int32, C.c_int32, _pp]),
    "cgp_host_csr_view": (_st, [_p, _i32p, _i64p, C.POINTER(_i32p), C.POINTER(_i32p), C.POINTER(_u32p)]),
    "cgp_host_csr_destroy": (None, [_p]),
    "cgp_sample_sources": (_st, [_p, C.c_int32, C.c_uint64, _i32p]),
    "cgp_poisson_arrivals": (_st, [C.c_int32, C.c_double, C.c_uint64, _dp]),
    "cgp_graph_create_csr": (_st, [C.c_int32, C.c_int64, _i32p, _i32p, _u32p, _pp]),
    "cgp_graph_from_host": (_st, [_p, _pp]),
    "cgp_graph_create_csr_ex": (_st, [C.c_int32, C.c_int64, _i32p, _i32p, _u32p, C.c_int32, _pp]),
    "cgp_graph_from_host_ex": (_st, [_p, C.c_int32, _pp]),
    "cgp_graph_create_csr_device": (_st, [C.c_int32, C.c_int64, _p, _p, _p, _pp]),
    "cgp_graph_destroy": (None, [_p]),
    "cgp_graph_num_vertices": (C.c_int32, [_p]),
    "cgp_graph_num_edges": (C.c_int64, [_p]),
    "cgp_graph_is_symmetric": (C.c_int32, [_p]),
    "cgp_engine_opts_default": (None, [C.POINTER(cgp_engine_opts)]),
    "cgp_engine_create": (_st, [_p, C.POINTER(cgp_engine_opts), _pp]),
    "cgp_engine_destroy": (None, [_p]),
    "cgp_engine_get_opts": (_st, [_p, C.POINTER(cgp_engine_opts)]),
    "cgp_engine_set_opts": (_st, [_p, C.POINTER(cgp_engine_opts)]),
    "cgp_run_batch": (_st, [_p, C.POINTER(cgp_query), C.c_int32, _p, C.c_int32, C.POINTER(cgp_run_stats)]),
    "cgp_run_batch_static": (_st, [_p, C.POINTER(cgp_query), C.c_int32, _p, C.POINTER(cgp_run_stats)]),
    "cgp_reached_edges": (_st, [_p, _p, C.c_int32, _u64p]),
    "cgp_last_run_report": (_st, [_p, C.POINTER(cgp_iteration_record), C.c_int32, _i32p]),
    "cgp_last_run_report_text": (_st, [_p, C.c_int32, C.c_char_p, C.c_char_p, C.c_uint64, _u64p]),
    "cgp_engine_autotune": (_st, [_p, C.POINTER(cgp_query), C.c_int32, C.POINTER(cgp_engine_opts)]),
    "cgp_serve_start": (_st, [_p, C.POINTER(cgp_serve_opts)]),
    "cgp_serve_submit": (_st, [_p, C.POINTER(cgp_query), C.c_int64, C.c_double]),
    "cgp_serve_step": (_st, [_p, C.c_int32, C.POINTER(cgp_completion), C.c_int32, _i32p]),
    "cgp_serve_clock": (C.c_double, [_p]),
    "cgp_serve_pending": (_st, [_p, _i32p, _i32p]),
    "cgp_serve_capacity": (_st, [_p, _i32p, _u64p]),
    "cgp_serve_release_result": (None, [_u32p]),
    "cgp_serve_stop": (_st, [_p]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "cgp_amd: %s is not built (run __graft_entry__.build() or make -C cgp_amd/csrc)" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class CgpError(RuntimeError):
    """Raised for a non-OK cgp_status; carries the status and cgp_last_error()."""

    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status
        self.message = message


def check(status):
    if status != CGP_OK:
        msg = lib.cgp_last_error()
        raise CgpError(status, msg.decode() if msg else "")
