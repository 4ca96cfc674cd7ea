"""ctypes binding of lib/libindigo2_hip.so (the C ABI in include/indigo_hip.h
and include/indigo_graph.h).

The library is the product: every GPU result comes from its HIP kernels.
Loading fails loudly when the shared object has not been built; there is no
fallback implementation.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_longlong, c_size_t, c_uint, c_ulonglong, c_void_p

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libindigo2_hip.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_OUT_OF_MEMORY = -2
ERR_DEVICE = -3
ERR_UNSORTED = -4
ERR_TIMEOUT = -5
ERR_CAPACITY = -6
ERR_NO_DEVICE = -7

HINT_BFS_DIRECTION = 1
HINT_SSSP_DELTA = 2
HINT_SMALL_FRONTIER = 3
HINT_BUDGET_MS = 4
HINT_BLOCKS_PER_CU = 5
HINT_TC_LDS_LIST = 6
HINT_RESIDENCY_MS = 7
HINT_MAX_WORKGROUPS = 8
HINT_VARIANT = 9
HINT_BLOCK_SIZE = 10
HINT_GRAPH = 11
HINT_TC_HUBS = 12
VARIANT_VERTEX = 0
VARIANT_EDGE = 1

CC_NOALLOC_MAX_NODES = 1 << 22


class IndigoError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"indigo error {code}: {message}")
        self.code = code


class EgrMapStruct(ctypes.Structure):
    """struct indigo_egr_map: a read-only mapping of an .egr file."""

    _fields_ = [
        ("nodes", c_int),
        ("edges", c_int),
        ("nindex", POINTER(c_int)),
        ("nlist", POINTER(c_int)),
        ("eweight", POINTER(c_int)),
        ("base", c_void_p),
        ("length", c_size_t),
    ]


class HostGraphStruct(ctypes.Structure):
    """struct indigo_host_graph (layout of the reference's struct ECLgraph)."""

    _fields_ = [
        ("nodes", c_int),
        ("edges", c_int),
        ("nindex", POINTER(c_int)),
        ("nlist", POINTER(c_int)),
        ("eweight", POINTER(c_int)),
    ]


class GraphRefStruct(ctypes.Structure):
    """struct indigo_graph_ref: one device-resident graph of a batch."""

    _fields_ = [
        ("nodes", c_int),
        ("edges", c_int),
        ("d_nindex", c_void_p),
        ("d_nlist", c_void_p),
        ("d_eweight", c_void_p),
    ]


QUERY_BFS, QUERY_SSSP, QUERY_CC, QUERY_MIS, QUERY_PR = 1, 2, 3, 4, 5
MAX_BFS_SOURCES, MAX_SSSP_SOURCES, MAX_GANG = 64, 32, 256


class QueryStruct(ctypes.Structure):
    """struct indigo_query: one query of a batch."""

    _fields_ = [
        ("alg", c_int),
        ("graph", c_int),
        ("nsrc", c_int),
        ("gang", c_int),
        ("src", c_int * MAX_BFS_SOURCES),
        ("d_out", c_void_p),
    ]


QUERY_TC = 6


class RequestStruct(ctypes.Structure):
    """struct indigo_request: one request of a serving run."""

    _fields_ = [
        ("alg", c_int),
        ("graph", c_int),
        ("src", c_int),
        ("reserved", c_int),
        ("arrival", c_double),
        ("d_out", c_void_p),
        ("admitted", c_double),
        ("finished", c_double),
        ("batch", c_int),
        ("slot", c_int),
    ]


class ServeConfigStruct(ctypes.Structure):
    _fields_ = [("device", c_int), ("slots", c_int), ("max_batch", c_int), ("mem_budget", c_size_t),
                ("progress_interval", c_double), ("cu_partition", c_int), ("batch_work_us", c_int)]


class ServeStatsStruct(ctypes.Structure):
    _fields_ = [("makespan", c_double), ("launches", c_int), ("tc_launches", c_int), ("peak_bytes", c_size_t),
                ("max_waiting", c_int), ("slots", c_int)]


_PI = POINTER(c_int)
_PU = POINTER(c_uint)
_PF = POINTER(c_float)
_PD = POINTER(c_double)
_PLL = POINTER(c_longlong)
_PHG = POINTER(HostGraphStruct)

# name -> argtypes; every function returns int
_SIGNATURES = {
    "indigo_version": [],
    "indigo_device_info": [c_int, _PI, _PI, c_char_p, c_int],
    "indigo_malloc": [POINTER(c_void_p), c_size_t],
    "indigo_free": [c_void_p],
    "indigo_memcpy_h2d": [c_void_p, c_void_p, c_size_t],
    "indigo_memcpy_d2h": [c_void_p, c_void_p, c_size_t],
    # drop-in entries
    "indigo_bfs_vertex": [c_int, c_int, c_int, c_void_p, c_void_p, _PI, _PD],
    "indigo_bfs_edge": [c_int, c_int, c_int, c_void_p, c_void_p, _PI, _PI, _PD],
    "indigo_sssp_vertex": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, _PI, _PD],
    "indigo_sssp_edge": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, _PI, _PI, _PD],
    "indigo_cc_vertex": [c_int, c_int, c_void_p, c_void_p, _PI, _PD],
    "indigo_cc_edge": [c_int, c_int, c_void_p, c_void_p, _PI, _PI, _PD],
    "indigo_mis_vertex": [c_int, c_int, c_void_p, c_void_p, c_int, _PU, _PI, _PD],
    "indigo_mis_edge": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, _PU, _PI, _PD],
    "indigo_pr": [c_int, c_int, _PI, _PI, _PI, _PF, _PI, _PD],
    "indigo_tc_vertex": [_PLL, c_int, c_void_p, c_void_p, _PD],
    "indigo_tc_edge": [_PLL, c_int, c_void_p, c_void_p, c_void_p, _PD],
    # resident layer
    "indigo_ctx_create": [c_int, POINTER(c_void_p)],
    "indigo_ctx_destroy": [c_void_p],
    "indigo_shutdown": [],
    "indigo_ctx_last_ms": [c_void_p, _PF],
    "indigo_ctx_last_stats": [c_void_p, _PI, _PI],
    "indigo_ctx_trim": [c_void_p],
    "indigo_ctx_last_grid": [c_void_p, _PI, _PI],
    "indigo_ctx_last_kernel": [c_void_p, _PF, ctypes.c_char_p, c_int],
    "indigo_ctx_copy_ms": [c_void_p, c_void_p, c_void_p, c_size_t, _PF],
    "indigo_batch": [c_void_p, POINTER(GraphRefStruct), c_int, POINTER(QueryStruct), c_int, c_int],
    "indigo_batch_gang": [c_int, c_int, c_int],
    "indigo_graph_upload": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, POINTER(GraphRefStruct),
                            POINTER(ctypes.c_double)],
    "indigo_graph_free": [POINTER(GraphRefStruct)],
    "indigo_build_sp": [c_void_p, c_int, c_int, c_void_p, c_void_p],
    "indigo_unpermute": [c_void_p, c_int, c_void_p, c_void_p, c_void_p],
    "indigo_egr_open": [c_char_p, POINTER(EgrMapStruct)],
    "indigo_egr_close": [POINTER(EgrMapStruct)],
    "indigo_serve": [POINTER(GraphRefStruct), c_int, POINTER(RequestStruct), c_int, POINTER(ServeConfigStruct),
                     POINTER(ServeStatsStruct)],
    "indigo_ctx_last_trace": [c_void_p, c_int, _PF, _PU, _PD, _PI, _PI],
    "indigo_ctx_set_hint": [c_void_p, c_int, c_int],
    "indigo_bfs": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p],
    "indigo_sssp": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p],
    "indigo_cc": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p],
    "indigo_mis": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p],
    "indigo_pagerank": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, _PI],
    "indigo_tc": [c_void_p, c_int, c_int, c_void_p, c_void_p, _PLL],
    "indigo_cc_noalloc": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p],
    # host graph tools
    "indigo_build_csr": [c_int, c_longlong, _PI, _PI, c_int, c_ulonglong, c_int, c_int, c_int, _PHG],
    "indigo_gen_grid": [c_int, c_int, c_ulonglong, c_int, c_int, _PHG],
    "indigo_gen_rmat": [c_int, c_int, c_double, c_double, c_double, c_ulonglong, c_int, c_int, c_int, _PHG],
    "indigo_gen_er": [c_int, c_longlong, c_ulonglong, c_int, c_int, c_int, _PHG],
    "indigo_gen_road": [c_int, c_int, c_double, c_double, c_ulonglong, c_int, c_int, c_int, _PHG],
    "indigo_graph_copy": [_PHG, _PHG],
    "indigo_graph_release": [_PHG],
    "indigo_graph_reorder": [_PHG, c_int, _PHG, _PI],
    "indigo_read_egr": [c_char_p, _PHG],
    "indigo_write_egr": [c_char_p, _PHG],
    "indigo_graph_check": [_PHG],
}

# Symbols that must exist once the corresponding layer is built (checked by
# the CPU test that loads the library).
EXPORTED = tuple(_SIGNATURES) + ("indigo_last_error",)

_lib = None

# Live objects that hold library resources, released by the exit hook before
# the interpreter tears modules down and before the HIP runtime unloads.
_contexts: "weakref.WeakSet" = weakref.WeakSet()
_buffers: "weakref.WeakSet" = weakref.WeakSet()
_shut_down = False


def track_context(obj) -> None:
    _contexts.add(obj)


def track_buffer(obj) -> None:
    _buffers.add(obj)


def shutting_down() -> bool:
    """True once the exit hook has run: finalizers must not call the library."""
    return _shut_down


def _shutdown() -> None:
    """Exit hook: destroys the live contexts (each waits for its stream),
    frees the live device buffers, then releases the library's default
    context and waits for the device (indigo_shutdown).  Runs from atexit,
    i.e. before module teardown and before the C runtime's exit handlers, so
    no finalizer reaches the HIP runtime after it has begun to unload."""
    global _shut_down
    if _shut_down:
        return
    if _lib is not None:
        for group in (_contexts, _buffers):
            for obj in list(group):
                try:
                    obj.close()
                except Exception:
                    pass
        _lib.indigo_shutdown()
    _shut_down = True


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the HIP library (once).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C parallelization-artifact-sc23_amd`"
        )
    lib = ctypes.CDLL(path)
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = argtypes
        fn.restype = c_int
    lib.indigo_last_error.argtypes = []
    lib.indigo_last_error.restype = c_char_p
    _lib = lib
    atexit.register(_shutdown)
    return lib


def check(rc: int) -> None:
    if rc != OK:
        msg = load().indigo_last_error().decode(errors="replace")
        raise IndigoError(rc, msg)


def call(name: str, *args) -> None:
    """Calls a C-ABI function by name and raises IndigoError on failure."""
    check(getattr(load(), name)(*args))
