"""ctypes binding of the CPU oracle (liboracle.so, built from indigo_oracle.c)
and of the reference's own C++ code compiled into _ref/ (when built).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py, and only as the checker or the CPU
baseline - never on the product path.

Besides the bindings, synth_weights restates the reference harnesses'
weight synthesis in numpy; it is pinned by hand-worked values, two of them
with a wrapping 32-bit product (tests/test_harness_host.py).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_longlong, c_uint, c_ulonglong

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

_PI = POINTER(c_int)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_bfs.argtypes = [c_int, _PI, _PI, c_int, _PI]
        lib.orc_sssp.argtypes = [c_int, _PI, _PI, _PI, c_int, _PI]
        lib.orc_cc.argtypes = [c_int, _PI, _PI, _PI]
        lib.orc_mis.argtypes = [c_int, _PI, _PI, c_int, _PI]
        lib.orc_mis_hash.argtypes = [c_uint]
        lib.orc_mis_hash.restype = c_uint
        lib.orc_pr_ref32.argtypes = [c_int, _PI, _PI, c_int, c_double, POINTER(c_float)]
        lib.orc_pr_ref32.restype = c_int
        lib.orc_pr64.argtypes = [c_int, _PI, _PI, c_int, POINTER(c_double)]
        lib.orc_pr64.restype = c_double
        lib.orc_tc.argtypes = [c_int, _PI, _PI]
        lib.orc_tc.restype = c_ulonglong
        for name in ("orc_bfs", "orc_sssp", "orc_cc", "orc_mis"):
            getattr(lib, name).restype = None
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(_PI)


def _csr(g):
    return (np.ascontiguousarray(g.nindex, dtype=np.int32), np.ascontiguousarray(g.nlist, dtype=np.int32))


def bfs(g, src: int) -> np.ndarray:
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.int32)
    load().orc_bfs(g.nodes, _p(ni), _p(nl), src, _p(out))
    return out


def sssp(g, src: int) -> np.ndarray:
    ni, nl = _csr(g)
    w = np.ascontiguousarray(g.eweight, dtype=np.int32)
    out = np.empty(g.nodes, np.int32)
    load().orc_sssp(g.nodes, _p(ni), _p(nl), _p(w), src, _p(out))
    return out


def cc(g) -> np.ndarray:
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.int32)
    load().orc_cc(g.nodes, _p(ni), _p(nl), _p(out))
    return out


def mis(g, signed_priority: bool = False) -> np.ndarray:
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.int32)
    load().orc_mis(g.nodes, _p(ni), _p(nl), int(signed_priority), _p(out))
    return out


def mis_hash(v: int) -> int:
    return int(load().orc_mis_hash(v & 0xFFFFFFFF))


def pr_ref32(g, max_iter: int = 100, eps: float = 1e-4) -> tuple[np.ndarray, int]:
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.float32)
    it = load().orc_pr_ref32(g.nodes, _p(ni), _p(nl), max_iter, eps, out.ctypes.data_as(POINTER(c_float)))
    return out, it


def pr64(g, iters: int) -> tuple[np.ndarray, float]:
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.float64)
    diff = load().orc_pr64(g.nodes, _p(ni), _p(nl), iters, out.ctypes.data_as(POINTER(c_double)))
    return out, diff


def tc(g) -> int:
    ni, nl = _csr(g)
    return int(load().orc_tc(g.nodes, _p(ni), _p(nl)))


def synth_weights(g) -> np.ndarray:
    """Weights the reference harnesses make for a graph stored without them
    (include/sssp_vertex_cuda.h:127-137, include/mis_vertex_cuda.h:117-127):
    w[j] = 1 + ((i * nei) % nodes) for entry j = (i, nei) of the CSR, the
    product in 32-bit int (wrapping, as gcc and clang compile it), % truncating
    toward zero as in C, then negated if negative."""
    ni = np.asarray(g.nindex, dtype=np.int64)
    nl = np.asarray(g.nlist, dtype=np.int64)
    rows = np.repeat(np.arange(g.nodes, dtype=np.int64), np.diff(ni))
    prod = (rows * nl) & 0xFFFFFFFF
    prod = np.where(prod >= 1 << 31, prod - (1 << 32), prod)
    w = 1 + np.fmod(prod, g.nodes)
    return np.abs(w).astype(np.int32)


# ---- the compiled reference (oracle/_ref) ------------------------------------
def ref_available(alg: str) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libref_{alg}.so"))


_ref = {}


def _ref_lib(alg: str):
    if alg not in _ref:
        path = os.path.join(REF_DIR, f"libref_{alg}.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} not built (make -C oracle with the reference present)")
        lib = ctypes.CDLL(path)
        if alg == "bfs":
            lib.ref_bfs_threads.argtypes = [c_int, c_int, _PI, _PI, c_int, _PI, c_int]
            lib.ref_bfs_threads.restype = c_double
            lib.ref_bfs_serial.argtypes = [c_int, c_int, _PI, _PI, c_int, _PI]
            lib.ref_bfs_serial.restype = None
        elif alg == "sssp":
            lib.ref_sssp_threads.argtypes = [c_int, c_int, _PI, _PI, _PI, c_int, _PI, c_int]
            lib.ref_sssp_threads.restype = c_double
            lib.ref_sssp_serial.argtypes = [c_int, c_int, _PI, _PI, _PI, c_int, _PI]
            lib.ref_sssp_serial.restype = None
        elif alg == "cc":
            lib.ref_cc_threads.argtypes = [c_int, c_int, _PI, _PI, _PI, c_int]
            lib.ref_cc_threads.restype = c_double
        elif alg == "mis":
            lib.ref_mis_threads.argtypes = [c_int, c_int, _PI, _PI, POINTER(c_uint), _PI, c_int]
            lib.ref_mis_threads.restype = c_double
        elif alg == "mis_omp":
            lib.ref_mis_omp.argtypes = [c_int, c_int, _PI, _PI, POINTER(c_uint), _PI, c_int]
            lib.ref_mis_omp.restype = c_double
        elif alg == "pr":
            lib.ref_pr_threads.argtypes = [c_int, c_int, _PI, _PI, POINTER(c_float), _PI, c_int]
            lib.ref_pr_threads.restype = c_double
        elif alg == "bfs_omp":
            lib.ref_bfs_omp.argtypes = [c_int, c_int, _PI, _PI, c_int, _PI, c_int]
            lib.ref_bfs_omp.restype = c_double
        elif alg == "sssp_omp":
            lib.ref_sssp_omp.argtypes = [c_int, c_int, _PI, _PI, _PI, c_int, _PI, c_int]
            lib.ref_sssp_omp.restype = c_double
        elif alg == "cc_omp":
            lib.ref_cc_omp.argtypes = [c_int, c_int, _PI, _PI, _PI, c_int]
            lib.ref_cc_omp.restype = c_double
        elif alg == "pr_omp":
            lib.ref_pr_omp.argtypes = [c_int, c_int, _PI, _PI, POINTER(c_float), _PI, c_int]
            lib.ref_pr_omp.restype = c_double
        elif alg == "tc_omp":
            lib.ref_tc_omp.argtypes = [c_int, _PI, _PI, POINTER(c_longlong), c_int]
            lib.ref_tc_omp.restype = c_double
        elif alg == "tc":
            lib.ref_tc_threads.argtypes = [c_int, _PI, _PI, POINTER(c_longlong), c_int]
            lib.ref_tc_threads.restype = c_double
            lib.ref_tc_serial.argtypes = [c_int, _PI, _PI]
            lib.ref_tc_serial.restype = c_longlong
        _ref[alg] = lib
    return _ref[alg]


def ref_bfs(g, src: int, threads: int = 1, serial: bool = False):
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.int32)
    lib = _ref_lib("bfs")
    if serial:
        lib.ref_bfs_serial(g.nodes, g.edges, _p(ni), _p(nl), src, _p(out))
        return out, 0.0
    rt = lib.ref_bfs_threads(g.nodes, g.edges, _p(ni), _p(nl), src, _p(out), threads)
    return out, rt


def ref_sssp(g, src: int, threads: int = 1, serial: bool = False):
    ni, nl = _csr(g)
    w = np.ascontiguousarray(g.eweight, dtype=np.int32)
    out = np.empty(g.nodes, np.int32)
    lib = _ref_lib("sssp")
    if serial:
        lib.ref_sssp_serial(g.nodes, g.edges, _p(ni), _p(nl), _p(w), src, _p(out))
        return out, 0.0
    rt = lib.ref_sssp_threads(g.nodes, g.edges, _p(ni), _p(nl), _p(w), src, _p(out), threads)
    return out, rt


def ref_cc(g, threads: int = 1):
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.int32)
    rt = _ref_lib("cc").ref_cc_threads(g.nodes, g.edges, _p(ni), _p(nl), _p(out), threads)
    return out, rt


def ref_mis(g, threads: int = 1):
    ni, nl = _csr(g)
    st = np.empty(g.nodes, np.int32)
    pr = np.empty(g.nodes, np.uint32)
    rt = _ref_lib("mis").ref_mis_threads(g.nodes, g.edges, _p(ni), _p(nl), pr.ctypes.data_as(POINTER(c_uint)),
                                         _p(st), threads)
    return st, pr, rt


def ref_mis_omp(g, threads: int = 1):
    """The reference's best published OpenMP MIS variant (oracle/_ref)."""
    ni, nl = _csr(g)
    st = np.empty(g.nodes, np.int32)
    pr = np.empty(g.nodes, np.uint32)
    rt = _ref_lib("mis_omp").ref_mis_omp(g.nodes, g.edges, _p(ni), _p(nl), pr.ctypes.data_as(POINTER(c_uint)),
                                         _p(st), threads)
    return st, pr, rt


def ref_pr(g, threads: int = 1):
    ni, nl = _csr(g)
    sc = np.empty(g.nodes, np.float32)
    it = c_int()
    rt = _ref_lib("pr").ref_pr_threads(g.nodes, g.edges, _p(ni), _p(nl), sc.ctypes.data_as(POINTER(c_float)),
                                       ctypes.byref(it), threads)
    return sc, it.value, rt


def ref_tc(g, threads: int = 1, serial: bool = False):
    ni, nl = _csr(g)
    lib = _ref_lib("tc")
    if serial:
        return int(lib.ref_tc_serial(g.nodes, _p(ni), _p(nl))), 0.0
    c = c_longlong()
    rt = lib.ref_tc_threads(g.nodes, _p(ni), _p(nl), ctypes.byref(c), threads)
    return int(c.value), rt


def ref_tc_omp(g, threads: int = 1):
    """The reference's best published OpenMP TC variant (oracle/_ref)."""
    ni, nl = _csr(g)
    c = c_longlong()
    rt = _ref_lib("tc_omp").ref_tc_omp(g.nodes, _p(ni), _p(nl), ctypes.byref(c), threads)
    return int(c.value), rt


def ref_bfs_omp(g, src: int, threads: int = 1):
    """The OpenMP BFS variant that won the CPU sweep (oracle/_ref)."""
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.int32)
    rt = _ref_lib("bfs_omp").ref_bfs_omp(g.nodes, g.edges, _p(ni), _p(nl), src, _p(out), threads)
    return out, rt


def ref_sssp_omp(g, src: int, threads: int = 1):
    """The OpenMP edge-centric SSSP variant that won the CPU sweep (oracle/_ref)."""
    ni, nl = _csr(g)
    w = np.ascontiguousarray(g.eweight, dtype=np.int32)
    out = np.empty(g.nodes, np.int32)
    rt = _ref_lib("sssp_omp").ref_sssp_omp(g.nodes, g.edges, _p(ni), _p(nl), _p(w), src, _p(out), threads)
    return out, rt


def ref_cc_omp(g, threads: int = 1):
    """The OpenMP edge-centric CC variant that won the CPU sweep (oracle/_ref)."""
    ni, nl = _csr(g)
    out = np.empty(g.nodes, np.int32)
    rt = _ref_lib("cc_omp").ref_cc_omp(g.nodes, g.edges, _p(ni), _p(nl), _p(out), threads)
    return out, rt


def ref_pr_omp(g, threads: int = 1):
    """The OpenMP PageRank variant that won the CPU sweep (oracle/_ref)."""
    ni, nl = _csr(g)
    sc = np.empty(g.nodes, np.float32)
    it = c_int()
    rt = _ref_lib("pr_omp").ref_pr_omp(g.nodes, g.edges, _p(ni), _p(nl), sc.ctypes.data_as(POINTER(c_float)),
                                       ctypes.byref(it), threads)
    return sc, it.value, rt
