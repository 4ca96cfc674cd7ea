"""Host CSR graphs (ECLgraph layout), synthetic generators, .egr I/O and
device-resident copies.

HostGraph mirrors the reference's struct ECLgraph (include/ECLgraph.h:44-50):
nodes, edges, nindex[nodes+1], nlist[edges], optional eweight[edges], all
int32.  The generators and the CSR builder run in the library's host C++
(include/indigo_graph.h); DeviceGraph uploads a HostGraph into HBM.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib


@dataclass
class HostGraph:
    nodes: int
    edges: int
    nindex: np.ndarray
    nlist: np.ndarray
    eweight: Optional[np.ndarray] = None
    name: str = ""

    def __post_init__(self):
        self.nindex = np.ascontiguousarray(self.nindex, dtype=np.int32)
        self.nlist = np.ascontiguousarray(self.nlist, dtype=np.int32)
        if self.eweight is not None:
            self.eweight = np.ascontiguousarray(self.eweight, dtype=np.int32)
        if self.nindex.shape != (self.nodes + 1,) or self.nlist.shape != (self.edges,):
            raise ValueError("CSR arrays do not match nodes/edges")

    @property
    def degree(self) -> np.ndarray:
        return np.diff(self.nindex).astype(np.int32)

    def sp(self) -> np.ndarray:
        """Edge-source array sp[e] = v for e in [nindex[v], nindex[v+1])
        (reference bfs_edge_cuda.h:140-146)."""
        return np.repeat(np.arange(self.nodes, dtype=np.int32), self.degree)

    def with_weights(self, weights: np.ndarray) -> "HostGraph":
        return HostGraph(self.nodes, self.edges, self.nindex, self.nlist, weights, self.name)

    def _struct(self) -> _lib.HostGraphStruct:
        s = _lib.HostGraphStruct()
        s.nodes = self.nodes
        s.edges = self.edges
        s.nindex = self.nindex.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        s.nlist = self.nlist.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        s.eweight = (
            self.eweight.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if self.eweight is not None else None
        )
        return s

    def check(self) -> int:
        """0 if simple, sorted, symmetric (with symmetric weights); else the
        code of include/indigo_graph.h indigo_graph_check."""
        s = self._struct()
        return _lib.load().indigo_graph_check(ctypes.byref(s))

    def write_egr(self, path: str) -> None:
        s = self._struct()
        _lib.call("indigo_write_egr", str(path).encode(), ctypes.byref(s))


def _take(s: _lib.HostGraphStruct, name: str = "") -> HostGraph:
    """Copies a library-allocated graph into numpy arrays and releases it."""
    try:
        n, e = s.nodes, s.edges
        nindex = np.ctypeslib.as_array(s.nindex, shape=(n + 1,)).copy()
        nlist = np.ctypeslib.as_array(s.nlist, shape=(max(e, 1),))[:e].copy()
        ew = None
        if bool(s.eweight):
            ew = np.ctypeslib.as_array(s.eweight, shape=(max(e, 1),))[:e].copy()
        return HostGraph(n, e, nindex, nlist, ew, name)
    finally:
        _lib.load().indigo_graph_release(ctypes.byref(s))


def from_edges(nodes: int, src, dst, symmetrize: bool = True, wseed: int = 0, wmin: int = 0, wmax: int = -1,
               threads: int = 0, name: str = "") -> HostGraph:
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    s = _lib.HostGraphStruct()
    pi = ctypes.POINTER(ctypes.c_int)
    _lib.call("indigo_build_csr", nodes, len(src), src.ctypes.data_as(pi), dst.ctypes.data_as(pi),
              1 if symmetrize else 0, wseed, wmin, wmax, threads, ctypes.byref(s))
    return _take(s, name)


def grid(rows: int, cols: int, seed: int = 0, wmin: int = 0, wmax: int = -1) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_grid", rows, cols, seed, wmin, wmax, ctypes.byref(s))
    return _take(s, f"grid{rows}x{cols}")


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 1,
         wmin: int = 0, wmax: int = -1, threads: int = 0) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_rmat", scale, edge_factor, a, b, c, seed, wmin, wmax, threads, ctypes.byref(s))
    return _take(s, f"rmat{scale}_ef{edge_factor}")


def erdos_renyi(nodes: int, samples: int, seed: int = 1, wmin: int = 0, wmax: int = -1,
                threads: int = 0) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_er", nodes, samples, seed, wmin, wmax, threads, ctypes.byref(s))
    return _take(s, f"er{nodes}")


def road(rows: int, cols: int, keep_p: float = 0.85, diag_p: float = 0.02, seed: int = 1, wmin: int = 0,
         wmax: int = -1, threads: int = 0) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_road", rows, cols, keep_p, diag_p, seed, wmin, wmax, threads, ctypes.byref(s))
    return _take(s, f"road{rows}x{cols}")


def read_egr(path: str) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_read_egr", str(path).encode(), ctypes.byref(s))
    return _take(s, str(path))


class DeviceBuffer:
    """A device allocation made through the C ABI (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _lib.call("indigo_malloc", ctypes.byref(p), max(self.nbytes, 4))
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            _lib.call("indigo_memcpy_h2d", buf.ptr, arr.ctypes.data, arr.nbytes)
        return buf

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _lib.call("indigo_memcpy_d2h", out.ctypes.data, self.ptr, out.nbytes)
        return out

    def close(self) -> None:
        if self.ptr:
            _lib.load().indigo_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """A HostGraph resident in device memory (nindex, nlist, eweight)."""

    def __init__(self, g: HostGraph):
        self.nodes = g.nodes
        self.edges = g.edges
        self.name = g.name
        self.d_nindex = DeviceBuffer.from_array(g.nindex)
        self.d_nlist = DeviceBuffer.from_array(g.nlist)
        self.d_eweight = DeviceBuffer.from_array(g.eweight) if g.eweight is not None else None

    @property
    def footprint_bytes(self) -> int:
        b = self.d_nindex.nbytes + self.d_nlist.nbytes
        if self.d_eweight is not None:
            b += self.d_eweight.nbytes
        return b

    def close(self) -> None:
        for b in (self.d_nindex, self.d_nlist, self.d_eweight):
            if b is not None:
                b.close()
