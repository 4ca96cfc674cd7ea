"""Host CSR graphs (ECLgraph layout), synthetic generators, .egr I/O and
device-resident copies.

HostGraph mirrors the reference's struct ECLgraph (include/ECLgraph.h:44-50):
nodes, edges, nindex[nodes+1], nlist[edges], optional eweight[edges], all
int32.  The generators and the CSR builder run in the library's host C++
(include/indigo_graph.h); DeviceGraph uploads a HostGraph into HBM.
"""
from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass, field
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
    # how the graph was made (generator, sizes, seed, weights), written
    # beside an .egr file as its metadata sidecar
    meta: dict = field(default_factory=dict)

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
        return HostGraph(self.nodes, self.edges, self.nindex, self.nlist, weights, self.name, dict(self.meta))

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

    def reordered(self, threads: int = 0) -> tuple["HostGraph", np.ndarray]:
        """Degree-ordered copy (indigo_graph_reorder): (graph, perm) with
        perm[v] the new id of vertex v, new ids by degree, largest first."""
        s = self._struct()
        out = _lib.HostGraphStruct()
        perm = np.empty(self.nodes, dtype=np.int32)
        _lib.call("indigo_graph_reorder", ctypes.byref(s), threads, ctypes.byref(out),
                  perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return _take(out, f"{self.name}-degree-ordered"), perm

    def check(self) -> int:
        """0 if simple, sorted, symmetric (with symmetric weights); else the
        code of include/indigo_graph.h indigo_graph_check."""
        s = self._struct()
        return _lib.load().indigo_graph_check(ctypes.byref(s))

    def write_egr(self, path: str) -> None:
        """Writes the ECLgraph file and, when the graph carries metadata, the
        sidecar `<path>.meta.json` (generator, parameters, seed, sizes)."""
        s = self._struct()
        _lib.call("indigo_write_egr", str(path).encode(), ctypes.byref(s))
        if self.meta:
            with open(meta_path(path), "w") as f:
                json.dump(dict(self.meta, nodes=self.nodes, edges=self.edges,
                               weights=self.eweight is not None), f, indent=1, sort_keys=True)
                f.write("\n")


def meta_path(path: str) -> str:
    return f"{path}.meta.json"


def read_meta(path: str) -> Optional[dict]:
    """The metadata sidecar of an .egr file, or None."""
    try:
        with open(meta_path(path)) as f:
            return json.load(f)
    except FileNotFoundError:
        return None


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


def _with_meta(g: HostGraph, **meta) -> HostGraph:
    g.meta = meta
    return g


def grid(rows: int, cols: int, seed: int = 0, wmin: int = 0, wmax: int = -1) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_grid", rows, cols, seed, wmin, wmax, ctypes.byref(s))
    return _with_meta(_take(s, f"grid{rows}x{cols}"), generator="grid", rows=rows, cols=cols, seed=seed,
                      wmin=wmin, wmax=wmax)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19, seed: int = 1,
         wmin: int = 0, wmax: int = -1, threads: int = 0) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_rmat", scale, edge_factor, a, b, c, seed, wmin, wmax, threads, ctypes.byref(s))
    return _with_meta(_take(s, f"rmat{scale}_ef{edge_factor}"), generator="rmat", scale=scale,
                      edge_factor=edge_factor, a=a, b=b, c=c, seed=seed, wmin=wmin, wmax=wmax)


def erdos_renyi(nodes: int, samples: int, seed: int = 1, wmin: int = 0, wmax: int = -1,
                threads: int = 0) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_er", nodes, samples, seed, wmin, wmax, threads, ctypes.byref(s))
    return _with_meta(_take(s, f"er{nodes}"), generator="erdos_renyi", vertices=nodes, samples=samples, seed=seed,
                      wmin=wmin, wmax=wmax)


def road(rows: int, cols: int, keep_p: float = 0.85, diag_p: float = 0.02, seed: int = 1, wmin: int = 0,
         wmax: int = -1, threads: int = 0) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_gen_road", rows, cols, keep_p, diag_p, seed, wmin, wmax, threads, ctypes.byref(s))
    return _with_meta(_take(s, f"road{rows}x{cols}"), generator="road", rows=rows, cols=cols, keep_p=keep_p,
                      diag_p=diag_p, seed=seed, wmin=wmin, wmax=wmax)


def read_egr(path: str) -> HostGraph:
    s = _lib.HostGraphStruct()
    _lib.call("indigo_read_egr", str(path).encode(), ctypes.byref(s))
    g = _take(s, str(path))
    g.meta = read_meta(path) or {}
    return g


class DeviceBuffer:
    """A device allocation made through the C ABI (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _lib.call("indigo_malloc", ctypes.byref(p), max(self.nbytes, 4))
        self.ptr = p.value
        _lib.track_buffer(self)

    @classmethod
    def adopt(cls, ptr: int, nbytes: int) -> "DeviceBuffer":
        """Takes ownership of a device allocation made by the library."""
        buf = cls.__new__(cls)
        buf.nbytes, buf.ptr = int(nbytes), ptr
        _lib.track_buffer(buf)
        return buf

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
        if self.ptr and not _lib.shutting_down():
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
        self.upload_seconds = None
        self.d_nindex = DeviceBuffer.from_array(g.nindex)
        self.d_nlist = DeviceBuffer.from_array(g.nlist)
        self.d_eweight = DeviceBuffer.from_array(g.eweight) if g.eweight is not None else None

    @classmethod
    def _from_ref(cls, ref: _lib.GraphRefStruct, name: str, seconds: float) -> "DeviceGraph":
        self = cls.__new__(cls)
        self.nodes, self.edges, self.name, self.upload_seconds = ref.nodes, ref.edges, name, seconds
        self.d_nindex = DeviceBuffer.adopt(ref.d_nindex, 4 * (ref.nodes + 1))
        self.d_nlist = DeviceBuffer.adopt(ref.d_nlist, 4 * ref.edges)
        self.d_eweight = DeviceBuffer.adopt(ref.d_eweight, 4 * ref.edges) if ref.d_eweight else None
        return self

    @classmethod
    def upload(cls, ctx, g: HostGraph) -> "DeviceGraph":
        """Staged upload on ctx's stream (indigo_graph_upload)."""
        ref, sec = _lib.GraphRefStruct(), ctypes.c_double()
        pi = ctypes.c_void_p
        _lib.call("indigo_graph_upload", ctx.h, g.nodes, g.edges, pi(g.nindex.ctypes.data), pi(g.nlist.ctypes.data),
                  pi(g.eweight.ctypes.data) if g.eweight is not None else None, ctypes.byref(ref), ctypes.byref(sec))
        return cls._from_ref(ref, g.name, sec.value)

    @classmethod
    def from_file(cls, ctx, path: str) -> "DeviceGraph":
        """Maps the .egr file (indigo_egr_open) and uploads it straight from
        the mapping through the staging buffers: no host copy of the graph
        is made."""
        m = _lib.EgrMapStruct()
        _lib.call("indigo_egr_open", str(path).encode(), ctypes.byref(m))
        try:
            ref, sec = _lib.GraphRefStruct(), ctypes.c_double()
            _lib.call("indigo_graph_upload", ctx.h, m.nodes, m.edges, ctypes.cast(m.nindex, ctypes.c_void_p),
                      ctypes.cast(m.nlist, ctypes.c_void_p),
                      ctypes.cast(m.eweight, ctypes.c_void_p) if m.eweight else None, ctypes.byref(ref),
                      ctypes.byref(sec))
        finally:
            _lib.load().indigo_egr_close(ctypes.byref(m))
        return cls._from_ref(ref, str(path), sec.value)

    def build_sp(self, ctx) -> "DeviceBuffer":
        """The edge-source array on the device (indigo_build_sp)."""
        sp = DeviceBuffer(4 * max(1, self.edges))
        _lib.call("indigo_build_sp", ctx.h, self.nodes, self.edges, self.d_nindex.ptr, sp.ptr)
        return sp

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
class ReorderedGraph:
    """A degree-ordered device copy of a graph (indigo_graph_reorder) and its
    permutation: BFS, SSSP and PageRank run on it through Context.*_reordered
    and come back in the original ids (indigo_unpermute).  Hubs sit at the
    smallest ids, so the per-vertex data that many lists gather share cache
    lines.  Built once per graph: reorder_seconds is the host relabelling."""

    def __init__(self, g: HostGraph, threads: int = 0):
        import time

        t = time.perf_counter()
        h, perm = g.reordered(threads)
        self.reorder_seconds = time.perf_counter() - t
        self.nodes, self.edges, self.name = g.nodes, g.edges, h.name
        self.perm = perm
        self.graph = DeviceGraph(h)
        self.d_perm = DeviceBuffer.from_array(perm)
        self.scratch = DeviceBuffer(4 * max(1, g.nodes))

    def close(self) -> None:
        self.graph.close()
        self.d_perm.close()
        self.scratch.close()


