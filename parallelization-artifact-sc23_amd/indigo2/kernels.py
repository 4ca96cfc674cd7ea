"""Python mirror of the six graph kernels over the C ABI.

Two forms, matching include/indigo_hip.h:

* ``Context``: the resident layer.  Graphs are DeviceGraph objects, outputs
  stay in device memory unless downloaded, and ``last_ms`` is the HIP-event
  time of the last call on the context's stream.
* ``GPUbfs_vertex`` ... ``PR_GPU``: the reference's entry-function names and
  argument meaning (include/*_cuda.h), filling host numpy outputs and
  returning the elapsed seconds, for harness-style use in tests.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .graph import DeviceBuffer, DeviceGraph, HostGraph, ReorderedGraph

INF = np.iinfo(np.int32).max
UNDECIDED, INCLUDED, EXCLUDED = 0, 1, 2  # mis_vertex_cuda.h:50-52


def _pi(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _lib.call("indigo_ctx_create", device, ctypes.byref(h))
        self.h = h.value
        _lib.track_context(self)

    def close(self) -> None:
        if self.h and not _lib.shutting_down():
            _lib.load().indigo_ctx_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_hint(self, key: int, value: int) -> None:
        _lib.call("indigo_ctx_set_hint", self.h, key, value)

    @property
    def last_ms(self) -> float:
        ms = ctypes.c_float()
        _lib.call("indigo_ctx_last_ms", self.h, ctypes.byref(ms))
        return ms.value

    @property
    def last_stats(self) -> tuple[int, int]:
        launches, iters = ctypes.c_int(), ctypes.c_int()
        _lib.call("indigo_ctx_last_stats", self.h, ctypes.byref(launches), ctypes.byref(iters))
        return launches.value, iters.value

    @property
    def last_kernel(self) -> tuple[str, float]:
        """(name, ms) of the main kernel launch of the last call."""
        ms = ctypes.c_float()
        buf = ctypes.create_string_buffer(64)
        _lib.call("indigo_ctx_last_kernel", self.h, ctypes.byref(ms), buf, 64)
        return buf.value.decode(), ms.value

    def copy_ms(self, dst: DeviceBuffer, src: DeviceBuffer, nbytes: int) -> float:
        """Device-to-device copy on the context stream, timed with its events."""
        ms = ctypes.c_float()
        _lib.call("indigo_ctx_copy_ms", self.h, dst.ptr, src.ptr, nbytes, ctypes.byref(ms))
        return ms.value

    def last_trace(self) -> list[tuple[float, int, float, int]]:
        """Per-level (start_us, frontier, degree_sum, mode) of the last BFS /
        SSSP call; mode 0 top-down, 1 bottom-up, 2 one-workgroup, 3 split."""
        m = 256
        us = (ctypes.c_float * m)()
        nf = (ctypes.c_uint * m)()
        ds = (ctypes.c_double * m)()
        mo = (ctypes.c_int * m)()
        n = ctypes.c_int()
        _lib.call("indigo_ctx_last_trace", self.h, m, us, nf, ds, mo, ctypes.byref(n))
        return [(us[i], nf[i], ds[i], mo[i]) for i in range(n.value)]

    def trim(self) -> None:
        _lib.call("indigo_ctx_trim", self.h)

    # ---- device-output forms ------------------------------------------------
    def bfs_into(self, g: DeviceGraph, src: int, out: DeviceBuffer) -> None:
        _lib.call("indigo_bfs", self.h, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, src, out.ptr)

    def sssp_into(self, g: DeviceGraph, src: int, out: DeviceBuffer) -> None:
        if g.d_eweight is None:
            raise ValueError("sssp needs edge weights")
        _lib.call("indigo_sssp", self.h, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, g.d_eweight.ptr, src,
                  out.ptr)

    def cc_into(self, g: DeviceGraph, out: DeviceBuffer) -> None:
        _lib.call("indigo_cc", self.h, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, out.ptr)

    def mis_into(self, g: DeviceGraph, out: DeviceBuffer, signed_priority: bool = False) -> None:
        _lib.call("indigo_mis", self.h, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, int(signed_priority),
                  out.ptr)

    def pagerank_into(self, g: DeviceGraph, out: DeviceBuffer) -> int:
        it = ctypes.c_int()
        _lib.call("indigo_pagerank", self.h, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, out.ptr,
                  ctypes.byref(it))
        return it.value

    def tc(self, g: DeviceGraph) -> int:
        c = ctypes.c_longlong()
        _lib.call("indigo_tc", self.h, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, ctypes.byref(c))
        return int(c.value)

    def batch(self, graphs: list[DeviceGraph], queries: list["Query"], max_workgroups: int = 0) -> None:
        """Runs all queries in one persistent launch (indigo_batch); each
        query's output buffer receives its result."""
        gt = graph_table(graphs)
        qt = (_lib.QueryStruct * max(1, len(queries)))()
        for i, q in enumerate(queries):
            qt[i].alg, qt[i].graph, qt[i].nsrc, qt[i].gang = q.alg, q.graph, len(q.sources), q.gang
            for j, s_ in enumerate(q.sources):
                qt[i].src[j] = s_
            qt[i].d_out = q.out.ptr
        _lib.call("indigo_batch", self.h, gt, len(graphs), qt, len(queries), max_workgroups)

    # ---- degree-ordered copies: results in the original ids -----------------
    def unpermute(self, n: int, d_perm: DeviceBuffer, src: DeviceBuffer, out: DeviceBuffer) -> None:
        """out[v] = src[perm[v]] for v < n (indigo_unpermute)."""
        _lib.call("indigo_unpermute", self.h, n, d_perm.ptr, src.ptr, out.ptr)

    def bfs_reordered_into(self, rg: ReorderedGraph, src: int, out: DeviceBuffer) -> float:
        """BFS from original vertex src on rg; out in the original ids.
        Returns the device ms of the search and the read-back together."""
        self.bfs_into(rg.graph, int(rg.perm[src]), rg.scratch)
        ms = self.last_ms
        self.unpermute(rg.nodes, rg.d_perm, rg.scratch, out)
        return ms + self.last_ms

    def sssp_reordered_into(self, rg: ReorderedGraph, src: int, out: DeviceBuffer) -> float:
        self.sssp_into(rg.graph, int(rg.perm[src]), rg.scratch)
        ms = self.last_ms
        self.unpermute(rg.nodes, rg.d_perm, rg.scratch, out)
        return ms + self.last_ms

    def pagerank_reordered_into(self, rg: ReorderedGraph, out: DeviceBuffer) -> tuple[int, float]:
        it = self.pagerank_into(rg.graph, rg.scratch)
        ms = self.last_ms
        self.unpermute(rg.nodes, rg.d_perm, rg.scratch, out)
        return it, ms + self.last_ms

    def bfs_reordered(self, rg: ReorderedGraph, src: int) -> np.ndarray:
        out = DeviceBuffer(4 * rg.nodes)
        self.bfs_reordered_into(rg, src, out)
        return out.to_array(np.int32, rg.nodes)

    def sssp_reordered(self, rg: ReorderedGraph, src: int) -> np.ndarray:
        out = DeviceBuffer(4 * rg.nodes)
        self.sssp_reordered_into(rg, src, out)
        return out.to_array(np.int32, rg.nodes)

    def pagerank_reordered(self, rg: ReorderedGraph) -> tuple[np.ndarray, int]:
        out = DeviceBuffer(4 * rg.nodes)
        it, _ = self.pagerank_reordered_into(rg, out)
        return out.to_array(np.float32, rg.nodes), it

    # ---- host-output conveniences -------------------------------------------
    def bfs(self, g: DeviceGraph, src: int) -> np.ndarray:
        out = DeviceBuffer(4 * g.nodes)
        self.bfs_into(g, src, out)
        return out.to_array(np.int32, g.nodes)

    def sssp(self, g: DeviceGraph, src: int) -> np.ndarray:
        out = DeviceBuffer(4 * g.nodes)
        self.sssp_into(g, src, out)
        return out.to_array(np.int32, g.nodes)

    def cc(self, g: DeviceGraph) -> np.ndarray:
        out = DeviceBuffer(4 * g.nodes)
        self.cc_into(g, out)
        return out.to_array(np.int32, g.nodes)

    def mis(self, g: DeviceGraph, signed_priority: bool = False) -> np.ndarray:
        out = DeviceBuffer(4 * g.nodes)
        self.mis_into(g, out, signed_priority)
        return out.to_array(np.int32, g.nodes)

    def pagerank(self, g: DeviceGraph) -> tuple[np.ndarray, int]:
        out = DeviceBuffer(4 * g.nodes)
        it = self.pagerank_into(g, out)
        return out.to_array(np.float32, g.nodes), it


class Query:
    """One query of a batch: alg is "bfs", "sssp", "cc", "mis" or "pr"; graph
    is the index into the batch's graph list; BFS / SSSP take up to 64 / 32
    sources and write an int32 [len(sources), nodes] slab; PR writes float
    scores followed by the iteration count.  gang = 0 lets the engine size
    the team of workgroups; an autotuner passes its own choice."""

    ALGS = {"bfs": _lib.QUERY_BFS, "sssp": _lib.QUERY_SSSP, "cc": _lib.QUERY_CC, "mis": _lib.QUERY_MIS,
            "pr": _lib.QUERY_PR}

    def __init__(self, alg: str, graph: int, g: DeviceGraph, sources=(), gang: int = 0):
        self.name = alg
        self.gang = int(gang)
        self.alg = self.ALGS[alg]
        self.graph = graph
        self.sources = [int(s) for s in sources]
        rows = len(self.sources) if alg in ("bfs", "sssp") else 1
        self.nodes = g.nodes
        self.out = DeviceBuffer(4 * (g.nodes * max(rows, 1) + (1 if alg == "pr" else 0)))

    def result(self):
        """int32 slab / labels / status; for PR (float32 scores, iterations)."""
        if self.name == "pr":
            a = self.out.to_array(np.int32, self.nodes + 1)
            return a[: self.nodes].view(np.float32).copy(), int(a[self.nodes])
        rows = len(self.sources) if self.name in ("bfs", "sssp") else 1
        a = self.out.to_array(np.int32, rows * self.nodes)
        return a.reshape(rows, self.nodes) if self.name in ("bfs", "sssp") else a


def batch_gang(alg: str, nodes: int, edges: int) -> int:
    """The batched engine's own gang size for a query (indigo_batch_gang)."""
    return int(_lib.load().indigo_batch_gang(Query.ALGS[alg], int(nodes), int(edges)))


SERVE_ALGS = {"bfs": _lib.QUERY_BFS, "sssp": _lib.QUERY_SSSP, "cc": _lib.QUERY_CC, "mis": _lib.QUERY_MIS,
              "pr": _lib.QUERY_PR, "tc": _lib.QUERY_TC}


def graph_table(graphs: list[DeviceGraph]):
    gt = (_lib.GraphRefStruct * max(1, len(graphs)))()
    for i, g in enumerate(graphs):
        gt[i].nodes, gt[i].edges = g.nodes, g.edges
        gt[i].d_nindex, gt[i].d_nlist = g.d_nindex.ptr, g.d_nlist.ptr
        gt[i].d_eweight = g.d_eweight.ptr if g.d_eweight is not None else None
    return gt


def serve(graphs: list[DeviceGraph], algs, graph_ids, sources, arrivals, outs=None, slots: int = 4,
          max_batch: int = 256, mem_budget: int = 1 << 30, device: int = 0, progress: float = 0.0,
          cu_partition: bool = False, batch_work_us: int = 0):
    """Replays requests through indigo_serve.  algs: names ("bfs", "sssp",
    "cc", "mis", "pr", "tc"); outs: optional DeviceBuffer per request (None
    for a discarded result).  Returns (admitted, finished, batch, slot) numpy
    arrays in seconds and the run statistics as a dict."""
    n = len(algs)
    rq = (_lib.RequestStruct * max(1, n))()
    for i in range(n):
        rq[i].alg = SERVE_ALGS[algs[i]]
        rq[i].graph = int(graph_ids[i])
        rq[i].src = int(sources[i])
        rq[i].arrival = float(arrivals[i])
        rq[i].d_out = outs[i].ptr if outs is not None and outs[i] is not None else None
    cfg = _lib.ServeConfigStruct(device, slots, max_batch, mem_budget, progress, int(cu_partition), batch_work_us)
    st = _lib.ServeStatsStruct()
    _lib.call("indigo_serve", graph_table(graphs), len(graphs), rq, n, ctypes.byref(cfg), ctypes.byref(st))
    admitted = np.array([rq[i].admitted for i in range(n)])
    finished = np.array([rq[i].finished for i in range(n)])
    batch = np.array([rq[i].batch for i in range(n)])
    slot = np.array([rq[i].slot for i in range(n)])
    stats = {"makespan": st.makespan, "launches": st.launches, "tc_launches": st.tc_launches,
             "peak_bytes": st.peak_bytes, "max_waiting": st.max_waiting, "slots": st.slots}
    return admitted, finished, batch, slot, stats


# ---- the reference's entry functions (include/*_cuda.h) ----------------------
def GPUbfs_vertex(src: int, g: DeviceGraph, dist: np.ndarray) -> float:
    """include/bfs_vertex_cuda.h:61."""
    sec = ctypes.c_double()
    _lib.call("indigo_bfs_vertex", src, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, _pi(dist),
              ctypes.byref(sec))
    return sec.value


def GPUbfs_edge(src: int, g: DeviceGraph, dist: np.ndarray, sp: np.ndarray) -> float:
    """include/bfs_edge_cuda.h:61 (sp on the host)."""
    sec = ctypes.c_double()
    _lib.call("indigo_bfs_edge", src, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, _pi(sp), _pi(dist),
              ctypes.byref(sec))
    return sec.value


def GPUsssp_vertex(src: int, g: DeviceGraph, dist: np.ndarray) -> float:
    """include/sssp_vertex_cuda.h:60."""
    sec = ctypes.c_double()
    _lib.call("indigo_sssp_vertex", src, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, g.d_eweight.ptr,
              _pi(dist), ctypes.byref(sec))
    return sec.value


def GPUsssp_edge(src: int, g: DeviceGraph, dist: np.ndarray, sp: np.ndarray) -> float:
    """include/sssp_edge_cuda.h:60 (sp on the host)."""
    sec = ctypes.c_double()
    _lib.call("indigo_sssp_edge", src, g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, g.d_eweight.ptr, _pi(sp),
              _pi(dist), ctypes.byref(sec))
    return sec.value


def GPUcc_vertex(g: DeviceGraph, label: np.ndarray) -> float:
    """include/cc_vertex_cuda.h:57."""
    sec = ctypes.c_double()
    _lib.call("indigo_cc_vertex", g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, _pi(label), ctypes.byref(sec))
    return sec.value


def GPUcc_edge(g: DeviceGraph, label: np.ndarray, sp: np.ndarray) -> float:
    """include/cc_edge_cuda.h:57 (sp on the host)."""
    sec = ctypes.c_double()
    _lib.call("indigo_cc_edge", g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, _pi(sp), _pi(label),
              ctypes.byref(sec))
    return sec.value


def GPUmis_vertex(g: DeviceGraph, priority: np.ndarray, status: np.ndarray, signed_priority: bool = False) -> float:
    """include/mis_vertex_cuda.h:54."""
    sec = ctypes.c_double()
    _lib.call("indigo_mis_vertex", g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, int(signed_priority),
              priority.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _pi(status), ctypes.byref(sec))
    return sec.value


def GPUmis_edge(g: DeviceGraph, d_sp: DeviceBuffer, priority: np.ndarray, status: np.ndarray,
                signed_priority: bool = False) -> float:
    """include/mis_edge_cuda.h:54 (sp on the device)."""
    sec = ctypes.c_double()
    _lib.call("indigo_mis_edge", g.nodes, g.edges, g.d_nindex.ptr, g.d_nlist.ptr, d_sp.ptr, int(signed_priority),
              priority.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), _pi(status), ctypes.byref(sec))
    return sec.value


def PR_GPU(g: HostGraph, scores: np.ndarray, degree: np.ndarray) -> tuple[float, int]:
    """include/pr_cuda.h:48 (host graph; scores in/out).  Returns (seconds,
    iterations)."""
    sec = ctypes.c_double()
    it = ctypes.c_int()
    _lib.call("indigo_pr", g.nodes, g.edges, _pi(g.nindex), _pi(g.nlist), _pi(degree),
              scores.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(it), ctypes.byref(sec))
    return sec.value, it.value


def GPUtc_vertex(g: DeviceGraph) -> tuple[int, float]:
    """include/tc_vertex_cuda.h:42; returns (count, seconds)."""
    sec = ctypes.c_double()
    c = ctypes.c_longlong()
    _lib.call("indigo_tc_vertex", ctypes.byref(c), g.nodes, g.d_nindex.ptr, g.d_nlist.ptr, ctypes.byref(sec))
    return int(c.value), sec.value


def GPUtc_edge(g: DeviceGraph, d_sp: DeviceBuffer) -> tuple[int, float]:
    """include/tc_edge_cuda.h:42 (sp on the device); returns (count, seconds)."""
    sec = ctypes.c_double()
    c = ctypes.c_longlong()
    _lib.call("indigo_tc_edge", ctypes.byref(c), g.edges, g.d_nindex.ptr, g.d_nlist.ptr, d_sp.ptr,
              ctypes.byref(sec))
    return int(c.value), sec.value
