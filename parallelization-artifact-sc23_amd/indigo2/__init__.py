"""indigo2: MI355X (gfx950) implementation of Indigo2's six graph kernels.

Python host mirror over the C ABI of lib/libindigo2_hip.so.  The package
directory ``parallelization-artifact-sc23_amd`` is put on ``sys.path`` by
callers (tests, bench.py, __graft_entry__.py) so that ``import indigo2``
works; importing does not load the shared library until first use.
"""
from . import _lib
from ._lib import IndigoError, load
from .graph import DeviceBuffer, DeviceGraph, HostGraph, ReorderedGraph, erdos_renyi, from_edges, grid, read_egr, rmat, road
from .kernels import (
    EXCLUDED,
    INCLUDED,
    INF,
    UNDECIDED,
    Context,
    Query,
    serve,
    GPUbfs_edge,
    GPUbfs_vertex,
    GPUcc_edge,
    GPUcc_vertex,
    GPUmis_edge,
    GPUmis_vertex,
    GPUsssp_edge,
    GPUsssp_vertex,
    GPUtc_edge,
    GPUtc_vertex,
    PR_GPU,
)

__all__ = [
    "_lib", "IndigoError", "load", "DeviceBuffer", "DeviceGraph", "HostGraph", "ReorderedGraph", "erdos_renyi", "from_edges",
    "grid", "read_egr", "rmat", "road", "EXCLUDED", "INCLUDED", "INF", "UNDECIDED", "Context", "Query", "serve", "GPUbfs_edge",
    "GPUbfs_vertex", "GPUcc_edge", "GPUcc_vertex", "GPUmis_edge", "GPUmis_vertex", "GPUsssp_edge",
    "GPUsssp_vertex", "GPUtc_edge", "GPUtc_vertex", "PR_GPU",
]
