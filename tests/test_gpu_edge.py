"""The edge-centric, topology-driven variants (INDIGO_HINT_VARIANT = edge; the
suite's E_Topo_Push family: one lane per CSR entry over a device edge-source
array) against the oracle, bit for bit, for BFS, SSSP, CC and MIS in both
priority orders, at each workgroup size the autotuner may pick, and on
persistent grids capped to one and to seven workgroups."""
import numpy as np
import pytest

from graphs import catalog, weighted

pytestmark = pytest.mark.gpu


def _set(ig, ctx, variant, block=0, cap=0):
    ctx.set_hint(ig._lib.HINT_VARIANT, variant)
    ctx.set_hint(ig._lib.HINT_BLOCK_SIZE, block)
    ctx.set_hint(ig._lib.HINT_MAX_WORKGROUPS, cap)


def _check_all(ig, orc, ctx, g, tag):
    dg = ig.DeviceGraph(g)
    for src in sorted({0, g.nodes // 2, g.nodes - 1}):
        assert np.array_equal(ctx.bfs(dg, src), orc.bfs(g, src)), (tag, g.name, "bfs", src)
        assert np.array_equal(ctx.sssp(dg, src), orc.sssp(g, src)), (tag, g.name, "sssp", src)
    assert np.array_equal(ctx.cc(dg), orc.cc(g)), (tag, g.name, "cc")
    for signed in (False, True):
        assert np.array_equal(ctx.mis(dg, signed), orc.mis(g, signed)), (tag, g.name, "mis", signed)
    dg.close()


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_edge_variant_catalog(ig, orc, ctx, block):
    _set(ig, ctx, ig._lib.VARIANT_EDGE, block)
    try:
        for g in catalog(ig, weights=True):
            _check_all(ig, orc, ctx, g, block)
    finally:
        _set(ig, ctx, 0)


def test_edge_variant_larger_graphs(ig, orc, ctx):
    _set(ig, ctx, ig._lib.VARIANT_EDGE)
    try:
        for g in (ig.rmat(16, 16, seed=21, wmin=1, wmax=255), ig.road(128, 128, seed=22, wmin=1, wmax=1000)):
            _check_all(ig, orc, ctx, g, "large")
    finally:
        _set(ig, ctx, 0)


@pytest.mark.parametrize("cap", [1, 7])
def test_edge_variant_small_grids(ig, orc, ctx, cap):
    _set(ig, ctx, ig._lib.VARIANT_EDGE, 0, cap)
    try:
        _check_all(ig, orc, ctx, ig.rmat(12, 8, seed=23, wmin=1, wmax=255), cap)
        _check_all(ig, orc, ctx, weighted(ig, ig.grid(40, 25), seed=5), cap)
    finally:
        _set(ig, ctx, 0)


def test_edge_variant_reports_sweeps(ig, ctx):
    """BFS on a 30 x 30 grid from vertex 0 sweeps once per level (59 levels)
    plus the quiet sweep that ends it."""
    g = ig.grid(30, 30)
    _set(ig, ctx, ig._lib.VARIANT_EDGE)
    try:
        dist = ctx.bfs(ig.DeviceGraph(g), 0)
        assert dist.max() == 58
        assert ctx.last_stats[1] == 59
        assert ctx.last_kernel[0] == "edge_bfs"
    finally:
        _set(ig, ctx, 0)
