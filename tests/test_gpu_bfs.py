"""BFS parity on the GPU: HIP kernels vs the CPU oracle (bit-exact int32
levels), through the resident layer and the drop-in entries."""
import numpy as np
import pytest

from graphs import catalog

pytestmark = pytest.mark.gpu


def _dev(ig, g):
    return ig.DeviceGraph(g)


def test_bfs_catalog(ig, orc, ctx):
    for g in catalog(ig):
        dg = _dev(ig, g)
        for src in sorted({0, g.nodes // 2, g.nodes - 1}):
            got = ctx.bfs(dg, src)
            want = orc.bfs(g, src)
            assert np.array_equal(got, want), (g.name, src)


@pytest.mark.parametrize("direction,small", [(1, 1), (2, 1), (1, 0), (0, 0)])
def test_bfs_modes(ig, orc, ctx, direction, small):
    ctx.set_hint(ig._lib.HINT_BFS_DIRECTION, direction)
    ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, small)
    try:
        for g in (ig.rmat(14, 16, seed=5), ig.grid(200, 300), ig.road(128, 128, seed=9)):
            dg = _dev(ig, g)
            got = ctx.bfs(dg, 0)
            assert np.array_equal(got, orc.bfs(g, 0)), (g.name, direction, small)
    finally:
        ctx.set_hint(ig._lib.HINT_BFS_DIRECTION, 0)
        ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, 0)


def test_bfs_grid_1024_closed_form(ig, ctx):
    """1024x1024 4-neighbour grid from vertex 0: dist(r, c) = r + c."""
    g = ig.grid(1024, 1024)
    got = ctx.bfs(_dev(ig, g), 0)
    r, c = np.divmod(np.arange(g.nodes), 1024)
    assert np.array_equal(got, (r + c).astype(np.int32))


def test_bfs_rmat18_properties(ig, orc, ctx):
    g = ig.rmat(18, 16, seed=1)
    got = ctx.bfs(_dev(ig, g), 0)
    assert np.array_equal(got, orc.bfs(g, 0))


def test_bfs_dropin_entries(ig, orc):
    g = ig.rmat(12, 16, seed=3)
    dg = _dev(ig, g)
    dist = np.empty(g.nodes, np.int32)
    secs = ig.GPUbfs_vertex(0, dg, dist)
    assert secs > 0
    assert np.array_equal(dist, orc.bfs(g, 0))
    dist2 = np.empty(g.nodes, np.int32)
    ig.GPUbfs_edge(5, dg, dist2, g.sp())
    assert np.array_equal(dist2, orc.bfs(g, 5))


def test_bfs_bad_source(ig, ctx):
    g = ig.grid(4, 4)
    dg = _dev(ig, g)
    with pytest.raises(ig.IndigoError) as e:
        ctx.bfs(dg, 16)
    assert e.value.code == ig._lib.ERR_INVALID_ARG
