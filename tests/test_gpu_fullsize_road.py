"""Parity at full size on the high-diameter configuration: the 4096 x 4096
road-like grid (16,777,216 vertices; each 4-neighbour edge kept with
probability 0.85, a diagonal edge for 2% of the vertices, weights uniform
in [1, 1000], seed 1; 57,695,488 CSR entries, about 7,600 BFS levels).

BFS, SSSP and CC from vertex 0, exactly as the benchmark calls them, are
compared bit for bit with the oracle; BFS and SSSP also from a second
source in the far corner, and once more with the one-workgroup phase off so
that the grid-wide path runs every level."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def road4096(ig):
    print("generating road 4096x4096", flush=True)
    g = ig.road(4096, 4096, keep_p=0.85, diag_p=0.02, seed=1, wmin=1, wmax=1000)
    assert (g.nodes, g.edges) == (16777216, 57695488)
    dg = ig.DeviceGraph(g)
    yield g, dg
    dg.close()


@pytest.fixture(scope="module")
def oracle_from_0(orc, road4096):
    g, _ = road4096
    return {"bfs": orc.bfs(g, 0), "sssp": orc.sssp(g, 0)}


def test_road4096_bfs(orc, ctx, road4096, oracle_from_0):
    g, dg = road4096
    assert np.array_equal(ctx.bfs(dg, 0), oracle_from_0["bfs"])
    far = g.nodes - 1 - 4096 // 2
    assert np.array_equal(ctx.bfs(dg, far), orc.bfs(g, far))


def test_road4096_sssp(orc, ctx, road4096, oracle_from_0):
    g, dg = road4096
    assert np.array_equal(ctx.sssp(dg, 0), oracle_from_0["sssp"])
    far = g.nodes - 1 - 4096 // 2
    assert np.array_equal(ctx.sssp(dg, far), orc.sssp(g, far))


def test_road4096_cc(orc, ctx, road4096):
    g, dg = road4096
    labels = ctx.cc(dg)
    assert np.array_equal(labels, orc.cc(g))
    # vertex 0 lies in the giant component the benchmark reports
    assert np.count_nonzero(labels == labels[0]) == 16768341


def test_road4096_grid_wide_levels(ig, ctx, road4096, oracle_from_0):
    g, dg = road4096
    ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, 1)
    try:
        assert np.array_equal(ctx.bfs(dg, 0), oracle_from_0["bfs"])
        assert np.array_equal(ctx.sssp(dg, 0), oracle_from_0["sssp"])
    finally:
        ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, 0)
