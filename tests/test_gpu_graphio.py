"""Staged graph uploads, mapped .egr files and the device-side edge-source
array: the device copies equal the host arrays bit for bit, and the kernels
give the oracle's results on graphs that arrive this way."""
import numpy as np
import pytest

from graphs import catalog

pytestmark = pytest.mark.gpu


def _same_arrays(dg, g):
    assert np.array_equal(dg.d_nindex.to_array(np.int32, g.nodes + 1), g.nindex)
    assert np.array_equal(dg.d_nlist.to_array(np.int32, g.edges), g.nlist)
    if g.eweight is None:
        assert dg.d_eweight is None
    else:
        assert np.array_equal(dg.d_eweight.to_array(np.int32, g.edges), g.eweight)


def test_staged_upload_spans_several_chunks(ig, orc, ctx):
    # 128 MB of nlist and of weights: several 64 MiB staging chunks each
    g = ig.rmat(20, 16, seed=5, wmin=1, wmax=255)
    dg = ig.DeviceGraph.upload(ctx, g)
    assert dg.upload_seconds > 0
    _same_arrays(dg, g)
    assert np.array_equal(ctx.bfs(dg, 0), orc.bfs(g, 0))
    dg.close()


def test_mapped_file_upload(ig, orc, ctx, tmp_path):
    for g in (ig.rmat(14, 16, seed=9, wmin=1, wmax=255), ig.grid(40, 50)):
        p = tmp_path / f"{g.name}.egr"
        g.write_egr(str(p))
        dg = ig.DeviceGraph.from_file(ctx, str(p))
        _same_arrays(dg, g)
        assert np.array_equal(ctx.cc(dg), orc.cc(g))
        if g.eweight is not None:
            assert np.array_equal(ctx.sssp(dg, 0), orc.sssp(g, 0))
        dg.close()


def test_device_sp_equals_host_sp(ig, ctx):
    """Isolated vertices at the ends and inside, a single vertex, R-MAT."""
    lone = ig.from_edges(6, [1, 2], [2, 3])  # vertices 0, 4, 5 have no entries
    single = ig.from_edges(1, [], [])
    for g in [lone, single, ig.rmat(16, 16, seed=2)] + catalog(ig):
        dg = ig.DeviceGraph(g)
        sp = dg.build_sp(ctx)
        assert np.array_equal(sp.to_array(np.int32, g.edges), g.sp()), g.name
        dg.close()


def test_sp_feeds_the_edge_entries(ig, orc, ctx):
    g = ig.rmat(15, 16, seed=6)
    dg = ig.DeviceGraph(g)
    d_sp = dg.build_sp(ctx)
    count, _ = ig.GPUtc_edge(dg, d_sp)
    assert count == orc.tc(g)
    dg.close()
