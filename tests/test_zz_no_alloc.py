"""The allocation-free CC entry (indigo_cc_noalloc) through the C ABI: its
labels must equal indigo_cc's bit for bit, on two calls in a row whose
second graph is smaller than the first (the shared scratch keeps the first
call's contents), on both link orders (Afforest for average degree >= 8, the
plain order below), and a graph above the scratch limit is refused before
any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ig, ctx, dg, out):
    ig._lib.call("indigo_cc_noalloc", ctx.h, dg.nodes, dg.edges, dg.d_nindex.ptr, dg.d_nlist.ptr, out.ptr)


def test_cc_noalloc_matches_cc_on_shrinking_inputs(ig, orc, ctx):
    graphs = [ig.rmat(20, 16, seed=3), ig.road(512, 512, keep_p=0.7, seed=5), ig.rmat(12, 8, seed=4)]
    for g in graphs:
        dg = ig.DeviceGraph(g)
        want = ctx.cc(dg)
        out = ig.DeviceBuffer(4 * g.nodes)
        _noalloc(ig, ctx, dg, out)
        got = out.to_array(np.int32, g.nodes)
        assert np.array_equal(got, want), g.name
        # and once more on the same graph right after
        _noalloc(ig, ctx, dg, out)
        assert np.array_equal(out.to_array(np.int32, g.nodes), want), g.name
        out.close()
        dg.close()
    # the smallest graph against the oracle as well
    g = graphs[-1]
    dg = ig.DeviceGraph(g)
    out = ig.DeviceBuffer(4 * g.nodes)
    _noalloc(ig, ctx, dg, out)
    assert np.array_equal(out.to_array(np.int32, g.nodes), orc.cc(g))


def test_cc_noalloc_refuses_graphs_above_the_limit(ig, ctx):
    g = ig.grid(4, 4)
    dg = ig.DeviceGraph(g)
    out = ig.DeviceBuffer(4 * g.nodes)
    with pytest.raises(ig.IndigoError) as e:
        # the size check comes first: no kernel reads these arrays
        ig._lib.call("indigo_cc_noalloc", ctx.h, ig._lib.CC_NOALLOC_MAX_NODES + 1, dg.edges, dg.d_nindex.ptr,
                     dg.d_nlist.ptr, out.ptr)
    assert e.value.code == ig._lib.ERR_CAPACITY
    assert "scratch limit" in str(e.value)
    # the context still works afterwards
    _noalloc(ig, ctx, dg, out)
    assert np.array_equal(out.to_array(np.int32, g.nodes), np.zeros(g.nodes, np.int32))
