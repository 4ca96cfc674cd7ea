"""CC and TC through captured HIP graphs (INDIGO_HINT_GRAPH): the first call
with a given graph, output and workspace captures the kernel sequence, later
calls replay it.  Every replay is checked against the eager result, with the
program's own eager work (other algorithms on the same context, host
synchronisation through the downloads) between the replays; a new output
buffer, another graph and a released workspace each force a fresh capture."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs(ig):
    return [ig.rmat(14, 16, seed=41), ig.road(128, 128, seed=42), ig.grid(33, 17), ig.rmat(9, 8, seed=43)]


def test_cc_and_tc_replays_equal_eager(ig, orc, ctx):
    gs = _graphs(ig)
    dgs = [ig.DeviceGraph(g) for g in gs]
    want_cc = [ctx.cc(dg) for dg in dgs]
    want_tc = [ctx.tc(dg) for dg in dgs]
    for g, w in zip(gs, want_cc):
        assert np.array_equal(w, orc.cc(g)), g.name
    outs = [ig.DeviceBuffer(4 * g.nodes) for g in gs]
    ctx.set_hint(ig._lib.HINT_GRAPH, 1)
    try:
        for rep in range(4):
            for i, (g, dg) in enumerate(zip(gs, dgs)):
                outs[i].to_array(np.int32, 1)  # scribble-free host round trip between replays
                ctx.cc_into(dg, outs[i])
                assert np.array_equal(outs[i].to_array(np.int32, g.nodes), want_cc[i]), (rep, g.name, "cc")
                assert ctx.tc(dg) == want_tc[i], (rep, g.name, "tc")
                # eager work of another algorithm on the same context between replays
                ctx.set_hint(ig._lib.HINT_GRAPH, 0)
                ctx.bfs(dg, 0)
                ctx.set_hint(ig._lib.HINT_GRAPH, 1)
        # a different output buffer and a released workspace capture anew
        other = ig.DeviceBuffer(4 * gs[0].nodes)
        ctx.cc_into(dgs[0], other)
        assert np.array_equal(other.to_array(np.int32, gs[0].nodes), want_cc[0])
        ctx.trim()
        for rep in range(2):
            ctx.cc_into(dgs[1], outs[1])
            assert np.array_equal(outs[1].to_array(np.int32, gs[1].nodes), want_cc[1]), rep
            assert ctx.tc(dgs[1]) == want_tc[1], rep
    finally:
        ctx.set_hint(ig._lib.HINT_GRAPH, 0)
    for b in outs:
        b.close()
    for d in dgs:
        d.close()
