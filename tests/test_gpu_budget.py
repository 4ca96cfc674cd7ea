"""Bounded device-side waits: a wait budget far below a run's length makes
the persistent kernels stop at their bounded waits (grid barriers, the
one-workgroup phases' hand-back and abort checks) and the call return
INDIGO_ERR_TIMEOUT; every kernel exits normally, and the same context then
gives exact results with the default budget."""
import numpy as np
import pytest

from graphs import weighted

pytestmark = pytest.mark.gpu


def test_wait_budget_stops_long_runs_and_the_context_recovers(ig, orc, ctx):
    road = ig.road(1024, 1024, seed=2, wmin=1, wmax=1000)  # SSSP: about 2900 iterations, 37 ms
    grid = ig.grid(1024, 1024)                             # BFS: 2047 levels, 20 ms
    droad, dgrid = ig.DeviceGraph(road), ig.DeviceGraph(grid)
    try:
        ctx.set_hint(ig._lib.HINT_BUDGET_MS, 1)
        for small in (0, 1):  # one-workgroup phase on, then off (grid barriers only)
            ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, small)
            print(f"budget 1 ms: sssp road1024 small={small}", flush=True)
            with pytest.raises(ig.IndigoError) as e:
                ctx.sssp(droad, 0)
            assert e.value.code == ig._lib.ERR_TIMEOUT
            print(f"budget 1 ms: bfs grid1024 small={small}", flush=True)
            with pytest.raises(ig.IndigoError) as e:
                ctx.bfs(dgrid, 0)
            assert e.value.code == ig._lib.ERR_TIMEOUT
    finally:
        ctx.set_hint(ig._lib.HINT_BUDGET_MS, 0)
        ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, 0)
    print("default budget: sssp road1024, bfs grid1024", flush=True)
    assert np.array_equal(ctx.sssp(droad, 0), orc.sssp(road, 0))
    r, c = np.divmod(np.arange(grid.nodes), 1024)
    assert np.array_equal(ctx.bfs(dgrid, 0), r + c)
    small = weighted(ig, ig.grid(40, 30), seed=5, wmax=9)
    assert np.array_equal(ctx.sssp(ig.DeviceGraph(small), 0), orc.sssp(small, 0))
