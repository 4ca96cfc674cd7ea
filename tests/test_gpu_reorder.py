"""BFS, SSSP and PageRank on degree-ordered device copies (ReorderedGraph:
indigo_graph_reorder on the host, the kernels on the copy, indigo_unpermute
back to the original ids) against the oracle on the original graph: BFS and
SSSP bit for bit from several sources, PageRank within max-abs 1e-6 and
L1 1e-4 of the FP64 recurrence at the same iteration count.  Also the
argument checks of indigo_unpermute."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PR_MAX_ABS = 1e-6
PR_L1 = 1e-4


def _graphs(ig):
    return [
        ig.rmat(16, 16, seed=21, wmin=1, wmax=255),
        ig.road(128, 96, seed=22, wmin=1, wmax=1000),
        ig.grid(64, 33, wmin=1, wmax=9),
        ig.erdos_renyi(20000, 80000, seed=23, wmin=1, wmax=255),
    ]


def test_reordered_equals_oracle(ig, orc, ctx):
    for g in _graphs(ig):
        rg = ig.ReorderedGraph(g)
        try:
            for s in (0, g.nodes // 3, g.nodes - 1):
                print(f"{g.name} from {s}", flush=True)
                assert np.array_equal(ctx.bfs_reordered(rg, s), orc.bfs(g, s)), (g.name, s)
                assert np.array_equal(ctx.sssp_reordered(rg, s), orc.sssp(g, s)), (g.name, s)
            scores, iters = ctx.pagerank_reordered(rg)
            ref, _ = orc.pr64(g, iters)
            err = np.abs(scores.astype(np.float64) - ref)
            assert err.max() <= PR_MAX_ABS and err.sum() <= PR_L1, (g.name, err.max(), err.sum())
        finally:
            rg.close()


def test_unpermute_arguments(ig, ctx):
    vals = np.arange(100, 116, dtype=np.int32)
    perm = np.random.default_rng(3).permutation(16).astype(np.int32)
    a = ig.DeviceBuffer.from_array(vals)
    p = ig.DeviceBuffer.from_array(perm)
    b = ig.DeviceBuffer(64)
    ctx.unpermute(16, p, a, b)
    assert np.array_equal(b.to_array(np.int32, 16), vals[perm])
    with pytest.raises(ig.IndigoError):
        ctx.unpermute(16, p, a, a)  # in place
    with pytest.raises(ig.IndigoError):
        ctx.unpermute(-1, p, a, b)
    ctx.unpermute(0, p, a, b)  # nothing to do
    for x in (a, b, p):
        x.close()
