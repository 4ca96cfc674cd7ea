"""GPU parity of the batched engine (indigo_batch): every query of a mixed
batch - many graphs of uneven sizes, multi-source BFS (up to 64 sources, with
duplicates), multi-source SSSP (up to 32), CC and MIS - must equal the
single-query resident kernels on the same graph, row by row and bit for bit;
a few rows are also checked against the oracle directly."""
import numpy as np
import pytest

from graphs import catalog, clique, weighted

pytestmark = pytest.mark.gpu


def _mixed_graphs(ig):
    gs = [weighted(ig, g, seed=i + 1) for i, g in enumerate(catalog(ig))]
    gs += [ig.rmat(s, 8, seed=40 + s, wmin=1, wmax=255) for s in (8, 11, 14)]
    gs += [ig.erdos_renyi(5000, 20000, seed=5, wmin=1, wmax=255), weighted(ig, clique(ig, 70), seed=3)]
    return gs


def _queries(ig, gs, dgs, rng, nbfs=64, nsssp=32):
    qs = []
    for gi, (g, dg) in enumerate(zip(gs, dgs)):
        nb = int(rng.integers(1, nbfs + 1))
        srcs = rng.integers(0, g.nodes, nb)
        srcs[: min(2, nb)] = 0  # duplicates and vertex 0 included
        qs.append(ig.Query("bfs", gi, dg, srcs))
        ns = int(rng.integers(1, nsssp + 1))
        qs.append(ig.Query("sssp", gi, dg, rng.integers(0, g.nodes, ns)))
        qs.append(ig.Query("cc", gi, dg))
        qs.append(ig.Query("mis", gi, dg))
    return qs


def _check_against_plain(ctx, gs, dgs, qs):
    for q in qs:
        g, dg = gs[q.graph], dgs[q.graph]
        got = q.result()
        if q.name == "bfs":
            for i, s in enumerate(q.sources):
                assert np.array_equal(got[i], ctx.bfs(dg, s)), (g.name, "bfs", s)
        elif q.name == "sssp":
            for i, s in enumerate(q.sources):
                assert np.array_equal(got[i], ctx.sssp(dg, s)), (g.name, "sssp", s)
        elif q.name == "cc":
            assert np.array_equal(got, ctx.cc(dg)), (g.name, "cc")
        else:
            assert np.array_equal(got, ctx.mis(dg)), (g.name, "mis")


def test_batch_mixed_matches_plain(ig, orc, ctx):
    rng = np.random.default_rng(7)
    gs = _mixed_graphs(ig)
    dgs = [ig.DeviceGraph(g) for g in gs]
    qs = _queries(ig, gs, dgs, rng)
    print(f"batch of {len(qs)} queries over {len(gs)} graphs", flush=True)
    ctx.batch(dgs, qs)
    _check_against_plain(ctx, gs, dgs, qs)
    # direct oracle rows
    for q in qs[:8]:
        g = gs[q.graph]
        if q.name == "bfs":
            assert np.array_equal(q.result()[0], orc.bfs(g, q.sources[0]))
        elif q.name == "sssp":
            assert np.array_equal(q.result()[0], orc.sssp(g, q.sources[0]))
        elif q.name == "cc":
            assert np.array_equal(q.result(), orc.cc(g))
        else:
            assert np.array_equal(q.result(), orc.mis(g))


def test_batch_large_graph_uses_gangs(ig, ctx):
    # rmat 17 ef8 has ~2M CSR entries: gangs of 8 workgroups
    g = ig.rmat(17, 8, seed=3, wmin=1, wmax=255)
    dg = ig.DeviceGraph(g)
    rng = np.random.default_rng(3)
    qs = [ig.Query("bfs", 0, dg, rng.integers(0, g.nodes, 64)), ig.Query("sssp", 0, dg, rng.integers(0, g.nodes, 8)),
          ig.Query("cc", 0, dg), ig.Query("mis", 0, dg)]
    print("batch on rmat17", flush=True)
    ctx.batch([dg], qs)
    _check_against_plain(ctx, [g], [dg], qs)


@pytest.mark.parametrize("cap", [1, 8])
def test_batch_capped_grid(ig, ctx, cap):
    rng = np.random.default_rng(11)
    gs = [ig.rmat(12, 8, seed=1, wmin=1, wmax=255), ig.grid(40, 30, seed=2, wmin=1, wmax=100),
          ig.rmat(16, 8, seed=2, wmin=1, wmax=255)]
    dgs = [ig.DeviceGraph(g) for g in gs]
    qs = _queries(ig, gs, dgs, rng, nbfs=8, nsssp=4)
    print(f"batch capped at {cap} workgroups", flush=True)
    ctx.batch(dgs, qs, max_workgroups=cap)
    _check_against_plain(ctx, gs, dgs, qs)


def test_batch_rejects_bad_queries(ig, ctx):
    g = ig.grid(4, 4)
    dg = ig.DeviceGraph(g)
    q = ig.Query("bfs", 0, dg, [0])
    q.sources = []
    with pytest.raises(ig.IndigoError):
        ctx.batch([dg], [q])
    q2 = ig.Query("bfs", 1, dg, [0])
    with pytest.raises(ig.IndigoError):
        ctx.batch([dg], [q2])
    q3 = ig.Query("sssp", 0, dg, [0])
    with pytest.raises(ig.IndigoError):
        ctx.batch([dg], [q3])  # the grid graph has no weights
    q4 = ig.Query("bfs", 0, dg, [16])
    with pytest.raises(ig.IndigoError):
        ctx.batch([dg], [q4])


def test_batch_pagerank(ig, orc, ctx):
    """PR tasks: FP32 scores within 1e-6 per vertex and 1e-4 in L1 of the FP64
    oracle at the same iteration count; the count within +-1 of the FP32
    reference recurrence (both as for the single-query kernel)."""
    from test_gpu_kernels import _pr_check

    gs = [g for g in catalog(ig)] + [ig.rmat(15, 8, seed=9), ig.rmat(17, 8, seed=4)]
    dgs = [ig.DeviceGraph(g) for g in gs]
    qs = [ig.Query("pr", i, dgs[i]) for i in range(len(gs))]
    print(f"batch of {len(qs)} PR queries", flush=True)
    ctx.batch(dgs, qs)
    for q in qs:
        g = gs[q.graph]
        scores, iters = q.result()
        assert 1 <= iters <= 100, (g.name, iters)
        _pr_check(ig, orc, g, scores, iters)
        _, it32 = orc.pr_ref32(g)
        assert abs(it32 - iters) <= 1, (g.name, it32, iters)


@pytest.mark.parametrize("cap", [0, 100])
def test_batch_big_gangs(ig, orc, ctx, cap):
    """R-MAT 20 (31M CSR entries): PR, CC and MIS take gangs of 128
    workgroups and the two-level gang barrier, uncapped and with the grid
    capped at 100 (gangs of 100: not a power of two); BFS / SSSP 64 / 128."""
    from test_gpu_kernels import _pr_check

    g = ig.rmat(20, 16, seed=12, wmin=1, wmax=255)
    dg = ig.DeviceGraph(g)
    rng = np.random.default_rng(12)
    qs = [ig.Query("bfs", 0, dg, rng.integers(0, g.nodes, 16)), ig.Query("sssp", 0, dg, rng.integers(0, g.nodes, 4)),
          ig.Query("cc", 0, dg), ig.Query("mis", 0, dg), ig.Query("pr", 0, dg)]
    print(f"batch on rmat20, cap {cap}", flush=True)
    ctx.batch([dg], qs, max_workgroups=cap)
    _check_against_plain(ctx, [g], [dg], qs[:4])
    scores, iters = qs[4].result()
    _pr_check(ig, orc, g, scores, iters)


@pytest.mark.parametrize("gang", [1, 3, 37])
def test_batch_explicit_gangs(ig, ctx, gang):
    """Per-query gang sizes (the autotuner's knob, indigo_query.gang): any
    size from 1 to the grid, odd ones included, gives the same results."""
    g = ig.rmat(16, 8, seed=21, wmin=1, wmax=255)
    road = ig.road(120, 100, keep_p=0.8, seed=4, wmin=1, wmax=1000)
    gs = [g, road]
    dgs = [ig.DeviceGraph(x) for x in gs]
    rng = np.random.default_rng(gang)
    qs = []
    for gi, x in enumerate(gs):
        qs += [ig.Query("bfs", gi, dgs[gi], rng.integers(0, x.nodes, 32), gang=gang),
               ig.Query("sssp", gi, dgs[gi], rng.integers(0, x.nodes, 8), gang=gang),
               ig.Query("cc", gi, dgs[gi], gang=gang), ig.Query("mis", gi, dgs[gi], gang=gang)]
    print(f"batch with gangs of {gang}", flush=True)
    ctx.batch(dgs, qs)
    _check_against_plain(ctx, gs, dgs, qs)


def test_batch_rejects_bad_gang(ig, ctx):
    g = ig.grid(4, 4)
    dg = ig.DeviceGraph(g)
    for bad in (-1, ig._lib.MAX_GANG + 1):
        with pytest.raises(ig.IndigoError):
            ctx.batch([dg], [ig.Query("cc", 0, dg, gang=bad)])
