"""Every setting stored in the autotuner's database (indigo2/tuned_mi355x.json),
as the benchmark applies it, against the oracle: for each graph class and
algorithm, the tuned single-query hints (INDIGO_HINT_VARIANT, _BLOCK_SIZE,
_BLOCKS_PER_CU, _SSSP_DELTA, _BFS_DIRECTION, _SMALL_FRONTIER) on a graph of
that class, and for each class's batched gang scale a batch whose queries
run with the scaled gangs.  Integer outputs bit for bit; PageRank within
max-abs 1e-6 and L1 1e-4 of the FP64 recurrence at the same iteration count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graph_of(ig, klass):
    shape, size = klass.split("-")
    if shape == "deep":
        side = {"small": 120, "medium": 300, "large": 1024}[size]
        return ig.road(side, side, seed=31, wmin=1, wmax=1000)
    if shape == "skewed":
        return ig.rmat({"small": 12, "medium": 16, "large": 20}[size], 16, seed=32, wmin=1, wmax=255)
    n = {"small": 20000, "medium": 200000, "large": 1 << 20}[size]
    return ig.erdos_renyi(n, 4 * n, seed=33, wmin=1, wmax=255)


def _db():
    from indigo2.autotune import DEFAULT_DB, Tuning

    return Tuning.load(DEFAULT_DB)


def _single_cases():
    db = _db()
    return sorted((k, a) for k, algs in db.db["single"].items() for a, e in algs.items() if e["hints"])


def _batched_cases():
    db = _db()
    return sorted((k, a) for k, algs in db.db["batched"].items() for a, e in algs.items() if e["gang_scale"] != 1.0)


def test_database_has_tuned_settings():
    assert _single_cases() or _batched_cases()


@pytest.mark.parametrize("klass,alg", _single_cases())
def test_tuned_single_query_hints_match_oracle(ig, orc, ctx, klass, alg):
    db = _db()
    g = _graph_of(ig, klass)
    dg = ig.DeviceGraph(g)
    db.apply(ctx, alg, klass)
    try:
        if alg == "bfs":
            assert np.array_equal(ctx.bfs(dg, 0), orc.bfs(g, 0))
        elif alg == "sssp":
            assert np.array_equal(ctx.sssp(dg, 0), orc.sssp(g, 0))
        elif alg == "cc":
            assert np.array_equal(ctx.cc(dg), orc.cc(g))
        elif alg == "mis":
            assert np.array_equal(ctx.mis(dg), orc.mis(g))
        elif alg == "pr":
            scores, iters = ctx.pagerank(dg)
            ref, _ = orc.pr64(g, iters)
            err = np.abs(scores.astype(np.float64) - ref)
            assert err.max() <= 1e-6 and err.sum() <= 1e-4
    finally:
        for k in db.hints(alg, klass):
            ctx.set_hint(k, 0)
    dg.close()


@pytest.mark.parametrize("klass,alg", _batched_cases())
def test_tuned_batched_gangs_match_oracle(ig, orc, ctx, klass, alg):
    db = _db()
    g = _graph_of(ig, klass)
    dg = ig.DeviceGraph(g)
    gang = db.gang(alg, klass, g.nodes, g.edges)
    srcs = [0, g.nodes // 3, g.nodes - 1]
    q = ig.Query(alg, 0, dg, srcs if alg in ("bfs", "sssp") else (), gang=gang)
    ctx.batch([dg], [q])
    got = q.result()
    if alg == "bfs":
        for r, s in enumerate(srcs):
            assert np.array_equal(got[r], orc.bfs(g, s)), (klass, gang, s)
    elif alg == "sssp":
        for r, s in enumerate(srcs):
            assert np.array_equal(got[r], orc.sssp(g, s)), (klass, gang, s)
    elif alg == "cc":
        assert np.array_equal(got, orc.cc(g)), (klass, gang)
    else:
        assert np.array_equal(got, orc.mis(g)), (klass, gang)
    dg.close()
