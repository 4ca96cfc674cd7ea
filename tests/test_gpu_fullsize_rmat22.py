"""Parity at full size on the benchmark's own graph: R-MAT scale 22, edge
factor 16 (a/b/c/d .57/.19/.19/.05), seed 1, weights [1, 255]; 4,194,304
vertices and 128,307,226 CSR entries, with a 163K-entry hub.

Every algorithm runs through the resident layer exactly as the benchmark
calls it and is compared with the oracle on the same graph: BFS and SSSP
distances bit for bit from three sources (vertex 0, the second-largest hub
and a seeded non-isolated vertex) plus an isolated one, CC labels, MIS
status in both priority orders, PageRank against the FP64 recurrence at the
same iteration count (max-abs <= 1e-6 per vertex, L1 <= 1e-4), and the
triangle count against tests/golden/rmat22_ef16_s1.json (the reference's
own OpenMP TC and the oracle agree on it).  The JSON also holds sha256
digests of the CSR: a generator that drifted would fail here, not compare
against counts of another graph.  The reference harnesses check the same
outputs (include/bfs_vertex_cuda.h:174-184, include/tc_vertex_cuda.h:193-201,
include/pr_cuda.h:139-158)."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rmat22_ef16_s1.json")
PR_MAX_ABS = 1e-6
PR_L1 = 1e-4


@pytest.fixture(scope="module")
def gold():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def rmat22(ig, gold):
    print("generating rmat22 ef16 seed 1", flush=True)
    g = ig.rmat(gold["scale"], gold["edge_factor"], a=gold["a"], b=gold["b"], c=gold["c"], seed=gold["seed"],
                wmin=gold["wmin"], wmax=gold["wmax"])
    assert (g.nodes, g.edges, int(g.degree.max())) == (gold["nodes"], gold["edges"], gold["max_degree"])
    for name, arr in (("nindex", g.nindex), ("nlist", g.nlist), ("eweight", g.eweight)):
        digest = hashlib.sha256(np.ascontiguousarray(arr, dtype="<i4").tobytes()).hexdigest()
        assert digest == gold["sha256"][name], f"generator drift: {name} digest differs from the golden graph"
    dg = ig.DeviceGraph(g)
    yield g, dg
    dg.close()


def _sources(g):
    deg = g.degree
    hubs = np.argsort(deg, kind="stable")[::-1]
    rng = np.random.default_rng(22)
    nonisolated = np.flatnonzero(deg > 0)
    isolated = np.flatnonzero(deg == 0)
    return [0, int(hubs[1]), int(rng.choice(nonisolated)), int(isolated[len(isolated) // 2])]


def test_rmat22_bfs(orc, ctx, rmat22):
    g, dg = rmat22
    for src in _sources(g):
        print(f"rmat22 bfs from {src}", flush=True)
        assert np.array_equal(ctx.bfs(dg, src), orc.bfs(g, src)), src


def test_rmat22_sssp(orc, ctx, rmat22):
    g, dg = rmat22
    for src in _sources(g):
        print(f"rmat22 sssp from {src}", flush=True)
        assert np.array_equal(ctx.sssp(dg, src), orc.sssp(g, src)), src


def test_rmat22_cc(orc, ctx, rmat22):
    g, dg = rmat22
    assert np.array_equal(ctx.cc(dg), orc.cc(g))


@pytest.mark.parametrize("signed", [False, True])
def test_rmat22_mis(orc, ctx, rmat22, signed):
    g, dg = rmat22
    assert np.array_equal(ctx.mis(dg, signed), orc.mis(g, signed))


def test_rmat22_pagerank(orc, ctx, rmat22):
    g, dg = rmat22
    scores, iters = ctx.pagerank(dg)
    ref, _ = orc.pr64(g, iters)
    err = np.abs(scores.astype(np.float64) - ref)
    assert err.max() <= PR_MAX_ABS, err.max()
    assert err.sum() <= PR_L1, err.sum()
    _, it32 = orc.pr_ref32(g)
    assert abs(it32 - iters) <= 1, (it32, iters)


def test_rmat22_reordered(ig, orc, ctx, rmat22):
    """BFS, SSSP and PageRank on the degree-ordered copy of the graph, read
    back in the original ids, as bench.py's reordered step runs them."""
    g, _ = rmat22
    rg = ig.ReorderedGraph(g)
    try:
        for src in _sources(g)[:3]:
            print(f"rmat22 reordered bfs and sssp from {src}", flush=True)
            assert np.array_equal(ctx.bfs_reordered(rg, src), orc.bfs(g, src)), src
            assert np.array_equal(ctx.sssp_reordered(rg, src), orc.sssp(g, src)), src
        scores, iters = ctx.pagerank_reordered(rg)
        ref, _ = orc.pr64(g, iters)
        err = np.abs(scores.astype(np.float64) - ref)
        assert err.max() <= PR_MAX_ABS, err.max()
        assert err.sum() <= PR_L1, err.sum()
    finally:
        rg.close()


def test_rmat22_tc(ctx, rmat22, gold):
    _, dg = rmat22
    assert ctx.tc(dg) == gold["triangles"]


def test_rmat22_dropin_bfs_and_tc(ig, orc, rmat22, gold):
    """The drop-in entries (host output, reference timing) at full size."""
    g, dg = rmat22
    dist = np.empty(g.nodes, np.int32)
    assert ig.GPUbfs_vertex(0, dg, dist) > 0
    assert np.array_equal(dist, orc.bfs(g, 0))
    count, secs = ig.GPUtc_vertex(dg)
    assert count == gold["triangles"] and secs > 0
