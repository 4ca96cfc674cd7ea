"""The degree-ordered relabelling (indigo_graph_reorder, host code) against
what it must preserve: a permutation whose new ids run by degree, largest
first, ties in id order; a simple sorted symmetric CSR with every edge and
weight carried over under the permutation; and the oracle's BFS, SSSP,
triangle count and FP64 PageRank read back through the permutation equal
to those of the original graph.  Bad input is refused."""
import numpy as np
import pytest

from graphs import cliques


def _graphs(ig):
    return [
        ig.rmat(12, 16, seed=3, wmin=1, wmax=255),
        ig.road(40, 30, seed=4, wmin=1, wmax=1000),
        ig.grid(17, 9, wmin=1, wmax=9),
        ig.erdos_renyi(3000, 9000, seed=5, wmin=1, wmax=255),
        cliques(ig, [3, 5, 9, 17, 4, 4], seed=6).with_weights(None),
        ig.from_edges(50, [1, 2, 3, 7], [2, 3, 1, 8], name="isolated-vertices"),
    ]


def _edges(g):
    src = np.repeat(np.arange(g.nodes, dtype=np.int64), g.degree)
    w = g.eweight if g.eweight is not None else np.zeros(g.edges, np.int32)
    return src, g.nlist.astype(np.int64), w


def test_reorder_is_a_degree_ordered_isomorphism(ig):
    for g in _graphs(ig):
        r, perm = g.reordered()
        assert (r.nodes, r.edges) == (g.nodes, g.edges), g.name
        assert np.array_equal(np.sort(perm), np.arange(g.nodes)), g.name
        order = np.empty_like(perm)
        order[perm] = np.arange(g.nodes)
        d = g.degree[order]
        assert np.all(d[:-1] >= d[1:]), g.name
        ties = d[:-1] == d[1:]
        assert np.all(order[:-1][ties] < order[1:][ties]), g.name
        assert r.check() == 0, g.name
        s0, d0, w0 = _edges(g)
        s1, d1, w1 = _edges(r)
        a = np.lexsort((w0, perm[d0], perm[s0]))
        assert np.array_equal(perm[s0][a], s1) and np.array_equal(perm[d0][a], d1), g.name
        if g.eweight is not None:
            assert np.array_equal(w0[a], w1), g.name
        else:
            assert r.eweight is None


def test_reorder_keeps_bfs_sssp_tc_pagerank(ig, orc):
    for g in _graphs(ig):
        r, perm = g.reordered()
        for s in (0, g.nodes // 2, g.nodes - 1):
            assert np.array_equal(orc.bfs(r, int(perm[s]))[perm], orc.bfs(g, s)), (g.name, s)
            if g.eweight is not None:
                assert np.array_equal(orc.sssp(r, int(perm[s]))[perm], orc.sssp(g, s)), (g.name, s)
        assert orc.tc(r) == orc.tc(g), g.name
        a, _ = orc.pr64(g, 20)
        b, _ = orc.pr64(r, 20)
        assert np.abs(b[perm] - a).max() <= 1e-12, g.name


def test_reorder_refuses_bad_input(ig):
    g = ig.grid(4, 4)
    bad = ig.HostGraph(g.nodes, g.edges, g.nindex, np.where(g.nlist == 5, 99, g.nlist).astype(np.int32))
    with pytest.raises(ig.IndigoError):
        bad.reordered()
    nix = g.nindex.copy()
    nix[3] = nix[4] + 1
    with pytest.raises(ig.IndigoError):
        ig.HostGraph(g.nodes, g.edges, nix, g.nlist).reordered()
