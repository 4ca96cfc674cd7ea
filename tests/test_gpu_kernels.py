"""Parity of SSSP, CC, MIS, PageRank and TC on the GPU against the CPU oracle:
bit-exact for the integer kernels, and for PageRank max-abs <= 1e-6 per vertex
and L1 <= 1e-4 against the FP64 oracle run for the same iteration count."""
import numpy as np
import pytest

from graphs import catalog, clique, cliques, weighted

pytestmark = pytest.mark.gpu

PR_MAX_ABS = 1e-6
PR_L1 = 1e-4


def test_sssp_catalog(ig, orc, ctx):
    for g in catalog(ig, weights=True):
        dg = ig.DeviceGraph(g)
        for src in sorted({0, g.nodes // 3, g.nodes - 1}):
            assert np.array_equal(ctx.sssp(dg, src), orc.sssp(g, src)), (g.name, src)


@pytest.mark.parametrize("delta", [0, 1, 7, 100000])
def test_sssp_delta_hints(ig, orc, ctx, delta):
    ctx.set_hint(ig._lib.HINT_SSSP_DELTA, delta)
    try:
        for g in (ig.rmat(14, 16, seed=3, wmin=1, wmax=255), ig.road(96, 96, seed=2, wmin=1, wmax=1000)):
            dg = ig.DeviceGraph(g)
            assert np.array_equal(ctx.sssp(dg, 0), orc.sssp(g, 0)), (g.name, delta)
    finally:
        ctx.set_hint(ig._lib.HINT_SSSP_DELTA, 0)


@pytest.mark.parametrize("small", [0, 1])
@pytest.mark.parametrize("delta", [0, 1, 300, 100000])
def test_sssp_deep_graphs(ig, orc, ctx, delta, small):
    """Deep sparse graphs run most iterations in the one-workgroup phase
    (small 0), and with it most far-pile splits (delta 1: thousands of them;
    100000: almost none), or all on the grid (small 1); every source's
    distances equal the oracle's."""
    ctx.set_hint(ig._lib.HINT_SSSP_DELTA, delta)
    ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, small)
    try:
        for g in (ig.road(256, 256, seed=7, wmin=1, wmax=1000), weighted(ig, ig.grid(300, 200), seed=3, wmax=50)):
            dg = ig.DeviceGraph(g)
            for src in (0, g.nodes // 2 + 17, g.nodes - 1):
                assert np.array_equal(ctx.sssp(dg, src), orc.sssp(g, src)), (g.name, delta, small, src)
    finally:
        ctx.set_hint(ig._lib.HINT_SSSP_DELTA, 0)
        ctx.set_hint(ig._lib.HINT_SMALL_FRONTIER, 0)


def test_sssp_rmat18(ig, orc, ctx):
    g = ig.rmat(18, 16, seed=1, wmin=1, wmax=255)
    assert np.array_equal(ctx.sssp(ig.DeviceGraph(g), 0), orc.sssp(g, 0))


def test_sssp_few_mid_lists(ig, orc, ctx):
    """A source joined to 20 sub-hubs of 3000 leaves each (lists of 64 to
    4095 entries: mid lists), with weighted cross edges among the leaves:
    the second near iteration holds only 20 mid lists, which SSSP spreads
    over up to 16 waves each.  Distances and BFS levels against the oracle."""
    rng = np.random.default_rng(11)
    hubs, leaves = 20, 3000
    n = 1 + hubs + hubs * leaves
    src = [np.zeros(hubs, np.int64)]
    dst = [np.arange(1, 1 + hubs)]
    for h in range(hubs):
        src.append(np.full(leaves, 1 + h))
        dst.append(1 + hubs + h * leaves + np.arange(leaves))
    cross = rng.integers(1 + hubs, n, size=(2, 20000))
    src.append(cross[0])
    dst.append(cross[1])
    s_, d_ = np.concatenate(src), np.concatenate(dst)
    keep = s_ != d_
    g = ig.from_edges(n, s_[keep].astype(np.int32), d_[keep].astype(np.int32), wseed=5, wmin=1, wmax=1000)
    dg = ig.DeviceGraph(g)
    assert np.array_equal(ctx.sssp(dg, 0), orc.sssp(g, 0))
    assert np.array_equal(ctx.bfs(dg, 0), orc.bfs(g, 0))


def test_cc_catalog(ig, orc, ctx):
    for g in catalog(ig):
        assert np.array_equal(ctx.cc(ig.DeviceGraph(g)), orc.cc(g)), g.name


def test_cc_large(ig, orc, ctx):
    for g in (ig.rmat(18, 16, seed=2), ig.road(256, 256, keep_p=0.5, seed=3)):
        assert np.array_equal(ctx.cc(ig.DeviceGraph(g)), orc.cc(g)), g.name


def test_cc_dense_components(ig, orc, ctx):
    """Average degree >= 8 takes the Afforest order (sampled links, giant
    root, skipped lists).  Here the largest component does not hold vertex 0,
    ids of several components interleave, and isolated vertices sit between
    them, so the skipped tree is neither the first nor the only one."""
    rng = np.random.default_rng(11)
    perm = rng.permutation(6000)
    src, dst = [], []
    for lo, hi, m in ((0, 1000, 9000), (1000, 5000, 40000), (5000, 5900, 7000)):
        a = rng.integers(lo, hi, m)
        b = rng.integers(lo, hi, m)
        src.append(perm[a])
        dst.append(perm[b])
    g = ig.from_edges(6000, np.concatenate(src), np.concatenate(dst))
    assert g.edges >= 8 * g.nodes
    assert np.array_equal(ctx.cc(ig.DeviceGraph(g)), orc.cc(g))


@pytest.mark.parametrize("signed", [False, True])
def test_mis_catalog(ig, orc, ctx, signed):
    for g in catalog(ig):
        assert np.array_equal(ctx.mis(ig.DeviceGraph(g), signed), orc.mis(g, signed)), (g.name, signed)


def test_mis_large(ig, orc, ctx):
    g = ig.rmat(18, 16, seed=4)
    st = ctx.mis(ig.DeviceGraph(g))
    assert np.array_equal(st, orc.mis(g))


def _pr_check(ig, orc, g, scores, iters):
    ref, _ = orc.pr64(g, iters)
    err = np.abs(scores.astype(np.float64) - ref)
    assert err.max() <= PR_MAX_ABS, (g.name, err.max())
    assert err.sum() <= PR_L1, (g.name, err.sum())


def test_pagerank_catalog(ig, orc, ctx):
    for g in catalog(ig):
        scores, iters = ctx.pagerank(ig.DeviceGraph(g))
        assert 1 <= iters <= 100
        _pr_check(ig, orc, g, scores, iters)
        # iteration count agrees with the reference recurrence (FP32) to +-1
        _, it32 = orc.pr_ref32(g)
        assert abs(it32 - iters) <= 1, (g.name, it32, iters)


def test_pagerank_rmat16(ig, orc, ctx):
    g = ig.rmat(16, 16, seed=5)
    scores, iters = ctx.pagerank(ig.DeviceGraph(g))
    _pr_check(ig, orc, g, scores, iters)


def test_tc_catalog(ig, orc, ctx):
    for g in catalog(ig):
        assert ctx.tc(ig.DeviceGraph(g)) == orc.tc(g), g.name
    assert ctx.tc(ig.DeviceGraph(clique(ig, 40))) == 40 * 39 * 38 // 6


def test_tc_clique_unions(ig, ctx):
    """Closed-form counts (the sum of C(k, 3)) on disjoint cliques, on
    consecutive and on shuffled ids.  In a clique the out-list of the j-th
    smallest id holds the k - 1 - j larger ones, so: K3..K17 give every task
    of up to 16 entries (counted 16 tasks at a time); K33..K200 wave-pass
    tasks whose probed lists run from short (the in-wave balancer, 1 to 4
    windows of pairs per step) to 96 entries and more (the whole-wave walk,
    ending on 1 to 4 windows); K300 and K2049 workgroup-pass tasks, the
    latter's largest table at load 1/4 (2048 entries in 8192 slots)."""
    sizes = [3, 4, 5, 9, 16, 17] * 12 + [33, 64, 65, 66, 129, 200, 300, 2049]
    want = sum(k * (k - 1) * (k - 2) // 6 for k in sizes)
    for seed in (None, 4):
        g = cliques(ig, sizes, seed=seed)
        assert ctx.tc(ig.DeviceGraph(g)) == want, seed


def test_tc_rmat16(ig, orc, ctx):
    g = ig.rmat(16, 16, seed=6)
    assert ctx.tc(ig.DeviceGraph(g)) == orc.tc(g)


@pytest.mark.parametrize("lds_list", [0, 600])
def test_tc_long_oriented_lists(ig, orc, ctx, lds_list):
    # oriented lists of 257..2048 entries take the 512-thread workgroup pass,
    # longer ones the 1024-thread pass; with the hint at 600 the lists above
    # 600 entries are hashed 600 entries at a time, every piece probed with
    # all of a part's items (tasks of hubs use a bitmap instead)
    ctx.set_hint(ig._lib.HINT_TC_LDS_LIST, lds_list)
    try:
        assert ctx.tc(ig.DeviceGraph(clique(ig, 1100))) == 1100 * 1099 * 1098 // 6
        g = ig.erdos_renyi(3000, 900000, seed=11)
        assert ctx.tc(ig.DeviceGraph(g)) == orc.tc(g)
        if lds_list == 0:
            assert ctx.tc(ig.DeviceGraph(clique(ig, 2100))) == 2100 * 2099 * 2098 // 6
    finally:
        ctx.set_hint(ig._lib.HINT_TC_LDS_LIST, 0)


def test_tc_rmat17_hubs_and_others(ig, orc, ctx):
    """R-MAT 17 has 131072 vertices: up to 65535 of highest degree are hubs,
    whose out-lists hold 16-bit keys and whose workgroup-pass tasks probe a
    bitmap, and the others keep 32-bit rows and hash tables, so tasks, items
    and probed rows of both kinds mix in every pass."""
    g = ig.rmat(17, 16, seed=7)
    assert ctx.tc(ig.DeviceGraph(g)) == orc.tc(g)


@pytest.mark.parametrize("hubs", [1, 5, 300, 4000])
def test_tc_hub_caps(ig, orc, ctx, hubs):
    """INDIGO_HINT_TC_HUBS moves the line between hubs (16-bit rows, bitmap
    probes) and the rest (32-bit rows, hash probes).  With a cap of 1 the
    clique unions have no hub at all (the largest clique's vertices tie on
    degree, and a tie is never split); caps of 5 to 4000 cut through ties and
    degree classes of R-MAT 14 and of a dense random graph (with 600-entry
    LDS pieces for the hashed tasks too)."""
    sizes = [3, 4, 5, 9, 16, 17] * 4 + [33, 64, 65, 200, 300, 2049]
    want = sum(k * (k - 1) * (k - 2) // 6 for k in sizes)
    ctx.set_hint(ig._lib.HINT_TC_HUBS, hubs)
    try:
        assert ctx.tc(ig.DeviceGraph(cliques(ig, sizes, seed=5))) == want
        g = ig.rmat(14, 16, seed=8)
        assert ctx.tc(ig.DeviceGraph(g)) == orc.tc(g)
        g = ig.erdos_renyi(3000, 900000, seed=12)
        want_er = orc.tc(g)
        assert ctx.tc(ig.DeviceGraph(g)) == want_er
        ctx.set_hint(ig._lib.HINT_TC_LDS_LIST, 600)
        assert ctx.tc(ig.DeviceGraph(g)) == want_er
    finally:
        ctx.set_hint(ig._lib.HINT_TC_HUBS, 0)
        ctx.set_hint(ig._lib.HINT_TC_LDS_LIST, 0)


@pytest.mark.parametrize("extra_b", [0, 7])
def test_tc_two_big_hubs(ig, ctx, extra_b):
    """Two adjacent hubs with 40000+ leaves each, 5000 of them shared: 5000
    triangles (hub, hub, shared leaf).  Degrees at or above 32767 are ranked
    from nindex rather than the clamped 4-byte record; with extra_b 0 the two
    degrees tie (the orientation falls back to ids), with 7 they differ."""
    n_a, n_b, shared = 40000, 40000 + extra_b, 5000
    leaves_a = np.arange(2, 2 + n_a)
    leaves_b = np.arange(2 + n_a - shared, 2 + n_a - shared + n_b)
    src = np.concatenate([[0], np.zeros(n_a, np.int64), np.ones(n_b, np.int64)])
    dst = np.concatenate([[1], leaves_a, leaves_b])
    g = ig.from_edges(int(dst.max()) + 1, src, dst)
    assert ctx.tc(ig.DeviceGraph(g)) == shared


def test_tc_rejects_unsorted(ig, ctx):
    g = ig.grid(8, 8)
    nl = g.nlist.copy()
    nl[g.nindex[9]], nl[g.nindex[9] + 1] = nl[g.nindex[9] + 1], nl[g.nindex[9]]
    bad = ig.HostGraph(g.nodes, g.edges, g.nindex, nl)
    with pytest.raises(ig.IndigoError) as e:
        ctx.tc(ig.DeviceGraph(bad))
    assert e.value.code == ig._lib.ERR_UNSORTED


def test_dropin_entries(ig, orc):
    g = ig.rmat(12, 16, seed=9, wmin=1, wmax=255)
    dg = ig.DeviceGraph(g)
    sp = g.sp()
    d_sp = ig.DeviceBuffer.from_array(sp)
    dist = np.empty(g.nodes, np.int32)
    ig.GPUsssp_vertex(3, dg, dist)
    assert np.array_equal(dist, orc.sssp(g, 3))
    ig.GPUsssp_edge(4, dg, dist, sp)
    assert np.array_equal(dist, orc.sssp(g, 4))
    lab = np.empty(g.nodes, np.int32)
    ig.GPUcc_vertex(dg, lab)
    assert np.array_equal(lab, orc.cc(g))
    ig.GPUcc_edge(dg, lab, sp)
    assert np.array_equal(lab, orc.cc(g))
    pri = np.empty(g.nodes, np.uint32)
    st = np.empty(g.nodes, np.int32)
    ig.GPUmis_vertex(dg, pri, st)
    assert np.array_equal(st, orc.mis(g))
    assert pri[5] == orc.mis_hash(5 + 712313887)
    ig.GPUmis_edge(dg, d_sp, pri, st, signed_priority=True)
    assert np.array_equal(st, orc.mis(g, True))
    scores = np.full(g.nodes, 1.0 / g.nodes, np.float32)
    _, iters = ig.PR_GPU(g, scores, g.degree)
    _pr_check(ig, orc, g, scores, iters)
    cnt, _ = ig.GPUtc_vertex(dg)
    assert cnt == orc.tc(g)
    cnt, _ = ig.GPUtc_edge(dg, d_sp)
    assert cnt == orc.tc(g)


def test_copy_ms_copies_and_times(ig, ctx):
    src = np.arange(1 << 20, dtype=np.int32)
    a = ig.DeviceBuffer.from_array(src)
    b = ig.DeviceBuffer(src.nbytes)
    ms = ctx.copy_ms(b, a, src.nbytes)
    assert ms > 0
    assert np.array_equal(b.to_array(np.int32, src.size), src)


@pytest.mark.parametrize("cap", [1, 7])
def test_persistent_grid_cap(ig, orc, ctx, cap):
    """INDIGO_HINT_MAX_WORKGROUPS (contexts that share the GPU, as the serving
    slots do): the persistent kernels give the same results on a grid of one
    workgroup and on an odd grid smaller than the eight barrier groups, and
    CC's and TC's ordinary kernels on grids held to the same threads."""
    g = ig.rmat(13, 16, seed=5, wmin=1, wmax=255)
    road = ig.road(80, 60, keep_p=0.8, seed=2, wmin=1, wmax=1000)
    ctx.set_hint(ig._lib.HINT_MAX_WORKGROUPS, cap)
    try:
        for h in (g, road):
            dg = ig.DeviceGraph(h)
            assert np.array_equal(ctx.bfs(dg, 0), orc.bfs(h, 0)), (h.name, "bfs")
            assert np.array_equal(ctx.sssp(dg, 0), orc.sssp(h, 0)), (h.name, "sssp")
            assert np.array_equal(ctx.mis(dg), orc.mis(h)), (h.name, "mis")
            scores, iters = ctx.pagerank(dg)
            _pr_check(ig, orc, h, scores, iters)
            assert np.array_equal(ctx.cc(dg), orc.cc(h)), (h.name, "cc")
            assert ctx.tc(dg) == orc.tc(h), (h.name, "tc")
    finally:
        ctx.set_hint(ig._lib.HINT_MAX_WORKGROUPS, 0)
