"""Parity at full size on the batched configuration: the 512-graph mix
(vertex counts log-uniform in [2^10, 2^20], half R-MAT edge factor 8, half
Erdos-Renyi of average degree 8, weights [1, 255]) with a 64-source BFS, a
16-source SSSP, a CC and an MIS query per graph, all 2048 queries in one
persistent launch as the benchmark runs them.

Every row of every query is compared bit for bit with the single-query
kernel run alone on the same graph and source; a seeded sample of rows per
algorithm (eight graphs, one source each, always including the largest
graph) is also compared with the oracle directly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALGS = ("bfs", "sssp", "cc", "mis")


@pytest.fixture(scope="module")
def mix(ig, ctx):
    from indigo2.workloads import batched_mix

    print("generating the 512-graph batched mix", flush=True)
    graphs, srcs = batched_mix()
    assert len(graphs) == 512
    dgs = [ig.DeviceGraph(g) for g in graphs]
    qs = {"bfs": [ig.Query("bfs", i, dgs[i], srcs[i][0]) for i in range(len(graphs))],
          "sssp": [ig.Query("sssp", i, dgs[i], srcs[i][1]) for i in range(len(graphs))],
          "cc": [ig.Query("cc", i, dgs[i]) for i in range(len(graphs))],
          "mis": [ig.Query("mis", i, dgs[i]) for i in range(len(graphs))]}
    print("running 2048 queries in one batched launch", flush=True)
    ctx.batch(dgs, [q for a in ALGS for q in qs[a]])
    yield graphs, dgs, qs
    for a in ALGS:
        for q in qs[a]:
            q.out.close()
    for d in dgs:
        d.close()


def _sample(graphs, k=8, seed=4):
    rng = np.random.default_rng(seed)
    picks = set(int(i) for i in rng.choice(len(graphs), k - 1, replace=False))
    picks.add(int(np.argmax([g.edges for g in graphs])))
    return sorted(picks)


@pytest.mark.parametrize("alg", ALGS)
def test_batched_mix_rows_equal_plain(ig, ctx, mix, alg):
    graphs, dgs, qs = mix
    buf = ig.DeviceBuffer(4 * max(g.nodes for g in graphs))
    rows = 0
    for q in qs[alg]:
        dg = dgs[q.graph]
        got = q.result()
        if alg in ("bfs", "sssp"):
            run = ctx.bfs_into if alg == "bfs" else ctx.sssp_into
            for r, s in enumerate(q.sources):
                run(dg, s, buf)
                assert np.array_equal(got[r], buf.to_array(np.int32, dg.nodes)), (q.graph, alg, s)
                rows += 1
        else:
            (ctx.cc_into if alg == "cc" else ctx.mis_into)(dg, buf)
            assert np.array_equal(got, buf.to_array(np.int32, dg.nodes)), (q.graph, alg)
            rows += 1
    buf.close()
    print(f"{alg}: {rows} rows equal the single-query kernels", flush=True)


@pytest.mark.parametrize("alg", ALGS)
def test_batched_mix_sample_equals_oracle(orc, mix, alg):
    graphs, _, qs = mix
    for i in _sample(graphs):
        g, q = graphs[i], qs[alg][i]
        got = q.result()
        if alg == "bfs":
            assert np.array_equal(got[5], orc.bfs(g, q.sources[5])), (i, alg)
        elif alg == "sssp":
            assert np.array_equal(got[3], orc.sssp(g, q.sources[3])), (i, alg)
        elif alg == "cc":
            assert np.array_equal(got, orc.cc(g)), (i, alg)
        else:
            assert np.array_equal(got, orc.mis(g)), (i, alg)
