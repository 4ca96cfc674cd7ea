"""GPU tests of the serving loop (indigo_serve): requests replayed with
overlapping arrivals on several slots must each give the same result as the
same query run alone (integer outputs bit for bit, PageRank within the FP
tolerance of the single-query kernel), under memory-budget admission."""
import numpy as np
import pytest

from graphs import weighted

pytestmark = pytest.mark.gpu


def _pool(ig):
    gs = [ig.rmat(s, 8, seed=20 + s, wmin=1, wmax=255) for s in (9, 10, 11, 12, 13)]
    gs += [ig.grid(r, c, seed=r, wmin=1, wmax=1000) for r, c in ((16, 16), (32, 20), (64, 64))]
    gs += [ig.erdos_renyi(3000, 12000, seed=4, wmin=1, wmax=255), weighted(ig, ig.road(30, 30, seed=2))]
    # deep, sparse graphs: thousands of levels for the frontier-driven tasks
    gs += [ig.grid(300, 300, seed=7, wmin=1, wmax=255), weighted(ig, ig.road(260, 260, seed=3))]
    return gs


def _requests(rng, gs, n, rate):
    algs = rng.choice(["bfs", "sssp", "cc", "pr", "tc", "mis"], size=n, p=[0.35, 0.25, 0.1, 0.1, 0.1, 0.1])
    gid = rng.integers(0, len(gs), n)
    src = np.array([rng.integers(0, gs[g].nodes) for g in gid])
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    return list(algs), gid, src, arrivals


def _out(ig, alg, g):
    if alg == "tc":
        return ig.DeviceBuffer(8)
    return ig.DeviceBuffer(4 * (g.nodes + (1 if alg == "pr" else 0)))


def _check_alone(ig, orc, ctx, gs, dgs, algs, gid, src, outs):
    from test_gpu_kernels import _pr_check

    for i, alg in enumerate(algs):
        g, dg = gs[gid[i]], dgs[gid[i]]
        if alg == "tc":
            got = int(outs[i].to_array(np.int64, 1)[0])
            assert got == ctx.tc(dg), (i, g.name)
        elif alg == "pr":
            a = outs[i].to_array(np.int32, g.nodes + 1)
            scores, iters = a[: g.nodes].view(np.float32), int(a[g.nodes])
            q = ig.Query("pr", 0, dg)
            ctx.batch([dg], [q])
            alone, it_alone = q.result()
            assert abs(iters - it_alone) <= 1, (i, g.name, iters, it_alone)
            _pr_check(ig, orc, g, scores, iters)
            if iters == it_alone:
                assert np.abs(scores - alone).max() <= 1e-6, (i, g.name)
        else:
            got = outs[i].to_array(np.int32, g.nodes)
            if alg == "bfs":
                want = ctx.bfs(dg, int(src[i]))
            elif alg == "sssp":
                want = ctx.sssp(dg, int(src[i]))
            elif alg == "cc":
                want = ctx.cc(dg)
            else:
                want = ctx.mis(dg)
            assert np.array_equal(got, want), (i, alg, g.name)


@pytest.mark.parametrize("cu_partition", [
    False,
    pytest.param(True, marks=pytest.mark.skip(
        reason="CU-partitioned slots (CU-masked streams) are switched off: a run of this case ended in a host "
               "segmentation fault inside indigo_serve whose cause is not established")),
])
def test_serve_mix_equals_alone(ig, orc, ctx, cu_partition):
    """Every slot may use every CU; with cu_partition the last slot would get
    CUs of its own (CU-masked stream) and serve short requests only."""
    rng = np.random.default_rng(1)
    gs = _pool(ig)
    dgs = [ig.DeviceGraph(g) for g in gs]
    algs, gid, src, arrivals = _requests(rng, gs, 400, rate=20000.0)
    outs = [_out(ig, a, gs[g]) for a, g in zip(algs, gid)]
    print(f"serving 400 requests on 4 slots, cu_partition={cu_partition}", flush=True)
    adm, fin, batch, slot, st = ig.serve(dgs, algs, gid, src, arrivals, outs, slots=4, max_batch=64, progress=5.0,
                                         cu_partition=cu_partition)
    assert st["slots"] == 4
    assert np.all(adm >= arrivals - 1e-9) and np.all(fin >= adm)
    assert batch.max() > 1  # requests were batched together
    assert len(set(slot.tolist())) > 1  # and spread over slots
    # overlap: some request was admitted while another slot's launch was running
    order = np.argsort(adm)
    assert any(adm[order[k + 1]] < fin[order[k]] and slot[order[k + 1]] != slot[order[k]]
               for k in range(len(order) - 1))
    assert st["launches"] + st["tc_launches"] >= 2
    _check_alone(ig, orc, ctx, gs, dgs, algs, gid, src, outs)


def test_serve_under_tight_memory_budget(ig, orc, ctx):
    rng = np.random.default_rng(2)
    gs = _pool(ig)
    dgs = [ig.DeviceGraph(g) for g in gs]
    algs, gid, src, arrivals = _requests(rng, gs, 120, rate=50000.0)
    outs = [_out(ig, a, gs[g]) for a, g in zip(algs, gid)]
    budget = 3 * 4 * (max(g.nodes for g in gs) + 1) + 4096
    print(f"serving 120 requests with a {budget}-byte budget", flush=True)
    adm, fin, batch, slot, st = ig.serve(dgs, algs, gid, src, arrivals, outs, slots=3, max_batch=16,
                                         mem_budget=budget)
    assert st["peak_bytes"] <= budget
    assert st["max_waiting"] > 3  # admission held requests back
    _check_alone(ig, orc, ctx, gs, dgs, algs, gid, src, outs)


def test_serve_refuses_the_cu_partition(ig, ctx):
    g = ig.rmat(10, 8, seed=1, wmin=1, wmax=255)
    dg = ig.DeviceGraph(g)
    with pytest.raises(ig.IndigoError) as e:
        ig.serve([dg], ["bfs"], [0], [0], [0.0], None, cu_partition=True)
    assert e.value.code == ig._lib.ERR_INVALID_ARG


def test_serve_rejects_what_cannot_fit(ig, ctx):
    g = ig.rmat(12, 8, seed=1, wmin=1, wmax=255)
    dg = ig.DeviceGraph(g)
    with pytest.raises(ig.IndigoError) as e:
        ig.serve([dg], ["bfs"], [0], [0], [0.0], None, slots=1, max_batch=4, mem_budget=1024)
    assert e.value.code == ig._lib.ERR_CAPACITY
    with pytest.raises(ig.IndigoError):
        ig.serve([dg], ["bfs"], [0], [g.nodes], [0.0], None)
    gu = ig.DeviceGraph(ig.grid(8, 8))
    with pytest.raises(ig.IndigoError):
        ig.serve([gu], ["sssp"], [0], [0], [0.0], None)
