#!/usr/bin/env python3
"""Benchmark: the six Indigo2 graph kernels on one MI355X.

Workload (one "step"): BFS, SSSP, CC, MIS, PageRank and TC, once each, on one
synthetic R-MAT graph - scale 22 (4,194,304 vertices), edge factor 16,
a/b/c/d = 0.57/0.19/0.19/0.05, seed 1, symmetrised, duplicates and self loops
removed, int32 CSR, SSSP weights uniform in [1, 255]; BFS/SSSP from vertex 0.
The graph is generated on the host and uploaded before timing; every timed
call reads it from HBM and writes its result to HBM.

GTEPS follows the reference convention: 1e-9 * (directed CSR entries) /
seconds, per kernel run.  `value` is the whole step: 6 * E / step time.

Prints ONE JSON line on stdout (progress goes to stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "parallelization-artifact-sc23_amd"))

import numpy as np  # noqa: E402

ALGS = ("bfs", "sssp", "cc", "mis", "pr", "tc")
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def tc_count_model(g) -> dict:
    """Work of tc_count_group<512, 8192>, the kernel that does most of TC.

    Restates the kernel's task split in numpy: oriented out-lists N+(v) by
    (degree, id); each oriented edge v -> u is taken by the endpoint with the
    longer out-list, iterating the shorter one (ties: v); the task of x is
    its out-list plus its in-list.  Tasks with 2 <= |N+(x)| <= 256 and
    (|N+| + |N-|) * |N+| <= W run in the wave pass (W = 2^13 below 2^24 CSR
    entries, else 2^14, as tc.hip's wave_work_for), the rest with
    |N+(x)| <= 2048 in this kernel, in parts of 2048 items.  Hubs are the at
    most 65535 vertices of largest degree (all vertices when they fit, else
    min(degree, 65536) >= the smallest bin whose suffix holds at most 65535
    vertices); a hub's out-list holds hubs only and is stored as 16-bit keys.
    Algorithmic bytes per launch: per probe the element of the iterated list
    (2 from a hub's list, else 4), 12 per item (its id and offset pair), per
    table entry (N+(x) per part) 2 for a hub's task, else 4, and 16 per part
    (task header)."""
    v = g.nodes
    deg = g.degree.astype(np.int64)
    if v <= 65535:
        hub = np.ones(v, dtype=bool)
    else:
        dcl = np.minimum(deg, 65536)
        suffix = np.cumsum(np.bincount(dcl, minlength=65537)[::-1])[::-1]
        ok = np.flatnonzero(suffix <= 65535)
        hub = dcl >= (ok[0] if ok.size else 65537)
    src = np.repeat(np.arange(v, dtype=np.int64), deg)
    dst = g.nlist.astype(np.int64)
    rank = deg * v + np.arange(v, dtype=np.int64)
    keep = rank[dst] > rank[src]
    s, d = src[keep], dst[keep]
    del src, dst, keep, rank
    oc = np.bincount(s, minlength=v).astype(np.int64)
    to_d = oc[s] < oc[d]
    inl = np.bincount(d[to_d], minlength=v).astype(np.int64)
    owner = np.where(to_d, d, s)
    walked = np.where(to_d, s, d)  # the endpoint whose out-list is iterated
    probes = np.bincount(owner, weights=oc[walked], minlength=v)
    pbytes = np.bincount(owner, weights=oc[walked] * np.where(hub[walked], 2.0, 4.0), minlength=v)
    items = oc + inl
    wave = (oc >= 2) & (oc <= 256) & (items * oc <= (1 << 13 if g.edges < (1 << 24) else 1 << 14))
    mid = (oc >= 2) & ~wave & (oc <= 2048)
    parts = (items + 2047) // 2048
    m = {"probes": float(probes[mid].sum()), "items": int(items[mid].sum()), "parts": int(parts[mid].sum()),
         "table_entries": int((oc * parts)[mid].sum()), "probes_all_passes": float(probes.sum()),
         "hub_tasks": int((mid & hub).sum()), "tasks": int(mid.sum())}
    tbytes = float((oc * parts * np.where(hub, 2, 4))[mid].sum())
    m["bytes"] = float(pbytes[mid].sum()) + 12 * m["items"] + tbytes + 16 * m["parts"]
    return m


def tc_merge_bytes(g) -> float:
    """Algorithmic bytes of a whole TC call under the merge model of the
    reference's vertex form (id order, include/tc_vertex_cuda.h:44-63):
    4E + 4V + 4 * sum_v sum_{u in N+(v)} (|N+(v) cap (u, inf)| + |N+(u)|),
    N+(x) = neighbours with a larger id."""
    deg = g.degree.astype(np.int64)
    src = np.repeat(np.arange(g.nodes, dtype=np.int64), deg)
    up = g.nlist.astype(np.int64) > src
    dplus = np.bincount(src[up], minlength=g.nodes).astype(np.int64)
    pairs = (dplus * (dplus - 1) // 2).sum()
    fan = dplus[g.nlist[up]].sum()
    return 4.0 * g.edges + 4.0 * g.nodes + 4.0 * float(pairs + fan)


def copy_peak_gbs(ig, ctx) -> float:
    """Device-to-device copy bandwidth (read + write bytes / time) of 1 GiB,
    best of 5, timed with HIP events on the context stream."""
    n = 1 << 30
    a, b = ig.DeviceBuffer(n), ig.DeviceBuffer(n)
    ctx.copy_ms(b, a, n)
    best = max(2.0 * n / (ctx.copy_ms(b, a, n) * 1e-3) / 1e9 for _ in range(5))
    del a, b
    return best


def measure_host_transfer(ig, ctx, g, dg) -> dict:
    """The PCIe side of the drop-in form, outside `value`: uploading the
    graph (nindex + nlist + eweight) through the library's pinned staging
    (indigo_graph_upload) and through plain pageable copies, building the
    edge-source array on the device, and copying one int32[V] result back."""
    nbytes = 4 * (g.nodes + 1) + 8 * g.edges
    staged = []
    for _ in range(2):
        d = ig.DeviceGraph.upload(ctx, g)
        staged.append(d.upload_seconds)
        d.close()
    t = time.perf_counter()
    d = ig.DeviceGraph(g)
    pageable = time.perf_counter() - t
    d.close()
    sp = dg.build_sp(ctx)
    sp_ms = ctx.last_ms
    sp.close()
    out = ig.DeviceBuffer(4 * g.nodes)
    t = time.perf_counter()
    out.to_array(np.int32, g.nodes)
    d2h = time.perf_counter() - t
    res = {"graph_bytes": nbytes, "staged_upload_gbs": round(nbytes / min(staged) / 1e9, 2),
           "pageable_upload_gbs": round(nbytes / pageable / 1e9, 2), "sp_build_ms": round(sp_ms, 3),
           "result_d2h_ms": round(1e3 * d2h, 3)}
    log(f"host transfer: {res}")
    return res


def kernel_bytes(alg: str, g, pr_iters: int, tc_model: dict) -> float:
    """Algorithmic HBM bytes of each algorithm's main kernel per launch: every
    CSR entry read once (4 B) plus the per-entry lookup it drives (4 B: the
    neighbour's distance / parent / status / contribution; SSSP also the
    weight), and the per-vertex index pair and result (8 B; PR reads the
    index pair, scores and writes scores and contributions, 16 B, per sweep)."""
    e, v = float(g.edges), float(g.nodes)
    if alg in ("bfs", "cc", "mis"):
        return 8 * e + 8 * v
    if alg == "sssp":
        return 12 * e + 8 * v
    if alg == "pr":
        return pr_iters * (8 * e + 16 * v)
    return tc_model["bytes"]


# one main kernel per algorithm for the traffic record; CC's timed span is its
# whole kernel sequence, recorded as the sum over its kernels per call
MAIN_KERNELS = {"bfs": "bfs_kernel", "sssp": "sssp_kernel", "mis": "mis_kernel", "pr": "pr_kernel",
                "tc": "tc_count_group<512, 8192>", "cc": "cc_init..cc_flatten"}
CC_SEQUENCE = ("cc_init", "cc_sample", "cc_compress", "cc_giant", "cc_link", "cc_link_heavy", "cc_flatten")


def read_traffic(path: str, workload: str, kernel: str) -> dict | None:
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE) on this workload, scaled by the same counters on
    a copy of known size; recorded by the profiling commands in the file."""
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("workload") != workload or kernel not in rec.get("kernels", {}):
        return None
    return dict(rec["kernels"][kernel], source=rec.get("source"))


def _pmc_rows(path: str, counter: str, match) -> list[tuple[str, float]]:
    import csv

    with open(path) as f:
        return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in csv.DictReader(f)
                if r["Counter_Name"] == counter and match(r["Kernel_Name"])]


def _pmc_values(path: str, counter: str, match) -> list[float]:
    return [v for _, v in _pmc_rows(path, counter, match)]


def _per_call_sum(rows: list[tuple[str, float]]) -> tuple[float, int]:
    """Sum over distinct kernels of each kernel's mean value per dispatch
    (one call of a kernel sequence), and the dispatches seen."""
    by = {}
    for name, v in rows:
        by.setdefault(name, []).append(v)
    return sum(statistics.mean(v) for v in by.values()), len(rows)


def record_traffic(paths: list[str], workload: str, out: str) -> int:
    """Writes the traffic record of every algorithm's main kernel from four
    rocprofv3 --pmc CSVs: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) on the
    workload, then of the 1 GiB copy of --copy-calibration.  Each counter is
    scaled by what it reads for the copy, which moves exactly 1 GiB each
    way."""
    kf, kw, cf, cw = paths
    is_copy = lambda name: "copy" in name.lower() or "elementwise" in name.lower()
    cfetch = [x for x in _pmc_values(cf, "FETCH_SIZE", is_copy) if x > 1e5]
    cwrite = [x for x in _pmc_values(cw, "WRITE_SIZE", is_copy) if x > 1e5]
    if not (cfetch and cwrite):
        print("record-traffic: no copy dispatches in the calibration CSVs", file=sys.stderr)
        return 1
    gib_kib = float(1 << 20)
    ffac, wfac = gib_kib / statistics.median(cfetch), gib_kib / statistics.median(cwrite)
    kernels = {}
    for alg, kernel in MAIN_KERNELS.items():
        if alg == "cc":
            match = lambda name: any(f"::{k}(" in name or f"::{k}<" in name for k in CC_SEQUENCE)
        else:  # plain or templated symbol (the kernel may be named without its template arguments)
            match = lambda name, k=kernel: f"::{k}(" in name or f"::{k}<" in name
        fetch, nf = _per_call_sum(_pmc_rows(kf, "FETCH_SIZE", match))
        write, nw = _per_call_sum(_pmc_rows(kw, "WRITE_SIZE", match))
        if not (nf and nw):
            continue
        rb = 1024.0 * fetch * ffac
        wb = 1024.0 * write * wfac
        kernels[kernel] = {"algorithm": alg, "bytes_per_launch": rb + wb, "read_bytes": rb, "write_bytes": wb,
                           "fetch_kib_per_dispatch": fetch, "write_kib_per_dispatch": write, "dispatches": [nf, nw]}
    if not kernels:
        print("record-traffic: no main-kernel dispatches in the CSVs", file=sys.stderr)
        return 1
    rec = {"workload": workload, "copy_scale": {"fetch": ffac, "write": wfac}, "kernels": kernels,
           "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, python3 bench.py --steps 2 --warmup 0 "
                     "--no-cpu-baseline --profile-only, scaled by a 1 GiB copy (bench.py --copy-calibration)"}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec))
    return 0


def run_batched(args) -> int:
    line = measure_batched()
    print(json.dumps(line), flush=True)
    return 0 if line["parity_spot_mismatches"] == 0 else 1


def measure_batched() -> dict:
    """Batched vs plain on the batched configuration: per algorithm, one
    indigo_batch launch over all graphs against the plain resident kernels
    called once per (graph, source).  GTEPS = 1e-9 * sum over queries of
    sources x CSR entries / wall seconds (device-resident inputs and
    outputs)."""
    import indigo2 as ig

    from indigo2.autotune import DEFAULT_DB, Tuning, classify
    from indigo2.workloads import batched_mix

    t0 = time.time()
    graphs, srcs = batched_mix()
    tv = sum(g.nodes for g in graphs)
    te = sum(g.edges for g in graphs)
    log(f"batched workload: {len(graphs)} graphs, V={tv} E={te} ({time.time() - t0:.1f}s)")
    dgs = [ig.DeviceGraph(g) for g in graphs]
    ctx = ig.Context(0)
    # gang sizes per graph class from the autotuner's database, when there is one
    tuning, classes = None, None
    if os.path.exists(DEFAULT_DB):
        tuning = Tuning.load(DEFAULT_DB)
        classes = classify(ctx, graphs, dgs)

    def gang(alg, i):
        return tuning.gang(alg, classes[i], graphs[i].nodes, graphs[i].edges) if tuning else 0

    out = {}
    plan = {"bfs": lambda i: srcs[i][0], "sssp": lambda i: srcs[i][1], "cc": lambda i: [], "mis": lambda i: []}
    queries_all = {}
    for alg, sel in plan.items():
        qs = [ig.Query(alg, i, dgs[i], sel(i), gang=gang(alg, i)) for i in range(len(graphs))]
        queries_all[alg] = qs
        teps = sum(max(1, len(q.sources)) * graphs[q.graph].edges for q in qs)
        ctx.batch(dgs, qs)  # warm-up
        t = time.perf_counter()
        ctx.batch(dgs, qs)
        wall_b = time.perf_counter() - t
        dev_b = ctx.last_ms * 1e-3
        # plain: one resident call per (graph, source)
        buf = ig.DeviceBuffer(4 * max(g.nodes for g in graphs))
        calls = 0
        t = time.perf_counter()
        dev_p = 0.0
        for q in qs:
            dg = dgs[q.graph]
            if alg == "bfs":
                for s_ in q.sources:
                    ctx.bfs_into(dg, int(s_), buf)
                    dev_p += ctx.last_ms
                    calls += 1
            elif alg == "sssp":
                for s_ in q.sources:
                    ctx.sssp_into(dg, int(s_), buf)
                    dev_p += ctx.last_ms
                    calls += 1
            elif alg == "cc":
                ctx.cc_into(dg, buf)
                dev_p += ctx.last_ms
                calls += 1
            else:
                ctx.mis_into(dg, buf)
                dev_p += ctx.last_ms
                calls += 1
        wall_p = time.perf_counter() - t
        out[alg] = {"queries": len(qs), "plain_calls": calls, "edges_traversed": teps,
                    "batched_gteps": round(1e-9 * teps / wall_b, 3), "plain_gteps": round(1e-9 * teps / wall_p, 3),
                    "speedup": round(wall_p / wall_b, 2), "batched_s": round(wall_b, 4), "plain_s": round(wall_p, 3),
                    "batched_kernel_s": round(dev_b, 4), "plain_kernel_s": round(dev_p * 1e-3, 3)}
        log(f"{alg}: batched {wall_b * 1e3:.1f} ms, plain {wall_p:.2f} s ({calls} calls)")
    # spot parity: a few rows of the batched results against the plain kernels
    rng = np.random.default_rng(5)
    mismatches = 0
    for alg in ("bfs", "sssp", "cc", "mis"):
        for i in rng.choice(len(graphs), 6, replace=False):
            q = queries_all[alg][i]
            got = q.result()
            dg = dgs[i]
            if alg == "bfs":
                want = ctx.bfs(dg, int(q.sources[7]))
                mismatches += int(not np.array_equal(got[7], want))
            elif alg == "sssp":
                want = ctx.sssp(dg, int(q.sources[3]))
                mismatches += int(not np.array_equal(got[3], want))
            elif alg == "cc":
                mismatches += int(not np.array_equal(got, ctx.cc(dg)))
            else:
                mismatches += int(not np.array_equal(got, ctx.mis(dg)))
    # everything in one launch
    qall = [q for alg in ("bfs", "sssp", "cc", "mis") for q in queries_all[alg]]
    teps_all = sum(out[a]["edges_traversed"] for a in out)
    t = time.perf_counter()
    ctx.batch(dgs, qall)
    wall_all = time.perf_counter() - t
    plain_all = sum(out[a]["plain_s"] for a in out)
    line = {
        "metric": "GTEPS per algorithm, batched vs plain on 1 MI355X",
        "value": round(1e-9 * teps_all / wall_all, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"batched: {len(graphs)} graphs, V log-uniform in [2^10, 2^20], half R-MAT ef8 / half "
                               f"Erdos-Renyi avg degree 8, weights [1,255]; 64 BFS + 16 SSSP sources, CC and MIS "
                               f"per graph", "vertices": tv, "csr_entries": te},
        "all_in_one_launch_s": round(wall_all, 4),
        "plain_total_s": round(plain_all, 3),
        "speedup_vs_plain": round(plain_all / wall_all, 2),
        "per_algorithm": out,
        "parity_spot_mismatches": mismatches,
        "autotuned": ({"db": os.path.relpath(DEFAULT_DB, REPO),
                       "classes": {c: classes.count(c) for c in sorted(set(classes))}} if tuning else None),
    }
    ctx.close()
    return line


def measure_serving(rates=(1000.0,), nq: int = 10000, seed: int = 5, slots: int = 4, max_batch: int = 256,
                    mem_budget: int = 8 << 30, batch_work_us: int = 0) -> dict:
    """10,000 seeded requests with Poisson arrivals at each rate over the pool;
    mix BFS 40% / SSSP 30% / CC 10% / PageRank 10% / TC 10%, uniform graph
    and source.  Latency = finish - arrival; aggregate GTEPS = 1e-9 * sum of
    the requests' CSR entries / makespan."""
    import indigo2 as ig

    from indigo2.workloads import SERVE_LONG_US, serving_cost_us, serving_pool, serving_requests

    t0 = time.time()
    gs, names = serving_pool()
    log(f"serving pool: 64 graphs, V={sum(g.nodes for g in gs)} E={sum(g.edges for g in gs)} "
        f"({time.time() - t0:.1f}s)")
    dgs = [ig.DeviceGraph(g) for g in gs]
    algs, gid, src, gaps = serving_requests(gs, nq, seed)
    edges = np.array([gs[g].edges for g in gid], dtype=np.float64)
    # warm-up: a short run touches every code path and sizes the workspaces
    log("serving warm-up: 200 requests at once")
    ig.serve(dgs, algs[:200], gid[:200], src[:200], np.zeros(200), None, slots=slots, max_batch=max_batch,
             mem_budget=mem_budget, progress=10.0)
    runs = []
    for rate in rates:
        arrivals = np.cumsum(gaps / rate)
        log(f"serving {nq} requests at {rate:.0f} q/s")
        adm, fin, batch, slot, st = ig.serve(dgs, algs, gid, src, arrivals, None, slots=slots,
                                             max_batch=max_batch, mem_budget=mem_budget, progress=10.0,
                                             batch_work_us=batch_work_us)
        lat = fin - arrivals
        per = {}
        for a in ("bfs", "sssp", "cc", "pr", "tc"):
            m = np.array([x == a for x in algs])
            per[a] = {"n": int(m.sum()), "p50_ms": round(1e3 * float(np.percentile(lat[m], 50)), 3),
                      "p99_ms": round(1e3 * float(np.percentile(lat[m], 99)), 3)}
        # by the loop's own cost classes (short requests must not queue behind long ones)
        long_m = np.array([serving_cost_us(gs[g], a) > SERVE_LONG_US for a, g in zip(algs, gid)])
        classes = {}
        wait, service = adm - arrivals, fin - adm
        for name, m in (("short", ~long_m), ("long", long_m)):
            if m.any():
                classes[name] = {"n": int(m.sum()), "p50_ms": round(1e3 * float(np.percentile(lat[m], 50)), 3),
                                 "p99_ms": round(1e3 * float(np.percentile(lat[m], 99)), 3),
                                 "wait_p50_ms": round(1e3 * float(np.percentile(wait[m], 50)), 3),
                                 "wait_p99_ms": round(1e3 * float(np.percentile(wait[m], 99)), 3),
                                 "service_p50_ms": round(1e3 * float(np.percentile(service[m], 50)), 3),
                                 "service_p99_ms": round(1e3 * float(np.percentile(service[m], 99)), 3)}
        runs.append({"rate_qps": rate, "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
                     "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
                     "mean_batch": round(float(np.mean(batch)), 2), "makespan_s": round(st["makespan"], 3),
                     "arrival_span_s": round(float(arrivals[-1]), 3),
                     "aggregate_gteps": round(1e-9 * edges.sum() / st["makespan"], 3),
                     "launches": st["launches"], "tc_launches": st["tc_launches"],
                     "peak_output_bytes": st["peak_bytes"], "max_waiting": st["max_waiting"],
                     "slots_used": st["slots"], "per_class": classes, "per_algorithm": per})
        log(f"serving at {rate:.0f} q/s: p50 {runs[-1]['p50_ms']} ms, p99 {runs[-1]['p99_ms']} ms, "
            f"{runs[-1]['aggregate_gteps']} GTEPS, mean batch {runs[-1]['mean_batch']}")
    return {"workload": "serving: 64 resident graphs (32 R-MAT ef16 scales 14-22, 32 grids 256^2-2048^2), "
                        f"{nq} requests, Poisson arrivals, mix BFS .4 / SSSP .3 / CC .1 / PR .1 / TC .1",
            "slots": slots, "max_batch": max_batch, "mem_budget": mem_budget,
            "batch_work_us": batch_work_us or "no cap", "runs": runs}


def measure_deep(kind: str, steps: int, warmup: int) -> dict:
    """High-diameter single-graph workloads, each algorithm from vertex 0.
    grid: BFS on the 1024 x 1024 4-neighbour grid (row-major ids, sorted
    lists); the result is checked against the closed form r + c.
    road: BFS, SSSP and CC on the 4096 x 4096 road-like grid (each
    4-neighbour edge kept with probability 0.85, one diagonal edge for 2% of
    the vertices, weights uniform [1, 1000], seed 1); reports the size of
    vertex 0's component.  Times are medians over `steps` calls with device-
    resident inputs and outputs; the roofline uses the same per-entry bytes
    as the R-MAT line."""
    import numpy as np

    import indigo2 as ig

    t0 = time.time()
    if kind == "grid":
        g = ig.grid(1024, 1024, seed=1, wmin=1, wmax=255)
        algs = ("bfs",)
        desc = "BFS from vertex 0 on the 1024x1024 4-neighbour grid"
    else:
        g = ig.road(4096, 4096, keep_p=0.85, diag_p=0.02, seed=1, wmin=1, wmax=1000)
        algs = ("bfs", "sssp", "cc")
        desc = ("BFS, SSSP and CC from vertex 0 on the 4096x4096 road-like grid (keep .85, diagonals .02, "
                "weights [1,1000], seed 1)")
    log(f"{kind}: V={g.nodes} E={g.edges} ({time.time() - t0:.1f}s)")
    dg = ig.DeviceGraph(g)
    out = ig.DeviceBuffer(4 * g.nodes)
    extra = {}
    with ig.Context(0) as ctx:
        calls = {"bfs": lambda: ctx.bfs_into(dg, 0, out), "sssp": lambda: ctx.sssp_into(dg, 0, out),
                 "cc": lambda: ctx.cc_into(dg, out)}
        res = {}
        for a in algs:
            for _ in range(warmup):
                calls[a]()
            ms, kms = [], []
            for _ in range(steps):
                calls[a]()
                ms.append(ctx.last_ms)
                name, k = ctx.last_kernel
                kms.append(k)
            b = kernel_bytes(a, g, 0, {})
            km = statistics.median(kms)
            res[a] = {"call_ms": round(statistics.median(ms), 4), "gteps": round(1e-9 * g.edges / (statistics.median(ms) * 1e-3), 4),
                      "kernel": name, "kernel_ms": round(km, 4), "algorithmic_bytes": b,
                      "frac": round(b / (km * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            log(f"{kind} {a}: {res[a]['call_ms']:.3f} ms, {res[a]['gteps']:.3f} GTEPS")
        if kind == "grid":
            dist = ctx.bfs(dg, 0)
            r, c = np.divmod(np.arange(g.nodes, dtype=np.int64), 1024)
            extra["closed_form_ok"] = bool(np.array_equal(dist, (r + c).astype(np.int32)))
            extra["levels"] = int(dist.max()) + 1
        else:
            labels = ctx.cc(dg)
            extra["component_of_0"] = int(np.count_nonzero(labels == labels[0]))
            dist = ctx.bfs(dg, 0)
            extra["bfs_levels"] = int(dist[dist != np.iinfo(np.int32).max].max()) + 1
    dg.close()
    return {"workload": desc, "vertices": g.nodes, "csr_entries": g.edges, "steps": steps, "per_algorithm": res,
            **extra}


def cpu_threads() -> int:
    """Host threads for the CPU baseline: OMP_NUM_THREADS when it is set (the
    GPU box sets it to the one GPU's share of the host), else every core."""
    n = os.cpu_count() or 1
    try:
        return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", "") or n)))
    except ValueError:
        return n


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scale: int, threads: int, runs: int = 3) -> dict | None:
    """The reference's own CPU variants (compiled from its sources into
    oracle/_ref) on the benchmark's own graph: R-MAT `scale`, edge factor 16,
    seed 1, weights [1, 255], source 0.  Per algorithm a C++-threads and an
    OpenMP variant - the winners of the all-variant sweep (oracle/sweep_cpu.py)
    on the GPU box's host, as scripts/run_codes.py picks the best variant per
    machine - are each timed `runs` times, and the faster median counts: the
    reference harness's own rule, median of 3 (include/bfs_vertex_cpp.h:106-133).
    TC is bounded: one run of the OpenMP variant only (the C++-threads
    variants are several times slower on this graph)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    try:
        import oracle as orc  # checker / baseline only
    except ImportError:
        return None
    if not all(orc.ref_available(a) for a in ALGS) or not orc.ref_available("tc_omp"):
        return None
    import indigo2 as ig

    g = ig.rmat(scale, 16, seed=1, wmin=1, wmax=255)
    log(f"cpu baseline: rmat{scale} V={g.nodes} E={g.edges}, {threads} threads on {cpu_model()}")
    cands = {
        "bfs": {"C++ threads": lambda: orc.ref_bfs(g, 0, threads)[-1], "OpenMP": lambda: orc.ref_bfs_omp(g, 0, threads)[-1]},
        "sssp": {"C++ threads": lambda: orc.ref_sssp(g, 0, threads)[-1],
                 "OpenMP": lambda: orc.ref_sssp_omp(g, 0, threads)[-1]},
        "cc": {"C++ threads": lambda: orc.ref_cc(g, threads)[-1], "OpenMP": lambda: orc.ref_cc_omp(g, threads)[-1]},
        "mis": {"C++ threads": lambda: orc.ref_mis(g, threads)[-1], "OpenMP": lambda: orc.ref_mis_omp(g, threads)[-1]},
        "pr": {"C++ threads": lambda: orc.ref_pr(g, threads)[-1], "OpenMP": lambda: orc.ref_pr_omp(g, threads)[-1]},
        "tc": {"OpenMP": lambda: orc.ref_tc_omp(g, threads)[-1]},
    }
    times, variants, all_runs = {}, {}, {}
    for a in ALGS:
        for kind, run in cands[a].items():
            if kind == "OpenMP" and a != "tc" and not orc.ref_available(f"{a}_omp"):
                continue
            ts = [run() for _ in range(1 if a == "tc" else runs)]
            all_runs[f"{a} {kind}"] = [round(t, 4) for t in ts]
            med = statistics.median(ts)
            if med > 0 and (a not in times or med < times[a]):
                times[a], variants[a] = med, kind
        log(f"cpu baseline {a}: {times[a]:.3f} s ({variants[a]})")
    per = {a: 1e-9 * g.edges / max(t, 1e-9) for a, t in times.items()}
    total = sum(times.values())
    return {
        "value": round(6 * 1e-9 * g.edges / total, 6),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "reference",
        "sample": f"the benchmark's own graph, rmat{scale} ef16 seed 1 (V={g.nodes}, E={g.edges}), source 0; per "
                  f"algorithm the faster median of {runs} runs of the C++-threads and the OpenMP variant that won "
                  f"an all-variant sweep on this host; TC bounded to one run of its OpenMP variant; {threads} threads of "
                  f"{os.cpu_count()} on {cpu_model()}",
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "variants": variants,
        "per_algorithm_gteps": {a: round(x, 6) for a, x in per.items()},
        "seconds": {a: round(t, 4) for a, t in times.items()},
        "runs": all_runs,
    }


def measure_reordered(ig, ctx, g, dg, run_hints, steps: int, warmup: int) -> dict:
    """The same six-algorithm step with BFS, SSSP and PageRank on a
    degree-ordered copy of the graph (ig.ReorderedGraph: relabelled once on
    the host, before the timed steps, like the upload), each call followed
    by indigo_unpermute so its result is in the original ids; CC, MIS and TC
    stay on the original graph (their results depend on the ids, and TC's
    lists are no faster relabelled).  Device ms per algorithm include the
    read-back; the step's wall time is timed as the main step's is."""
    t = time.perf_counter()
    rg = ig.ReorderedGraph(g)
    prep = time.perf_counter() - t
    out_i = ig.DeviceBuffer(4 * g.nodes)
    out_f = ig.DeviceBuffer(4 * g.nodes)

    def run(a):
        run_hints(a)
        if a == "bfs":
            return ctx.bfs_reordered_into(rg, 0, out_i)
        if a == "sssp":
            return ctx.sssp_reordered_into(rg, 0, out_i)
        if a == "pr":
            return ctx.pagerank_reordered_into(rg, out_f)[1]
        if a == "cc":
            ctx.cc_into(dg, out_i)
        elif a == "mis":
            ctx.mis_into(dg, out_i)
        else:
            ctx.tc(dg)
        return ctx.last_ms

    for _ in range(warmup):
        for a in ALGS:
            run(a)
    ms = {a: [] for a in ALGS}
    t = time.perf_counter()
    for _ in range(steps):
        for a in ALGS:
            ms[a].append(run(a))
    wall = time.perf_counter() - t
    out_i.close()
    out_f.close()
    rg.close()
    step = 1e3 * wall / steps
    med = {a: statistics.median(ms[a]) for a in ALGS}
    log(f"reordered step {step:.3f} ms: " + " ".join(f"{a}={med[a]:.3f}ms" for a in ALGS))
    return {
        "what": "BFS, SSSP and PageRank on a degree-ordered copy of the same graph (new ids by degree, largest "
                "first), results read back in the original ids inside each call; CC, MIS, TC on the original",
        "reorder_seconds_host": round(prep, 3),
        "ms_per_step": round(step, 3),
        "value": round(6 * 1e-9 * g.edges / (step * 1e-3), 4),
        "per_algorithm_ms": {a: round(med[a], 4) for a in ALGS},
    }


def cpu_baseline_isolated(scale: int) -> dict | None:
    """cpu_baseline in a child process: the reference's CPU code runs there,
    so a crash in it (some of its variants fault on some graphs) costs the
    baseline, not the benchmark.  The child touches no GPU."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--cpu-scale", str(scale)],
                       stdout=subprocess.PIPE, text=True, timeout=900)
    lines = r.stdout.strip().splitlines()
    if r.returncode != 0 or not lines:
        log(f"cpu baseline skipped: child exited with status {r.returncode}")
        return None
    return json.loads(lines[-1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--cpu-scale", type=int, default=None,
                    help="R-MAT scale of the CPU baseline (default: --scale, the benchmark's own graph)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain sub-measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sub-measurement")
    ap.add_argument("--no-deep", action="store_true",
                    help="skip the 1024^2 grid and 4096^2 road-like sub-measurements")
    ap.add_argument("--algs", default=",".join(ALGS), help="subset to run (profiling); the JSON line needs all six")
    ap.add_argument("--rates", default="500,1000,2000", help="serving workload: arrival rates in requests/s")
    ap.add_argument("--slots", type=int, default=4, help="serving workload: concurrent slots")
    ap.add_argument("--max-batch", type=int, default=256, help="serving workload: requests per launch")
    ap.add_argument("--batch-work-us", type=int, default=0,
                    help="serving workload: estimated whole-GPU microseconds per launch (0: no cap)")
    ap.add_argument("--workload", choices=["rmat", "batched", "serving", "grid", "road"], default="rmat",
                    help="rmat: the six algorithms on the R-MAT graph (the default line); batched: batched vs "
                         "plain; serving: request loop; grid: BFS on the 1024^2 grid; road: BFS/SSSP/CC on the "
                         "4096^2 road-like grid")
    ap.add_argument("--profile-only", action="store_true",
                    help="run the timed steps and stop (no TC model, no JSON): the command profiled by rocprofv3")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 5 times (the PMC scale reference) and exit")
    ap.add_argument("--record-traffic", nargs=4, metavar=("KFETCH", "KWRITE", "CFETCH", "CWRITE"),
                    help="write profiles/traffic_<workload>.json for the main kernels from four PMC CSVs")
    ap.add_argument("--cpu-baseline-only", action="store_true",
                    help="print the CPU baseline record alone (the child process of a full run)")
    args = ap.parse_args()
    if args.cpu_baseline_only:
        print(json.dumps(cpu_baseline(args.cpu_scale or args.scale, cpu_threads())), flush=True)
        return 0
    if args.record_traffic:
        workload = f"rmat{args.scale} ef{args.edge_factor}"
        return record_traffic(args.record_traffic, workload,
                              os.path.join(REPO, "profiles", f"traffic_{workload.replace(' ', '_')}.json"))
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every kernel here runs on a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "batched":
        return run_batched(args)
    if args.workload in ("grid", "road"):
        print(json.dumps(measure_deep(args.workload, args.steps, args.warmup)), flush=True)
        return 0
    if args.workload == "serving":
        print(json.dumps(measure_serving(tuple(float(r) for r in args.rates.split(",")), slots=args.slots,
                                         max_batch=args.max_batch, batch_work_us=args.batch_work_us)), flush=True)
        return 0
    if args.copy_calibration:
        import indigo2 as ig

        ctx = ig.Context(0)
        n = 1 << 30
        a, b = ig.DeviceBuffer(n), ig.DeviceBuffer(n)
        for _ in range(5):
            ctx.copy_ms(b, a, n)
        log("copied 5 x 1 GiB device to device")
        return 0
    algs = tuple(a for a in args.algs.split(",") if a)
    if not set(algs) <= set(ALGS):
        print(f"unknown algorithm in --algs {args.algs}", file=sys.stderr)
        return 2

    # the CPU baseline first, in a child process, before this process uses the GPU
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline_isolated(args.cpu_scale or args.scale)
        except Exception as exc:  # the baseline is reported, never required
            log(f"cpu baseline skipped: {exc}")

    import indigo2 as ig

    t0 = time.time()
    g = ig.rmat(args.scale, args.edge_factor, seed=1, wmin=1, wmax=255)
    log(f"rmat{args.scale} ef{args.edge_factor}: V={g.nodes} E={g.edges} maxdeg={int(g.degree.max())} "
        f"({time.time() - t0:.1f}s)")
    dg = ig.DeviceGraph(g)
    out_i = ig.DeviceBuffer(4 * g.nodes)
    out_f = ig.DeviceBuffer(4 * g.nodes)
    ctx = ig.Context(0)
    # per-algorithm hints for this graph's class from the autotuner's database
    from indigo2.autotune import DEFAULT_DB, Tuning, bfs_depth, features, graph_class

    tuning = Tuning.load(DEFAULT_DB) if os.path.exists(DEFAULT_DB) else None
    klass = graph_class(features(g, bfs_depth(ctx, dg, g))) if tuning else None

    def run(alg: str):
        if tuning is not None:
            tuning.apply(ctx, alg, klass)
        if alg == "bfs":
            ctx.bfs_into(dg, 0, out_i)
        elif alg == "sssp":
            ctx.sssp_into(dg, 0, out_i)
        elif alg == "cc":
            ctx.cc_into(dg, out_i)
        elif alg == "mis":
            ctx.mis_into(dg, out_i)
        elif alg == "pr":
            return ctx.pagerank_into(dg, out_f)
        else:
            return ctx.tc(dg)
        return None

    for w in range(args.warmup):
        for a in algs:
            run(a)
        log(f"warmup step {w + 1}/{args.warmup} done")

    ms = {a: [] for a in algs}
    kms = {a: [] for a in algs}
    kname = {}
    pr_iters = 0
    tri = 0
    t_start = time.perf_counter()
    for s in range(args.steps):
        for a in algs:
            r = run(a)
            ms[a].append(ctx.last_ms)
            kname[a], k = ctx.last_kernel
            kms[a].append(k)
            if a == "pr":
                pr_iters = r
            elif a == "tc":
                tri = r
        log(f"step {s + 1}/{args.steps}: " + " ".join(f"{a}={ms[a][-1]:.3f}ms" for a in algs))
    wall = time.perf_counter() - t_start
    if algs != ALGS or args.profile_only:
        log("profiling run: no JSON line")
        return 0
    ms_per_step = 1e3 * wall / args.steps
    value = 6 * 1e-9 * g.edges / (ms_per_step * 1e-3)

    med = {a: statistics.median(ms[a]) for a in ALGS}
    kmed = {a: statistics.mean(kms[a]) for a in ALGS}
    per_gteps = {a: round(1e-9 * g.edges / (med[a] * 1e-3), 4) for a in ALGS}
    log("modelling TC work")
    tcm = tc_count_model(g)
    copy_gbs = copy_peak_gbs(ig, ctx)
    host_io = measure_host_transfer(ig, ctx, g, dg)
    roof = {}
    for a in ALGS:
        b = kernel_bytes(a, g, pr_iters, tcm)
        gbs = b / (kmed[a] * 1e-3) / 1e9
        # the one-pass model per algorithm over the whole call (all launches); TC: the
        # merge model of the reference's vertex form (tc_merge_bytes)
        b_call = tc_merge_bytes(g) if a == "tc" else b
        roof[a] = {"kernel": kname[a], "kernel_ms": round(kmed[a], 4), "bytes": b, "achieved_gbs": round(gbs, 1),
                   "frac": round(gbs / HBM_PEAK_GBS, 4), "call_ms": round(med[a], 4), "call_model_bytes": b_call,
                   "call_frac": round(b_call / (med[a] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    dom = max(ALGS, key=lambda a: med[a])
    workload = f"rmat{args.scale} ef{args.edge_factor}"
    traffic_path = os.path.join(REPO, "profiles", f"traffic_{workload.replace(' ', '_')}.json")
    for a in ALGS:  # measured HBM bytes of each main kernel, where recorded
        t = read_traffic(traffic_path, workload, kname[a])
        roof[a]["traffic"] = t["bytes_per_launch"] if t else None
    traffic = read_traffic(traffic_path, workload, kname[dom])

    line = {
        "metric": "GTEPS per algorithm (six Indigo2 kernels), aggregate over one step on 1 MI355X",
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 (PageRank fp32, fp64 residual)",
        "data": "synthetic",
        "config": {
            "workload": f"all six algorithms on rmat{args.scale} ef{args.edge_factor} a/b/c/d=.57/.19/.19/.05 "
                        f"seed 1, symmetrised, weights [1,255], source 0",
            "vertices": g.nodes,
            "csr_entries": g.edges,
            "parallelism": "single GPU",
        },
        "per_algorithm_gteps": per_gteps,
        "per_algorithm_ms": {a: round(med[a], 4) for a in ALGS},
        "pr_iterations": pr_iters,
        "triangles": tri,
        "autotuned": ({"db": os.path.relpath(DEFAULT_DB, REPO), "class": klass,
                       "hints": {a: {k: v for k, v in tuning.hints(a, klass).items() if v} for a in ALGS}}
                      if tuning else None),
        "roofline": {
            "bound": "hbm",
            "kernel": kname[dom],
            "algorithm": dom,
            "achieved": roof[dom]["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": roof[dom]["frac"],
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "algorithmic_bytes": roof[dom]["bytes"],
            "kernel_ms": roof[dom]["kernel_ms"],
            "traffic_source": traffic.get("source") if traffic else None,
            "copy_peak_gbs": round(copy_gbs, 1),
            "frac_of_copy_peak": round(roof[dom]["achieved_gbs"] / copy_gbs, 4),
            "tc_model": tcm,
            "per_algorithm": roof,
        },
        "cpu_baseline": cpu,
    }
    try:
        line["reordered"] = measure_reordered(
            ig, ctx, g, dg, (lambda a: tuning.apply(ctx, a, klass)) if tuning else (lambda a: None),
            args.steps, args.warmup)
    except Exception as exc:  # reported beside the main line, never required by it
        log(f"reordered measurement skipped: {exc}")
        line["reordered"] = None
    ctx.close()
    del dg, out_i, out_f
    line["host_transfer"] = host_io
    if not args.no_deep:
        for kind in ("grid", "road"):  # the high-diameter configurations, a few seconds each
            try:
                line[kind] = measure_deep(kind, 3, 1)
            except Exception as exc:  # reported beside the main line, never required by it
                log(f"{kind} measurement skipped: {exc}")
                line[kind] = None
    if not args.no_batched:
        try:
            line["batched"] = measure_batched()
        except Exception as exc:  # reported beside the main line, never required by it
            log(f"batched measurement skipped: {exc}")
            line["batched"] = None
    if not args.no_serving:
        if time.time() - t0 > 420:
            log("serving measurement skipped: the run has already taken over 7 minutes")
            line["serving"] = None
        else:
            try:
                line["serving"] = measure_serving((150.0,), slots=args.slots, max_batch=args.max_batch)
            except Exception as exc:  # reported beside the main line, never required by it
                log(f"serving measurement skipped: {exc}")
                line["serving"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
