"""Autotuner: picks, per algorithm and graph class, the launch knobs the
kernels expose, from measurements on the GPU it runs on, and keeps them in a
JSON database that the Python layer applies at call time.

Graph classes are built from the graph properties that the suite's study
correlates with variant performance (analysis/analysis_median.ipynb, cell 1:
numv, nume, dAvg, dMax, share of degrees >= 32, share of degrees >= 512,
diameter).

The knobs stand for the suite's variant axes (the style tokens of
analysis_median.ipynb, cell 1, and the file names under codes/cuda/*):

  V / E and Topo / Data
                       INDIGO_HINT_VARIANT: the vertex-centric data-driven
                       kernels (degree-binned worklists) or the edge-centric
                       topology-driven ones (one lane per CSR entry, every
                       entry in every sweep; BFS, SSSP, CC, MIS)
  Push / Pull          INDIGO_HINT_BFS_DIRECTION: 1 top-down (push) only,
                       0 direction-optimising (push and pull)
  Data worklist width  INDIGO_HINT_SSSP_DELTA: the near/far split width of
                       the data-driven SSSP worklist
  Thread / Warp / Block
                       the gang size of a batched task (workgroups per task)
                       and INDIGO_HINT_SMALL_FRONTIER (levels run by one
                       workgroup while the frontier is small)
  ThreadsPerBlock      INDIGO_HINT_BLOCK_SIZE: 256, 512 or 1024 threads per
                       workgroup of the edge-centric kernels (the suite's
                       ThreadsPerBlock, include/pr_cuda.h:44)
  Persist (grid size)  INDIGO_HINT_BLOCKS_PER_CU: resident workgroups per CU
                       of the persistent grid

Priors come from the suite's own measurements: the best published CUDA
variant per algorithm on its road-like grid (2d-2e20) and on its R-MAT
(rmat22) (data/*_cuda_32.csv), mapped onto these
knobs (PRIORS).  Each algorithm's prior is measured first, and another
setting replaces it only when it is faster by the required margin.

Every candidate is checked against the default configuration's output
(bit-exact; PageRank to 1e-6 per vertex), so a tuned setting can change
speed only.

Use:
    python -m indigo2.autotune --out tuned.json            # R-MAT 22 and the batched mix
    db = Tuning.load("tuned.json"); db.apply(ctx, "bfs", feats)
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import statistics
import sys
import time

import numpy as np

from . import _lib
from .graph import DeviceGraph, HostGraph
from .kernels import Context, Query, batch_gang

DEFAULT_DB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_mi355x.json")

HINTS = {"BFS_DIRECTION": _lib.HINT_BFS_DIRECTION, "SSSP_DELTA": _lib.HINT_SSSP_DELTA,
         "SMALL_FRONTIER": _lib.HINT_SMALL_FRONTIER, "BLOCKS_PER_CU": _lib.HINT_BLOCKS_PER_CU,
         "VARIANT": _lib.HINT_VARIANT, "BLOCK_SIZE": _lib.HINT_BLOCK_SIZE}
# hints every single-query tuning resets before it applies its own
_TUNED_HINTS = tuple(HINTS.values())

BATCH_ALGS = ("bfs", "sssp", "cc", "mis")

# The suite's best published variant per algorithm and graph shape, mapped
# onto the knobs above: (hints, the variant, its row in the reference data).
# Shapes: "deep" is the road-like 2d-2e20 grid, "skewed" the R-MAT rmat22.
EDGE = {"VARIANT": _lib.VARIANT_EDGE}
PRIORS = {
    "deep": {
        "bfs": ({}, "V_Data_Push_NonDeterm_RMW_Persist_Warp_Dup", "data/bfs_cuda_32.csv:772"),
        "sssp": ({}, "V_Data_Push_NonDeterm_RW_Persist_Warp_NonDup", "data/sssp_cuda_32.csv:1168"),
        "cc": (EDGE, "E_Topo_Push_NonDeterm_RW_Thread", "data/cc_cuda_32.csv:1370"),
        "mis": ({}, "V_Topo_Push_NonDeterm_NonPersist_Thread", "data/mis_cuda_32.csv:140"),
        "pr": ({}, "Pull_Persist_Thread_Determ_Reduction", "data/pr_cuda_32.csv:184"),
    },
    "skewed": {
        "bfs": ({}, "V_Topo_Push_Determ_RW_NonPersist_Thread", "data/bfs_cuda_32.csv:290"),
        "sssp": ({}, "V_Data_Push_NonDeterm_RW_Persist_Warp_NonDup", "data/sssp_cuda_32.csv:1456"),
        "cc": ({}, "V_Data_Push_NonDeterm_RW_NonPersist_Warp_NonDup", "data/cc_cuda_32.csv:138"),
        "mis": ({}, "V_Topo_Push_NonDeterm_NonPersist_Thread", "data/mis_cuda_32.csv:364"),
        "pr": ({}, "Pull_NonPersist_Thread_Determ_Reduction", "data/pr_cuda_16.csv:76"),
    },
}


def prior(alg: str, klass: str) -> tuple[dict, str, str]:
    """(hints, reference variant, data row) of the suite's best published
    variant for this algorithm on graphs of this shape ({} when none)."""
    return PRIORS.get(klass.split("-")[0], {}).get(alg, ({}, "", ""))
GANG_SCALES = (0.25, 0.5, 2.0, 4.0)


def log(msg: str) -> None:
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


# ---- graph classes ----------------------------------------------------------
def features(g: HostGraph, diameter: int | None = None) -> dict:
    """The suite's graph properties (analysis_median.ipynb, cell 1).
    `diameter` is a lower bound (BFS levels from one vertex) or None."""
    deg = g.degree
    v = max(1, g.nodes)
    return {"numv": int(g.nodes), "nume": int(g.edges), "dAvg": float(g.edges) / v,
            "dMax": int(deg.max()) if g.nodes else 0, "d32": float(np.count_nonzero(deg >= 32)) / v,
            "d512": float(np.count_nonzero(deg >= 512)) / v, "diameter": diameter}


def bfs_depth(ctx: Context, dg: DeviceGraph, g: HostGraph) -> int:
    """Levels of a BFS from the highest-degree vertex: a lower bound on the
    diameter, measured on the GPU."""
    if g.nodes == 0:
        return 0
    dist = ctx.bfs(dg, int(np.argmax(g.degree)))
    reached = dist[dist != np.iinfo(np.int32).max]
    return int(reached.max()) + 1


def graph_class(f: dict) -> str:
    """shape-size.  shape: deep (BFS depth above 100: grids, roads), skewed
    (largest degree above 8x the mean: R-MAT, social), else uniform (ER);
    size by vertex count: small < 2^16 <= medium < 2^20 <= large."""
    d = f.get("diameter")
    if d is not None and d > 100:
        shape = "deep"
    elif f["dMax"] > 8 * max(f["dAvg"], 1.0):
        shape = "skewed"
    else:
        shape = "uniform"
    n = f["numv"]
    size = "small" if n < (1 << 16) else ("medium" if n < (1 << 20) else "large")
    return f"{shape}-{size}"


# ---- the database -----------------------------------------------------------
class Tuning:
    """Tuned settings: single[class][alg] = {"hints": {name: value}, ...} and
    batched[class][alg] = {"gang_scale": s, ...}.  Lookups fall back to a
    class of the same shape, then to the engine's defaults."""

    def __init__(self, db: dict | None = None):
        self.db = db or {"single": {}, "batched": {}}
        self.db.setdefault("single", {})
        self.db.setdefault("batched", {})

    @classmethod
    def load(cls, path: str = DEFAULT_DB) -> "Tuning":
        with open(path) as f:
            return cls(json.load(f))

    def save(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump(self.db, f, indent=1, sort_keys=True)
            f.write("\n")

    def _lookup(self, table: str, alg: str, klass: str) -> dict | None:
        t = self.db[table]
        if klass in t and alg in t[klass]:
            return t[klass][alg]
        shape = klass.split("-")[0]
        for k in sorted(t):
            if k.split("-")[0] == shape and alg in t[k]:
                return t[k][alg]
        return None

    def hints(self, alg: str, klass: str) -> dict[int, int]:
        """Hint key -> value for a single query (all tuned hints, 0 = default)."""
        out = {k: 0 for k in _TUNED_HINTS}
        e = self._lookup("single", alg, klass)
        if e:
            for name, val in e["hints"].items():
                out[HINTS[name]] = int(val)
        return out

    def apply(self, ctx: Context, alg: str, klass: str) -> None:
        for k, v in self.hints(alg, klass).items():
            ctx.set_hint(k, v)

    def gang(self, alg: str, klass: str, nodes: int, edges: int) -> int:
        """Gang size for a batched query: the engine's rule times the tuned
        scale, within 1..INDIGO_MAX_GANG; 0 (the engine's rule) if untuned."""
        e = self._lookup("batched", alg, klass)
        if not e or e["gang_scale"] == 1.0:
            return 0
        return scaled_gang(alg, nodes, edges, e["gang_scale"])


def scaled_gang(alg: str, nodes: int, edges: int, scale: float) -> int:
    k = batch_gang(alg, nodes, edges) * scale
    return int(min(_lib.MAX_GANG, max(1, round(k))))


# ---- single-query tuning ----------------------------------------------------
def _space(alg: str, g: HostGraph) -> list[dict]:
    """Candidate hint settings of one algorithm ({} is the default): the
    vertex-centric kernels over their knobs, and the edge-centric kernels at
    each workgroup size."""
    blocks = [0, 1, 2, 4]
    edge = [dict(EDGE, **({"BLOCK_SIZE": b} if b else {})) for b in (0, 256, 1024)]
    if alg == "bfs":
        axes = {"BFS_DIRECTION": [0, 1], "SMALL_FRONTIER": [0, 1], "BLOCKS_PER_CU": blocks}
    elif alg == "sssp":
        w = g.eweight
        mean_w = float(w.mean()) if w is not None and len(w) else 1.0
        auto = mean_w * 32.0 / max(1.0, g.edges / max(1, g.nodes))
        deltas = [0] + sorted({max(1, int(round(auto * m))) for m in (0.25, 0.5, 2.0, 4.0)})
        axes = {"SSSP_DELTA": deltas, "BLOCKS_PER_CU": blocks}
    elif alg == "mis":
        axes = {"BLOCKS_PER_CU": blocks}
    elif alg == "cc":
        axes = {}
    elif alg == "pr":
        return [{n: v for n, v in (("BLOCKS_PER_CU", b),) if v} for b in blocks]
    else:
        return [{}]
    names = list(axes)
    vertex = [{n: v for n, v in zip(names, vals) if v != 0} for vals in itertools.product(*(axes[n] for n in names))]
    return vertex + edge


def _run_single(ctx: Context, alg: str, dg: DeviceGraph, src: int):
    if alg == "bfs":
        return ctx.bfs(dg, src)
    if alg == "sssp":
        return ctx.sssp(dg, src)
    if alg == "cc":
        return ctx.cc(dg)
    if alg == "mis":
        return ctx.mis(dg)
    return ctx.pagerank(dg)[0]


def _same(alg: str, a: np.ndarray, b: np.ndarray) -> bool:
    if alg == "pr":
        return bool(np.abs(a.astype(np.float64) - b).max() <= 1e-6)
    return bool(np.array_equal(a, b))


def tune_single(ctx: Context, g: HostGraph, dg: DeviceGraph, klass: str, algs=("bfs", "sssp", "cc", "mis", "pr"),
                reps: int = 3, src: int = 0, gain: float = 0.03) -> dict:
    """Times every candidate of each algorithm (median of `reps` calls after
    one warm-up; ctx.last_ms, the device time of the call), the prior first,
    and keeps the prior unless another candidate whose output equals the
    default's beats it by `gain`."""
    out = {}
    for alg in algs:
        for k in _TUNED_HINTS:
            ctx.set_hint(k, 0)
        ref = _run_single(ctx, alg, dg, src)
        p_hints, p_variant, p_row = prior(alg, klass)
        space = _space(alg, g)
        cands = [p_hints] + [c for c in space if c != p_hints]
        results = []
        for cand in cands:
            for k in _TUNED_HINTS:
                ctx.set_hint(k, 0)
            for name, val in cand.items():
                ctx.set_hint(HINTS[name], val)
            res = _run_single(ctx, alg, dg, src)
            if not _same(alg, res, ref):
                raise RuntimeError(f"autotune: {alg} with {cand} changed the result")
            ts = []
            for _ in range(reps):
                _run_single(ctx, alg, dg, src)
                ts.append(ctx.last_ms)
            results.append((statistics.median(ts), cand))
            log(f"{klass} {alg} {cand or 'default'}: {results[-1][0]:.3f} ms")
        default_ms = next(t for t, c in results if not c)
        prior_ms = results[0][0]
        best_ms, best = min(results, key=lambda r: r[0])
        if best_ms > (1.0 - gain) * prior_ms:
            best_ms, best = prior_ms, p_hints
        out[alg] = {"hints": best, "ms": round(best_ms, 4), "default_ms": round(default_ms, 4),
                    "prior": {"hints": p_hints, "ms": round(prior_ms, 4), "variant": p_variant, "source": p_row},
                    "candidates": len(results)}
    for k in _TUNED_HINTS:
        ctx.set_hint(k, 0)
    return out


# ---- batched tuning ---------------------------------------------------------
def _set_gangs(qs, sub, graphs, scales: dict, classes: list[str]) -> None:
    for q, (alg, gi, _) in zip(qs, sub):
        s = scales.get(classes[gi], 1.0)
        q.gang = 0 if s == 1.0 else scaled_gang(alg, graphs[gi].nodes, graphs[gi].edges, s)


def _time_batch(ctx: Context, graphs, qs, reps: int) -> float:
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        ctx.batch(graphs, qs)
        ts.append(1e3 * (time.perf_counter() - t))
    return statistics.median(ts)


def _same_result(alg: str, a, b) -> bool:
    if alg == "pr":
        return bool(np.abs(a[0].astype(np.float64) - b[0]).max() <= 1e-6 and abs(a[1] - b[1]) <= 1)
    return bool(np.array_equal(a, b))


def tune_batched(ctx: Context, graphs: list[DeviceGraph], classes: list[str], specs, algs=BATCH_ALGS,
                 reps: int = 3, check: int = 12, seed: int = 1, gain: float = 0.02) -> dict:
    """Per algorithm, on the real mix (specs = [(alg, graph index, sources)];
    every query of one algorithm runs in one batch): each class's gang
    scale is tried at 1/4, 1/2, 2 and 4 times the engine's rule with every
    other class at the rule, and the best scale of each class that beats the
    default by `gain` is kept; the combination is then measured once more
    and kept only if it also beats the default by `gain` (batches of a few
    hundred ms vary by a few per cent from run to run, so single greedy
    steps are not trusted).  Finally `check` queries drawn at random, and
    the largest query of each class, are compared with the engine's
    defaults (bit for bit; PageRank to 1e-6)."""
    rng = np.random.default_rng(seed)
    out: dict[str, dict] = {}
    for alg in algs:
        sub = [s for s in specs if s[0] == alg]
        if not sub:
            continue
        present = sorted({classes[gi] for _, gi, _ in sub},
                         key=lambda c: -sum(graphs[gi].edges for _, gi, _ in sub if classes[gi] == c))
        qs = [Query(a, gi, graphs[gi], srcs) for a, gi, srcs in sub]
        ctx.batch(graphs, qs)  # warm-up, and the defaults' results
        picks = set(int(i) for i in rng.choice(len(qs), min(check, len(qs)), replace=False))
        for c in present:
            picks.add(max((i for i, (_, gi, _) in enumerate(sub) if classes[gi] == c),
                          key=lambda i: graphs[sub[i][1]].edges))
        want = {i: qs[i].result() for i in sorted(picks)}
        default_ms = _time_batch(ctx, graphs, qs, reps)
        log(f"batched {alg} default: {default_ms:.1f} ms")
        best = {c: 1.0 for c in present}
        for c in present:
            best_t = default_ms
            for sc in GANG_SCALES:
                _set_gangs(qs, sub, graphs, {c: sc}, classes)
                t = _time_batch(ctx, graphs, qs, reps)
                log(f"batched {alg} {c} x{sc}: {t:.1f} ms")
                if t < (1.0 - gain) * best_t:
                    best_t, best[c] = t, sc
        scales = {c: 1.0 for c in present}
        best_ms = default_ms
        if any(v != 1.0 for v in best.values()):
            _set_gangs(qs, sub, graphs, best, classes)
            t = _time_batch(ctx, graphs, qs, reps)
            log(f"batched {alg} combined {best}: {t:.1f} ms")
            if t < (1.0 - gain) * default_ms:
                scales, best_ms = best, t
        _set_gangs(qs, sub, graphs, scales, classes)
        ctx.batch(graphs, qs)
        for i, w in want.items():
            if not _same_result(alg, qs[i].result(), w):
                raise RuntimeError(f"autotune: batched {alg} query {i} changed under gang scales {scales}")
        for c in present:
            out.setdefault(c, {})[alg] = {"gang_scale": scales[c], "batch_ms": round(best_ms, 2),
                                          "batch_default_ms": round(default_ms, 2)}
        log(f"batched {alg}: {default_ms:.1f} -> {best_ms:.1f} ms, scales {scales}")
        del qs
    return out


def classify(ctx: Context, hosts: list[HostGraph], graphs: list[DeviceGraph]) -> list[str]:
    return [graph_class(features(h, bfs_depth(ctx, d, h))) for h, d in zip(hosts, graphs)]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="tune the kernels' knobs on this GPU and write the database")
    ap.add_argument("--out", default=DEFAULT_DB)
    ap.add_argument("--scale", type=int, default=22, help="R-MAT scale of the single-query tuning graph")
    ap.add_argument("--batched-graphs", type=int, default=512, help="graphs in the batched mix (0: skip)")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)

    from . import graph as gmod
    from .workloads import batched_mix, batched_specs

    db = Tuning()
    db.db["device"] = "MI355X (gfx950)"
    with Context(0) as ctx:
        cases = [("rmat", gmod.rmat(args.scale, 16, seed=1, wmin=1, wmax=255)),
                 ("road", gmod.road(1024, 1024, keep_p=0.85, diag_p=0.02, seed=1, wmin=1, wmax=1000))]
        for name, g in cases:
            dg = DeviceGraph(g)
            f = features(g, bfs_depth(ctx, dg, g))
            klass = graph_class(f)
            log(f"{name}: class {klass}, features {f}")
            entry = tune_single(ctx, g, dg, klass, reps=args.reps)
            for alg, e in entry.items():
                e["tuned_on"] = name
            db.db["single"].setdefault(klass, {}).update(entry)
            db.db.setdefault("features", {})[klass] = f
            dg.close()
        if args.batched_graphs > 0:
            hosts, srcs = batched_mix(args.batched_graphs)
            specs = batched_specs(srcs)
            graphs = [DeviceGraph(h) for h in hosts]
            classes = classify(ctx, hosts, graphs)
            log("batched classes: " + ", ".join(f"{c} {classes.count(c)}" for c in sorted(set(classes))))
            db.db["batched"] = tune_batched(ctx, graphs, classes, specs, reps=args.reps)
            for d in graphs:
                d.close()
    db.save(args.out)
    log(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
