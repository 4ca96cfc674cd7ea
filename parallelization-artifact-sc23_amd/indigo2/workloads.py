"""Seeded synthetic workloads shared by the benchmark and the autotuner."""
from __future__ import annotations

import numpy as np

from .graph import HostGraph, erdos_renyi, grid, rmat


def batched_mix(ngraphs: int = 512, seed: int = 2024) -> tuple[list[HostGraph], list[tuple[np.ndarray, np.ndarray]]]:
    """The batched configuration: `ngraphs` seeded graphs with vertex counts
    log-uniform in [2^10, 2^20]; even-indexed ones R-MAT edge factor 8
    (a/b/c/d .57/.19/.19/.05, 2^round(log2 n) vertices: the generator works on
    powers of two), odd-indexed ones Erdos-Renyi with 4n undirected samples
    (average degree 8); weights uniform in [1, 255].  Returns the graphs and,
    per graph, 64 BFS sources and 16 SSSP sources drawn uniformly."""
    rng = np.random.default_rng(seed)
    graphs, srcs = [], []
    for i in range(ngraphs):
        n = int(round(2.0 ** rng.uniform(10, 20)))
        if i % 2 == 0:
            g = rmat(int(round(np.log2(n))), 8, seed=seed + i, wmin=1, wmax=255)
        else:
            g = erdos_renyi(n, 4 * n, seed=seed + i, wmin=1, wmax=255)
        graphs.append(g)
        srcs.append((rng.integers(0, g.nodes, 64), rng.integers(0, g.nodes, 16)))
    return graphs, srcs


def batched_specs(srcs, algs=("bfs", "sssp", "cc", "mis")) -> list[tuple[str, int, list[int]]]:
    """(algorithm, graph index, sources) for every query of the batched
    configuration: 64-source BFS and 16-source SSSP, one CC and one MIS per
    graph."""
    out = []
    for alg in algs:
        for i, (b, s) in enumerate(srcs):
            src = b if alg == "bfs" else (s if alg == "sssp" else [])
            out.append((alg, i, [int(x) for x in src]))
    return out


def serving_pool(seed: int = 77) -> tuple[list[HostGraph], list[str]]:
    """The serving configuration's 64 resident graphs: 32 R-MAT (edge factor
    16, a/b/c/d .57/.19/.19/.05) with scales 14, 15, ..., 22 in turn, and 32
    square grids with sides 256, 512, 1024, 2048 in turn; weights [1, 255].
    Returns the graphs and their short names."""
    gs, names = [], []
    for i in range(32):
        sc = 14 + i % 9
        gs.append(rmat(sc, 16, seed=seed + i, wmin=1, wmax=255))
        names.append(f"rmat{sc}")
    for i in range(32):
        side = 256 << (i % 4)
        gs.append(grid(side, side, seed=seed + 100 + i, wmin=1, wmax=255))
        names.append(f"grid{side}")
    return gs, names


def serving_requests(graphs: list[HostGraph], nq: int, seed: int = 5):
    """`nq` seeded requests over the pool: algorithm mix BFS .4 / SSSP .3 /
    CC .1 / PageRank .1 / TC .1, graph and source uniform, and unit-mean
    exponential gaps (divide by the rate for Poisson arrivals).  Returns
    (algs, graph ids, sources, gaps)."""
    rng = np.random.default_rng(seed)
    algs = list(rng.choice(["bfs", "sssp", "cc", "pr", "tc"], size=nq, p=[0.4, 0.3, 0.1, 0.1, 0.1]))
    gid = rng.integers(0, len(graphs), nq)
    src = np.array([rng.integers(0, graphs[g].nodes) for g in gid])
    gaps = rng.exponential(1.0, nq)
    return algs, gid, src, gaps


# The serving loop's cost model (csrc/server.hip est_us): per algorithm a
# fixed part in microseconds and a rate in CSR entries per microsecond,
# fitted to the median service times of the serving pool at 150 requests/s.
_SERVE_COST = {"bfs": (2000.0, 10000.0), "sssp": (2000.0, 3000.0), "cc": (2000.0, 50000.0),
               "mis": (2000.0, 10000.0), "pr": (13000.0, 1600.0), "tc": (2000.0, 2000.0)}
SERVE_LONG_US = 8000.0


def serving_cost_us(g: HostGraph, alg: str) -> float:
    """The serving loop's estimate of one request's microseconds of service
    on a slot (the same rule as indigo_serve's admission): a fixed part plus
    CSR entries at the algorithm's rate, or 2 sqrt(V) levels at 10 / 20 us
    for BFS / SSSP on deep sparse graphs (average degree below 6, at least
    65536 vertices).  Requests above SERVE_LONG_US are long."""
    if g.nodes >= (1 << 16) and g.edges < 6 * g.nodes and alg in ("bfs", "sssp"):
        return 2.0 * float(np.sqrt(g.nodes)) * (10.0 if alg == "bfs" else 20.0)
    fixed, rate = _SERVE_COST[alg]
    return fixed + g.edges / rate
