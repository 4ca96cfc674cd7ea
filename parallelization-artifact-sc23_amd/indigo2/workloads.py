"""Seeded synthetic workloads shared by the benchmark and the autotuner."""
from __future__ import annotations

import numpy as np

from .graph import HostGraph, erdos_renyi, rmat


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
