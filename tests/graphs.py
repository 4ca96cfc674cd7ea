"""Small named test graphs (host CSR) shared by the CPU and GPU suites."""
import numpy as np


def path(ig, n):
    src = np.arange(n - 1)
    return ig.from_edges(n, src, src + 1, name=f"path{n}")


def star(ig, n):
    return ig.from_edges(n, np.zeros(n - 1, np.int32), np.arange(1, n), name=f"star{n}")


def clique(ig, n):
    s, d = np.triu_indices(n, 1)
    return ig.from_edges(n, s, d, name=f"K{n}")


def cliques(ig, sizes, seed=None):
    """Disjoint cliques of the given sizes on consecutive ids, or on ids
    shuffled by a seeded permutation."""
    src, dst, base = [], [], 0
    for k in sizes:
        s, d = np.triu_indices(k, 1)
        src.append(s + base)
        dst.append(d + base)
        base += k
    s, d = np.concatenate(src), np.concatenate(dst)
    if seed is not None:
        perm = np.random.default_rng(seed).permutation(base)
        s, d = perm[s], perm[d]
    return ig.from_edges(base, s.astype(np.int32), d.astype(np.int32), name=f"cliques{len(sizes)}")


def disjoint(ig):
    # two triangles, a path of 3, and isolated vertices 9..11
    s = np.array([0, 1, 2, 3, 4, 5, 6, 7], np.int32)
    d = np.array([1, 2, 0, 4, 5, 3, 7, 8], np.int32)
    return ig.from_edges(12, s, d, name="disjoint")


def single(ig):
    return ig.from_edges(1, np.zeros(0, np.int32), np.zeros(0, np.int32), name="single")


def weighted(ig, g, seed=7, wmin=1, wmax=255):
    """Same graph with symmetric uniform weights from the CSR builder."""
    s = np.repeat(np.arange(g.nodes, dtype=np.int32), g.degree)
    return ig.from_edges(g.nodes, s, g.nlist, symmetrize=True, wseed=seed, wmin=wmin, wmax=wmax, name=g.name + "_w")


def catalog(ig, weights=False):
    gs = [
        path(ig, 2), path(ig, 100), star(ig, 300), clique(ig, 12), disjoint(ig), single(ig),
        ig.grid(32, 32), ig.grid(7, 129), ig.rmat(10, 16, seed=1), ig.rmat(12, 8, seed=2),
        ig.erdos_renyi(3000, 12000, seed=3), ig.road(64, 64, seed=4),
    ]
    if weights:
        gs = [weighted(ig, g, seed=11 + i) for i, g in enumerate(gs)]
    return gs
