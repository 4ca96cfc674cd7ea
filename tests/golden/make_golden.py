#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/: small graphs as ECLgraph
.egr files and, for each, the outputs of the reference's own C++ code
(compiled from the reference sources into oracle/_ref by oracle/Makefile):

  bfs_<src>   CPUserialDijkstra with unit weights (include/bfs_vertex_cpp.h:52)
  sssp_<src>  CPUserialDijkstra with the file's weights (include/sssp_vertex_cpp.h:88)
  cc          CPUcc_vertex of CC_CPP_V_Topo_Push_..._Block_Atomic.cpp
  mis         CPPmis_vertex of MIS_CPP_V_Topo_Push_IntType_Atomic_NonDeterm_Block.cpp
              (status 1 included / 2 excluded, unsigned priorities)
  pr, pr_iterations  PR_CPU of PR_CPP_Pull_FloatType_AtomicAdd_Block_Determin.cpp
  tc          h_triCounting (include/tc_vertex_cpp.h:88)

Outputs go to <name>.npz (plain arrays, no pickles).  Run from the repo root
after `make -C oracle` with the reference present:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "parallelization-artifact-sc23_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import indigo2 as ig  # noqa: E402
import oracle as orc  # noqa: E402


def graphs():
    def edges(n, pairs, name, w=None):
        s = np.array([p[0] for p in pairs], np.int32)
        d = np.array([p[1] for p in pairs], np.int32)
        if w is None:
            return ig.from_edges(n, s, d, name=name)
        return ig.from_edges(n, s, d, wseed=w[0], wmin=w[1], wmax=w[2], name=name)

    out = [
        edges(2, [(0, 1)], "path2", (1, 1, 9)),
        edges(7, [(i, i + 1) for i in range(6)], "path7", (2, 1, 255)),
        edges(9, [(0, i) for i in range(1, 9)], "star9", (3, 1, 255)),
        edges(4, [(a, b) for a in range(4) for b in range(a + 1, 4)], "k4", (4, 1, 255)),
        edges(6, [(a, b) for a in range(6) for b in range(a + 1, 6)], "k6", (5, 1, 100)),
        edges(14, [(0, 1), (1, 2), (2, 0), (3, 4), (4, 5), (5, 3), (5, 6), (7, 8), (9, 10)], "disjoint14",
              (6, 1, 50)),
        edges(5, [(1, 3)], "isolated5", (7, 1, 255)),
        ig.grid(32, 32, seed=8, wmin=1, wmax=255),
        ig.grid(5, 17, seed=9, wmin=1, wmax=1000),
        ig.rmat(10, 16, seed=1, wmin=1, wmax=255),
        ig.rmat(9, 8, seed=2, wmin=1, wmax=255),
        ig.erdos_renyi(600, 2400, seed=3, wmin=1, wmax=255),
        ig.road(20, 20, seed=4, wmin=1, wmax=1000),
    ]
    names = ["path2", "path7", "star9", "k4", "k6", "disjoint14", "isolated5", "grid32x32", "grid5x17", "rmat10",
             "rmat9_ef8", "er600", "road20x20"]
    for g, n in zip(out, names):
        g.name = n
    return out


def sources(g):
    return sorted({0, g.nodes // 2, g.nodes - 1})


def main():
    for alg in ("bfs", "sssp", "cc", "mis", "pr", "tc"):
        if not orc.ref_available(alg):
            sys.exit(f"oracle/_ref/libref_{alg}.so missing: run `make -C oracle` with the reference present")
    for g in graphs():
        assert g.check() == 0, g.name
        g.write_egr(os.path.join(HERE, g.name + ".egr"))
        out = {}
        for s in sources(g):
            out[f"bfs_{s}"], _ = orc.ref_bfs(g, s, serial=True)
            out[f"sssp_{s}"], _ = orc.ref_sssp(g, s, serial=True)
        out["cc"], _ = orc.ref_cc(g, 2)
        out["mis"], _, _ = orc.ref_mis(g, 2)
        out["pr"], it, _ = orc.ref_pr(g, 2)
        out["pr_iterations"] = np.array([it], np.int32)
        out["tc"] = np.array([orc.ref_tc(g, serial=True)[0]], np.int64)
        np.savez(os.path.join(HERE, g.name + ".npz"), **out)
        print(f"{g.name}: V={g.nodes} E={g.edges} mis={int((out['mis'] == 1).sum())} tc={int(out['tc'][0])} "
              f"pr_it={it}")


if __name__ == "__main__":
    main()
