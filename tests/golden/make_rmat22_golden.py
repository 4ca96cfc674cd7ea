#!/usr/bin/env python3
"""Writes tests/golden/rmat22_ef16_s1.json: the benchmark graph's identity
and its triangle count, for the full-size GPU parity tests.

The graph is R-MAT scale 22, edge factor 16, a/b/c/d .57/.19/.19/.05,
seed 1, weights uniform in [1, 255], symmetrised, simple, sorted.  The
file holds:

  nodes, edges, max_degree   sizes of the generated CSR
  sha256                     digests of the raw little-endian int32 bytes of
                             nindex, nlist and eweight, so that any drift of
                             the generator fails the tests loudly instead of
                             comparing against a count of another graph
  triangles                  the count from the reference's own best
                             published OpenMP TC variant
                             (TC_OMP_V_IntType_ClauseRed_Dynamic, built from
                             the reference sources into oracle/_ref; its int
                             count holds 2,111,764,430 < 2^31), which must
                             equal the oracle's orc_tc (the restatement of
                             h_triCounting, include/tc_vertex_cpp.h:88-116)

The TC count is too slow to recompute inside a GPU test (about a minute on 8
threads for the reference, several for the oracle); BFS, SSSP, CC, MIS and
PageRank are recomputed by the oracle in the tests themselves.  Run from the
repo root after `make -C oracle` with the reference present:
    python tests/golden/make_rmat22_golden.py [threads]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "parallelization-artifact-sc23_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import indigo2 as ig  # noqa: E402
import oracle as orc  # noqa: E402

OUT = os.path.join(HERE, "rmat22_ef16_s1.json")
PARAMS = {"scale": 22, "edge_factor": 16, "a": 0.57, "b": 0.19, "c": 0.19, "seed": 1, "wmin": 1, "wmax": 255}


def csr_digests(g) -> dict:
    """sha256 of the raw int32 bytes of each CSR array."""
    return {name: hashlib.sha256(np.ascontiguousarray(arr, dtype="<i4").tobytes()).hexdigest()
            for name, arr in (("nindex", g.nindex), ("nlist", g.nlist), ("eweight", g.eweight))}


def make_graph():
    p = PARAMS
    return ig.rmat(p["scale"], p["edge_factor"], a=p["a"], b=p["b"], c=p["c"], seed=p["seed"], wmin=p["wmin"],
                   wmax=p["wmax"])


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    if not orc.ref_available("tc_omp"):
        sys.exit("oracle/_ref/libref_tc_omp.so missing: run `make -C oracle` with the reference present")
    g = make_graph()
    assert g.check() == 0
    rec = dict(PARAMS, nodes=g.nodes, edges=g.edges, max_degree=int(g.degree.max()), sha256=csr_digests(g))
    t = time.time()
    ref_count, _ = orc.ref_tc_omp(g, threads)
    print(f"reference TC_OMP_V_IntType_ClauseRed_Dynamic: {ref_count} ({time.time() - t:.1f} s)", flush=True)
    t = time.time()
    orc_count = orc.tc(g)
    print(f"oracle orc_tc: {orc_count} ({time.time() - t:.1f} s)", flush=True)
    if ref_count != orc_count:
        sys.exit(f"reference ({ref_count}) and oracle ({orc_count}) disagree")
    rec["triangles"] = orc_count
    rec["triangles_source"] = ("oracle/_ref TC_OMP_V_IntType_ClauseRed_Dynamic (reference sources) == oracle "
                               "orc_tc")
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
