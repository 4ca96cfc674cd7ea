#!/usr/bin/env python3
"""Runs every drop-in harness on every input graph, in the output layout of
the reference's scripts/run_codes.py, so that its read_throughput.py and
generate_*_csv.sh turn the results into CSV rows beside the CUDA ones.

For each graph and algorithm it appends to throughputs/hip/<alg>_<graph>_hip.out,
per harness source harness/<alg>/<NAME>.cpp:

    (blank line)
    compile : <NAME>.cpp
    <the harness's own output, which has one "Throughput: X gigaedges/s" line>

The argv per algorithm is the reference's (run_codes.py:65-71): BFS and
SSSP get `input source verify`, PR gets `input`, the others `input verify`.
Each harness is compiled against the prebuilt lib/libindigo2_hip.so first
(`make -C <package> lib`), with the same command as the package Makefile.

    python3 run_codes_hip.py verify [--inputs DIR] [--graphs a.egr,b.egr]
                             [--algs bfs,sssp,...] [--out DIR] [--make-inputs]

--make-inputs writes the synthetic inputs (1024x1024 grid, R-MAT scale 22
edge factor 16, 4096x4096 road-like grid) into the inputs folder first.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
ALGS = ["sssp", "bfs", "cc", "mis", "pr", "tc"]  # run_codes.py's order
SOURCE = "0"
SYNTHETIC = {
    "grid1024.sym.egr": ("grid", dict(rows=1024, cols=1024, seed=1, wmin=1, wmax=255)),
    "rmat22.sym.egr": ("rmat", dict(scale=22, edge_factor=16, seed=1, wmin=1, wmax=255)),
    "road4096.sym.egr": ("road", dict(rows=4096, cols=4096, seed=1, wmin=1, wmax=1000)),
}


def make_inputs(folder: str) -> None:
    sys.path.insert(0, PKG)
    import indigo2 as ig

    os.makedirs(folder, exist_ok=True)
    for name, (kind, kw) in SYNTHETIC.items():
        path = os.path.join(folder, name)
        if os.path.exists(path):
            continue
        if kind == "grid":
            g = ig.grid(kw["rows"], kw["cols"], seed=kw["seed"], wmin=kw["wmin"], wmax=kw["wmax"])
        elif kind == "rmat":
            g = ig.rmat(kw["scale"], kw["edge_factor"], seed=kw["seed"], wmin=kw["wmin"], wmax=kw["wmax"])
        else:
            g = ig.road(kw["rows"], kw["cols"], seed=kw["seed"], wmin=kw["wmin"], wmax=kw["wmax"])
        g.write_egr(path)
        print(f"wrote {path}: {g.nodes} nodes, {g.edges} edges", flush=True)


def compile_harness(src: str, exe: str) -> bool:
    cmd = ["g++", "-O2", "-std=c++17", f"-I{REPO}/include", f"-I{PKG}/csrc", f"-I{PKG}/harness", src, "-o", exe,
           f"-L{PKG}/lib", "-lindigo2_hip", f"-Wl,-rpath,{PKG}/lib", "-pthread"]
    return subprocess.run(cmd).returncode == 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("verify", choices=["0", "1"])
    ap.add_argument("thread_count", nargs="?", default=None, help="accepted for run_codes.py compatibility; unused")
    ap.add_argument("--inputs", default="inputs")
    ap.add_argument("--graphs", default=None, help="comma-separated file names in --inputs (default: all *.egr)")
    ap.add_argument("--algs", default=",".join(ALGS))
    ap.add_argument("--out", default="throughputs")
    ap.add_argument("--make-inputs", action="store_true")
    ap.add_argument("--timeout", type=float, default=600.0, help="seconds per harness run")
    args = ap.parse_args()

    if not os.path.exists(os.path.join(PKG, "lib", "libindigo2_hip.so")):
        sys.exit(f"Error: {PKG}/lib/libindigo2_hip.so missing: run `make -C {PKG} lib`")
    if args.make_inputs:
        make_inputs(args.inputs)
    graphs = args.graphs.split(",") if args.graphs else sorted(f for f in os.listdir(args.inputs) if f.endswith(".egr"))
    algs = [a for a in ALGS if a in args.algs.split(",")]
    hip_out = os.path.join(args.out, "hip")
    os.makedirs(hip_out, exist_ok=True)
    exe = os.path.abspath("minibenchmark")
    total = sum(len([f for f in os.listdir(os.path.join(PKG, "harness", a)) if f.endswith(".cpp")]) for a in algs)
    failures = 0
    for gi, graph in enumerate(graphs):
        input_path = os.path.join(args.inputs, graph)
        counter = 0
        for alg in algs:
            out_name = os.path.join(hip_out, f"{alg}_{graph}_hip.out")
            if os.path.isfile(out_name):
                os.remove(out_name)
            code_dir = os.path.join(PKG, "harness", alg)
            for code_file in sorted(f for f in os.listdir(code_dir) if f.endswith(".cpp")):
                counter += 1
                print(f"running {code_file}\n{counter} out of {total} programs, {gi + 1} out of {len(graphs)} inputs",
                      flush=True)
                with open(out_name, "a") as f:
                    f.write(f"\ncompile : {code_file}\n")
                if not compile_harness(os.path.join(code_dir, code_file), exe):
                    sys.exit("Error: compile failed")
                if alg in ("sssp", "bfs"):
                    argv = [exe, input_path, SOURCE, args.verify]
                elif alg == "pr":
                    argv = [exe, input_path]
                else:
                    argv = [exe, input_path, args.verify]
                with open(out_name, "a") as f:
                    try:
                        rc = subprocess.run(argv, stdout=f, stderr=subprocess.STDOUT, timeout=args.timeout).returncode
                    except subprocess.TimeoutExpired:
                        rc = -1
                        f.write(f"ERROR: timed out after {args.timeout:.0f} s\n")
                if rc != 0:
                    failures += 1
                    print(f"  {code_file} exited with {rc}", flush=True)
                os.remove(exe)
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
