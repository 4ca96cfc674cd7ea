"""Best-of-all-variants CPU baseline: builds every C++-threads and OpenMP
variant of the reference (codes/cpp/<alg>-cpp/*.cpp, codes/omp/<alg>-omp/*.cpp)
from its sources, runs each on one graph with the reference harness's own
argv, and reports the best throughput per algorithm.

TEST / MEASUREMENT INFRASTRUCTURE ONLY (a CPU baseline, never the product).
Needs the reference sources (REF, default /root/reference) and g++; every
binary goes to oracle/_ref/sweep/.  The compile lines are those of the
reference's scripts/run_codes.py:105 (g++ -O3 -pthread) and :145
(g++ -O3 -fopenmp); the run lines follow run_codes.py:106-111 and :147-152
(verify 0), except that OMP_NUM_THREADS really reaches the OpenMP binaries
(the script's `os.system('export ...')` at :146 sets it in a subshell only).
The throughput is the second-to-last token of the `Throughput` line, as
scripts/read_throughput.py:47-54 reads it; each harness prints the median of
its own timed runs.

    python3 oracle/sweep_cpu.py --rmat 16 --threads 8 --csv sweep.csv
    python3 oracle/sweep_cpu.py --egr graph.egr --algs tc --kinds omp

Where the reference sources are absent (the GPU box), --build-only here and
--prebuilt there run the binaries built beforehand into oracle/_ref/sweep/.
"""
from __future__ import annotations

import argparse
import csv
import os
import shutil
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_ref", "sweep")
ALGS = ("bfs", "sssp", "cc", "mis", "pr", "tc")
KINDS = ("cpp", "omp")


def variants(ref: str, alg: str, kind: str) -> list[str]:
    d = os.path.join(ref, "codes", kind, f"{alg}-{kind}")
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".cpp")) if os.path.isdir(d) else []


def prebuilt(alg: str, kind: str) -> list[str]:
    """Binaries already built into OUT/<kind>/ for one algorithm."""
    d = os.path.join(OUT, kind)
    if not os.path.isdir(d):
        return []
    pre = f"{alg}_{kind}_".upper()
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.upper().startswith(pre))


def compile_one(ref: str, src: str, kind: str) -> tuple[str, str | None]:
    """Builds src into OUT/<kind>/<name> unless an up-to-date binary exists;
    returns (binary, error or None)."""
    exe = os.path.join(OUT, kind, os.path.basename(src)[:-4])
    if os.path.exists(exe) and os.path.getmtime(exe) >= os.path.getmtime(src):
        return exe, None
    flag = "-pthread" if kind == "cpp" else "-fopenmp"
    r = subprocess.run(["g++", src, "-O3", flag, "-I" + os.path.join(ref, "include"), "-w", "-o", exe],
                       capture_output=True, text=True)
    return exe, None if r.returncode == 0 else (r.stderr.strip().splitlines() or ["compile failed"])[-1]


def argv_for(alg: str, kind: str, egr: str, source: int, threads: int) -> list[str]:
    if kind == "cpp":
        if alg in ("bfs", "sssp"):
            return [egr, str(source), "0", str(threads)]
        return [egr, str(threads)] if alg == "pr" else [egr, "0", str(threads)]
    if alg in ("bfs", "sssp"):
        return [egr, str(source), "0"]
    return [egr] if alg == "pr" else [egr, "0"]


def throughput(stdout: str) -> float | None:
    for line in stdout.splitlines():
        if "Throughput" in line:
            try:
                return float(line.split(" ")[-2])
            except (IndexError, ValueError):
                return None
    return None


def make_graph(scale: int, ef: int) -> str:
    sys.path.insert(0, os.path.join(ROOT, "parallelization-artifact-sc23_amd"))
    import indigo2 as ig  # host-side generator only; no GPU is used

    path = os.path.join(OUT, f"rmat{scale}_ef{ef}.egr")
    if not os.path.exists(path):
        ig.rmat(scale, ef, seed=1, wmin=1, wmax=255).write_egr(path)
    return path


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    src = ap.add_mutually_exclusive_group(required=True)
    src.add_argument("--egr", help="graph file (ECLgraph format)")
    src.add_argument("--rmat", type=int, help="generate an R-MAT graph of this scale (seed 1, weights 1..255)")
    ap.add_argument("--ef", type=int, default=16, help="R-MAT edge factor")
    ap.add_argument("--ref", default=os.environ.get("REF", "/root/reference"))
    ap.add_argument("--algs", default=",".join(ALGS))
    ap.add_argument("--kinds", default=",".join(KINDS))
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1), help="parallel compiles")
    ap.add_argument("--source", type=int, default=0)
    ap.add_argument("--timeout", type=float, default=120.0, help="seconds per variant run")
    ap.add_argument("--csv", default=None, help="write one row per variant")
    ap.add_argument("--build-only", action="store_true", help="compile every variant into oracle/_ref/sweep and stop")
    ap.add_argument("--prebuilt", action="store_true",
                    help="run the binaries already in oracle/_ref/sweep (no reference sources needed)")
    args = ap.parse_args()
    algs = [a for a in args.algs.split(",") if a]
    kinds = [k for k in args.kinds.split(",") if k]
    if args.prebuilt:
        todo = [(a, k, b) for a in algs for k in kinds for b in prebuilt(a, k)]
        built = [(b, None) for _, _, b in todo]
    else:
        if shutil.which("g++") is None or not os.path.isdir(os.path.join(args.ref, "codes")):
            print("needs g++ and the reference sources (or --prebuilt)", file=sys.stderr)
            return 2
        for k in kinds:
            os.makedirs(os.path.join(OUT, k), exist_ok=True)
        todo = [(a, k, s) for a in algs for k in kinds for s in variants(args.ref, a, k)]
        with ThreadPoolExecutor(args.jobs) as pool:
            built = list(pool.map(lambda t: compile_one(args.ref, t[2], t[1]), todo))
        if args.build_only:
            bad = [(s, e) for (_, _, s), (_, e) in zip(todo, built) if e is not None]
            print(f"built {len(todo) - len(bad)} of {len(todo)} variants into {OUT}", flush=True)
            for s, e in bad:
                print(f"  {os.path.basename(s)}: {e}")
            return 0
    if not todo:
        print("no variants to run", file=sys.stderr)
        return 2
    egr = os.path.abspath(args.egr) if args.egr else make_graph(args.rmat, args.ef)
    print(f"{len(todo)} variants, graph {egr}, {args.threads} threads", flush=True)

    best = {}
    env = dict(os.environ, OMP_NUM_THREADS=str(args.threads))
    # rows are written as they finish, so a run cut short keeps what it measured
    out = open(args.csv, "w", newline="") if args.csv else None
    w = csv.DictWriter(out, fieldnames=["algorithm", "kind", "variant", "gteps", "note"]) if out else None
    if w:
        w.writeheader()
    for (alg, kind, s), (exe, err) in zip(todo, built):
        name = os.path.basename(s)
        name = name if name.endswith(".cpp") else name + ".cpp"
        gteps, note = None, err
        if err is None:
            t0 = time.time()
            try:
                r = subprocess.run([exe] + argv_for(alg, kind, egr, args.source, args.threads), capture_output=True,
                                   text=True, timeout=args.timeout, env=env)
                gteps = throughput(r.stdout)
                note = None if gteps is not None else f"exit {r.returncode}, no Throughput line"
            except subprocess.TimeoutExpired:
                note = f"timeout after {args.timeout:.0f} s"
            print(f"{alg:4s} {kind:3s} {gteps if gteps is not None else '-':>10} {time.time() - t0:6.1f}s  {name}"
                  + (f"  ({note})" if note else ""), flush=True)
        if w:
            w.writerow({"algorithm": alg, "kind": kind, "variant": name, "gteps": gteps, "note": note or ""})
            out.flush()
        if gteps is not None and gteps > best.get((alg, kind), (0.0, ""))[0]:
            best[(alg, kind)] = (gteps, name)
    if out:
        out.close()
    print("best per algorithm:")
    for alg in algs:
        for kind in kinds:
            if (alg, kind) in best:
                g, n = best[(alg, kind)]
                print(f"  {alg:4s} {kind:3s} {g:.6f} GTEPS  {n}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
