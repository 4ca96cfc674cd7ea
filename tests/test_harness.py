"""The drop-in harness executables (parallelization-artifact-sc23_amd/harness)
and the run_codes twin: argv / stdout contract of the reference harnesses
(include/{alg}_cuda.h main()) and the output layout that the reference's
scripts/read_throughput.py parses (a "compile : <file>" header per program,
token [-2] of the line containing "Throughput" as the GTEPS value).

CPU tests check the usage errors (no GPU needed: argc is checked before the
device is touched); GPU tests run every harness with verification on."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "parallelization-artifact-sc23_amd")
BIN = os.path.join(PKG, "bin")
ARGC = {"BFS": 4, "SSSP": 4, "CC": 3, "MIS": 3, "TC": 3, "PR": 2}


def harnesses():
    srcs = []
    for alg in sorted(os.listdir(os.path.join(PKG, "harness"))):
        d = os.path.join(PKG, "harness", alg)
        if os.path.isdir(d):
            srcs += [os.path.splitext(f)[0] for f in sorted(os.listdir(d)) if f.endswith(".cpp")]
    return srcs


def throughputs(text):
    """read_throughput.py's rule: rows from 'compile' lines and 'Throughput' lines."""
    names = [l.split(" ")[-1].strip() for l in text.splitlines() if "compile" in l]
    values = [float(l.split(" ")[-2]) for l in text.splitlines() if "Throughput" in l]
    return names, values


def test_every_entry_form_has_a_harness():
    names = harnesses()
    assert len(names) == 11
    for alg, forms in {"BFS": "VE", "SSSP": "VE", "CC": "VE", "MIS": "VE", "TC": "VE"}.items():
        for f in forms:
            assert any(n.startswith(f"{alg}_HIP_{f}_") for n in names), (alg, f)
    assert any(n.startswith("PR_HIP_") for n in names)


@pytest.mark.skipif(not os.path.isdir(BIN), reason="harness executables not built (make -C <package> harness)")
def test_harness_usage_errors_without_gpu():
    for name in harnesses():
        exe = os.path.join(BIN, name)
        assert os.access(exe, os.X_OK), name
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0, name
        assert "USAGE" in (r.stdout + r.stderr), name
        assert "compile" not in r.stdout, name


def test_throughput_parse_rule():
    text = "\ncompile : A.cpp\nbfs (x)\nruntime: 0.5s\nThroughput: 2.500000 gigaedges/s\n"
    assert throughputs(text) == (["A.cpp"], [2.5])


@pytest.mark.gpu
def test_run_codes_twin_with_verification(ig, orc, tmp_path):
    g = ig.rmat(11, 8, seed=3, wmin=1, wmax=255)
    inputs = tmp_path / "inputs"
    inputs.mkdir()
    g.write_egr(str(inputs / "rmat11.sym.egr"))
    print("run_codes_hip on rmat11", flush=True)
    r = subprocess.run([sys.executable, os.path.join(PKG, "run_codes_hip.py"), "1", "--inputs", str(inputs),
                        "--out", str(tmp_path / "throughputs")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=110)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    expect_tc = orc.tc(g)
    for alg in ("bfs", "sssp", "cc", "mis", "pr", "tc"):
        text = (tmp_path / "throughputs" / "hip" / f"{alg}_rmat11.sym.egr_hip.out").read_text()
        names, values = throughputs(text)
        assert len(names) == len(values) >= 1, (alg, text)
        assert all(v > 0 for v in values), (alg, values)
        if alg in ("bfs", "sssp", "cc"):
            assert text.count("verification passed") == len(names), text
        elif alg == "mis":
            assert text.count("verification passed") == 3 * len(names), text
        elif alg == "pr":
            assert "All good." in text, text
        else:
            assert text.count(f"the pattern occurs {expect_tc} times") == len(names), text
        assert "ERROR" not in text, text


@pytest.mark.gpu
def test_bfs_harness_on_grid_prints_reference_lines(ig, tmp_path):
    g = ig.grid(64, 64)
    path = str(tmp_path / "grid64.egr")
    g.write_egr(path)
    exe = os.path.join(BIN, "BFS_HIP_V_DirOpt_Persist_Binned")
    print("BFS harness on grid64", flush=True)
    r = subprocess.run([exe, path, "0", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "nodes: 4096" in out and "source: 0" in out
    assert re.search(r"vertex \d+ has maximum distance 126", out), out
    assert "verification passed" in out
    assert len(throughputs(out)[1]) == 1
