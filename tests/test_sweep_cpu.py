"""The CPU-baseline sweep (oracle/sweep_cpu.py): argv per harness kind and the
throughput token, checked without the reference; and, where the reference
sources are present, one real variant built and run end to end."""
import importlib.util
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("REF", "/root/reference")


def _sweep():
    spec = importlib.util.spec_from_file_location("sweep_cpu", os.path.join(REPO, "oracle", "sweep_cpu.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_argv_follows_the_reference_harnesses():
    s = _sweep()
    # C++ threads harnesses take the thread count last (include/*_cpp.h usage lines)
    assert s.argv_for("bfs", "cpp", "g.egr", 5, 8) == ["g.egr", "5", "0", "8"]
    assert s.argv_for("sssp", "cpp", "g.egr", 0, 4) == ["g.egr", "0", "0", "4"]
    assert s.argv_for("pr", "cpp", "g.egr", 0, 4) == ["g.egr", "4"]
    for alg in ("cc", "mis", "tc"):
        assert s.argv_for(alg, "cpp", "g.egr", 0, 4) == ["g.egr", "0", "4"]
    # OpenMP harnesses take no thread count (OMP_NUM_THREADS instead)
    assert s.argv_for("bfs", "omp", "g.egr", 3, 8) == ["g.egr", "3", "0"]
    assert s.argv_for("pr", "omp", "g.egr", 0, 8) == ["g.egr"]
    assert s.argv_for("tc", "omp", "g.egr", 0, 8) == ["g.egr", "0"]


def test_throughput_is_the_second_to_last_token():
    s = _sweep()
    assert s.throughput("runtime: 0.1s\nThroughput: 1.250000 gigaedges/s\n") == 1.25
    assert s.throughput("no result\n") is None


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(os.path.join(REF, "codes", "omp", "pr-omp")),
                    reason="needs g++ and the reference sources")
def test_one_variant_end_to_end(tmp_path):
    s = _sweep()
    src = s.variants(REF, "pr", "omp")[0]
    os.makedirs(os.path.join(s.OUT, "omp"), exist_ok=True)
    exe, err = s.compile_one(REF, src, "omp")
    assert err is None, err
    golden = os.path.join(REPO, "tests", "golden")
    egr = sorted(os.path.join(golden, f) for f in os.listdir(golden) if f.endswith(".egr"))[0]
    r = subprocess.run([exe] + s.argv_for("pr", "omp", egr, 0, 2), capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert r.returncode == 0, r.stderr[-2000:]
    assert s.throughput(r.stdout) is not None
