"""Host parts of the drop-in harnesses, without a GPU.

tests/harness_host_check.cpp runs harness/common.hpp's weight synthesis,
edge-source array and median on an .egr file stored without weights.  The
weights are checked against the oracle's restatement of the reference's rule
(include/sssp_vertex_cuda.h:127-137, include/mis_vertex_cuda.h:117-127), which
is itself pinned here by values worked out by hand, two of them with a 32-bit
product that wraps; sp against its definition (include/bfs_edge_cuda.h:140-146)
and the medians against the reference's median (bfs_vertex_cuda.h:111-118:
the mean of the two middle values for an even count).  Built with ASan/UBSan.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallelization-artifact-sc23_amd")

# (i, nei, weight) on a 100,000-vertex graph:
#   70000 * 80000 = 5.6e9 wraps to  1305032704; % 100000 =  32704 -> 32705
#   50000 * 50001 = 2.5e9 wraps to -1794917296; % 100000 = -17296 -> |-17295|
#   3 * 7 = 21 -> 22
KNOWN = [(70000, 80000, 32705), (50000, 50001, 17295), (3, 7, 22)]


def _graph(ig):
    rng = np.random.default_rng(5)
    src = np.concatenate([[a for a, _, _ in KNOWN], rng.integers(0, 100000, 3000)])
    dst = np.concatenate([[b for _, b, _ in KNOWN], rng.integers(0, 100000, 3000)])
    return ig.from_edges(100000, src, dst, name="wrap100k")


def _weight(g, w, i, nei):
    lo, hi = g.nindex[i], g.nindex[i + 1]
    j = lo + int(np.searchsorted(g.nlist[lo:hi], nei))
    assert g.nlist[j] == nei
    return int(w[j])


def test_oracle_weight_synthesis_known_values(ig, orc):
    g = _graph(ig)
    w = orc.synth_weights(g)
    for i, nei, want in KNOWN:
        assert _weight(g, w, i, nei) == want
        # the product is symmetric, so both directions carry the same weight
        assert _weight(g, w, nei, i) == want
    assert w.min() >= 1 and w.max() <= g.nodes


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_harness_host_parts_match_the_reference_rules(ig, orc, tmp_path):
    g = _graph(ig)
    egr = str(tmp_path / "wrap100k.egr")
    g.write_egr(egr)
    exe = tmp_path / "harness_host_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "harness"),
           os.path.join(ROOT, "tests", "harness_host_check.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    out = tmp_path / "out.bin"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0")
    run = subprocess.run([str(exe), egr, str(out)], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, (run.stdout + run.stderr)[-4000:]
    assert "Generating weights." in run.stdout
    raw = out.read_bytes()
    e = g.edges
    w = np.frombuffer(raw, np.int32, e, 0)
    sp = np.frombuffer(raw, np.int32, e, 4 * e)
    med = np.frombuffer(raw, np.float64, 3, 8 * e)
    np.testing.assert_array_equal(w, orc.synth_weights(g))
    np.testing.assert_array_equal(sp, np.repeat(np.arange(g.nodes, dtype=np.int32), np.diff(g.nindex)))
    np.testing.assert_array_equal(sp, g.sp())
    assert list(med) == [2.0, 2.5, 7.0]
