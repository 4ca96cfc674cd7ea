"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

Compiles the host-only CSR builders, generators and .egr readers
(parallelization-artifact-sc23_amd/csrc/graphgen.cpp) with g++ and the
sanitizers, together with tests/asan_graph_host.cpp, and runs the cases it
holds: generated graphs pass the invariant check, damaged .egr files are
rejected by both readers without leaks or out-of-bounds reads, and the
reader's rules for weights follow include/ECLgraph.h:58-76 of the reference
(no weights, all weights, or an error for a partial weight array).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallelization-artifact-sc23_amd")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_graph_host_code_is_sanitizer_clean(tmp_path):
    exe = tmp_path / "asan_graph_host"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "csrc"),
           "-pthread", os.path.join(ROOT, "tests", "asan_graph_host.cpp"), os.path.join(PKG, "csrc", "graphgen.cpp"),
           "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    assert "0 failure(s)" in run.stdout
    assert "FAIL" not in run.stdout
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
