"""Host-side check of the persistent-grid arithmetic shared by every
persistent kernel (csrc/grid_plan.hpp): the per-CU residency, the grid under
the hints, the caller's bound and a learned cap, and the lowered cap after a
grid that did not assemble.  tests/grid_plan_check.cpp holds the cases; it is
compiled with g++ (no GPU) under ASan/UBSan."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parallelization-artifact-sc23_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_grid_plan_arithmetic(tmp_path):
    exe = tmp_path / "grid_plan_check"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-I" + CSRC,
                           os.path.join(ROOT, "tests", "grid_plan_check.cpp"), "-o", str(exe)])
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
