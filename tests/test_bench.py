"""The bench script's host-side contract, without a GPU: the one-GPU rule and
the traffic record computed from rocprofv3 --pmc CSVs."""
import csv
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _pmc_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for i, (kernel, counter, value) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": kernel, "Counter_Name": counter, "Counter_Value": value})


def test_traffic_record_scales_by_the_copy(tmp_path):
    import bench

    tc = "void indigo::(anonymous namespace)::tc_count_group<512, 8192>(int const*)"
    pr = "indigo::(anonymous namespace)::pr_kernel(indigo::PrArgs)"
    copy = "__amd_rocclr_copyBuffer"
    kib = float(1 << 20)  # a 1 GiB copy in KiB
    # the copy reads as half its size in FETCH and its size in WRITE: factors 2 and 1
    _pmc_csv(tmp_path / "cf.csv", [(copy, "FETCH_SIZE", kib / 2)] * 5 + [(copy, "FETCH_SIZE", 10.0)])
    _pmc_csv(tmp_path / "cw.csv", [(copy, "WRITE_SIZE", kib)] * 5)
    _pmc_csv(tmp_path / "kf.csv", [(tc, "FETCH_SIZE", 1000.0), (tc, "FETCH_SIZE", 3000.0), (pr, "FETCH_SIZE", 500.0),
                                   ("other_kernel", "FETCH_SIZE", 9e9)])
    _pmc_csv(tmp_path / "kw.csv", [(tc, "WRITE_SIZE", 10.0), (tc, "WRITE_SIZE", 30.0), (pr, "WRITE_SIZE", 50.0)])
    out = tmp_path / "traffic.json"
    rc = bench.record_traffic([str(tmp_path / n) for n in ("kf.csv", "kw.csv", "cf.csv", "cw.csv")],
                              "rmat22 ef16", str(out))
    assert rc == 0
    rec = json.loads(out.read_text())
    assert rec["copy_scale"] == {"fetch": 2.0, "write": 1.0}
    k = rec["kernels"]["tc_count_group<512, 8192>"]
    assert k["algorithm"] == "tc" and k["dispatches"] == [2, 2]
    assert k["read_bytes"] == pytest.approx(1024 * 2000.0 * 2.0)
    assert k["write_bytes"] == pytest.approx(1024 * 20.0)
    assert rec["kernels"]["pr_kernel"]["bytes_per_launch"] == pytest.approx(1024 * (500.0 * 2 + 50.0))
    got = bench.read_traffic(str(out), "rmat22 ef16", "pr_kernel")
    assert got["bytes_per_launch"] == pytest.approx(1024 * 1050.0)
    assert bench.read_traffic(str(out), "rmat20 ef16", "pr_kernel") is None


def test_traffic_record_needs_the_copy(tmp_path):
    import bench

    for n in ("kf.csv", "kw.csv", "cf.csv", "cw.csv"):
        _pmc_csv(tmp_path / n, [("x", "FETCH_SIZE", 1.0)])
    assert bench.record_traffic([str(tmp_path / n) for n in ("kf.csv", "kw.csv", "cf.csv", "cw.csv")], "w",
                                str(tmp_path / "t.json")) == 1


@pytest.mark.skipif(not os.path.exists(os.path.join(REPO, "oracle", "_ref", "libref_tc.so")),
                    reason="oracle/_ref not built (needs the reference sources)")
def test_cpu_baseline_runs_in_a_child_process():
    import bench

    rec = bench.cpu_baseline_isolated(12)
    assert rec is not None and rec["kind"] == "reference" and rec["value"] > 0
    assert set(rec["variants"]) == set(bench.ALGS)
    assert all(v in ("C++ threads", "OpenMP") for v in rec["variants"].values())


def _comb3(n):
    return n * (n - 1) * (n - 2) // 6


def test_tc_count_model_on_cliques():
    """The TC byte model against closed forms.  On a k-clique (all degrees
    equal, so the orientation is by id and |N+(i)| = k - 1 - i) every probe
    hits, so the probes of all passes are C(k, 3); a task has no in-list, and
    on a graph this small (under 2^24 CSR entries) it goes to the workgroup
    pass exactly when |N+| > 90 (wave tasks need |N+|^2 <= 2^13), so on a
    300-clique that pass owns the tasks with |N+| in 91..299, whose probes
    sum to C(300, 3) - C(91, 3).  Every vertex of a graph this small is a
    hub: 2 bytes per probe and per table entry."""
    sys.path.insert(0, os.path.join(REPO, "parallelization-artifact-sc23_amd"))
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import bench
    import indigo2 as ig
    from graphs import cliques

    sizes = [3, 5, 40, 300]
    m = bench.tc_count_model(cliques(ig, sizes))
    assert m["probes_all_passes"] == sum(_comb3(k) for k in sizes)
    mids = list(range(91, 300))  # |N+| of the workgroup-pass tasks
    assert m["tasks"] == m["hub_tasks"] == len(mids)
    assert m["probes"] == _comb3(300) - _comb3(91)
    assert m["items"] == sum(mids) and m["parts"] == len(mids) and m["table_entries"] == sum(mids)
    assert m["bytes"] == 2 * m["probes"] + 12 * m["items"] + 2 * m["table_entries"] + 16 * m["parts"]
