"""CPU tests of the host side: graph generators and CSR builder, .egr I/O in
the ECLgraph format, and the C ABI library (loads, exports every declared
symbol, fails cleanly without a GPU).  No GPU compute is issued here."""
import ctypes
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("indigo_hip.h", "indigo_graph.h"):
        text = open(os.path.join(REPO, "include", h)).read()
        names += re.findall(r"^\s*(?:const\s+)?[a-z_]+\*?\s+\*?(indigo_[a-z0-9_]+)\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(ig):
    lib = ig.load()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_constants_match_the_header(ig):
    """Every numeric INDIGO_* constant of include/indigo_hip.h (status codes,
    hints, query kinds, capacities) has the same value in the binding."""
    import re
    text = open(os.path.join(REPO, "include", "indigo_hip.h")).read()
    defs = re.findall(r"^#define INDIGO_(\w+)\s+\(?(-?\d+)\)?", text, re.M)
    assert len(defs) >= 20
    for name, value in defs:
        assert getattr(ig._lib, name) == int(value), name


def test_python_binding_covers_the_abi(ig):
    bound = set(ig._lib._SIGNATURES) | {"indigo_last_error"}
    assert set(declared_functions()) <= bound


def test_grid_generator_shape(ig):
    g = ig.grid(7, 11)
    assert g.nodes == 77 and g.edges == 2 * (7 * 10 + 6 * 11)
    assert g.check() == 0
    assert list(g.nlist[g.nindex[12]:g.nindex[13]]) == [1, 11, 13, 23]


def test_rmat_generator_properties(ig):
    g = ig.rmat(12, 16, seed=1, wmin=1, wmax=255)
    assert g.nodes == 4096 and g.check() == 0
    assert g.eweight.min() >= 1 and g.eweight.max() <= 255
    # skew: vertex 0 is the densest row of an unpermuted R-MAT
    assert g.degree[0] == g.degree.max()
    # deterministic in the seed, different across seeds
    assert np.array_equal(ig.rmat(12, 16, seed=1).nlist, g.nlist)
    assert not np.array_equal(ig.rmat(12, 16, seed=2).nlist[:1000], g.nlist[:1000])


def test_road_and_er_are_simple_and_symmetric(ig):
    for g in (ig.road(50, 60, seed=3, wmin=1, wmax=1000), ig.erdos_renyi(1000, 4000, seed=4, wmin=1, wmax=255)):
        assert g.check() == 0


def test_build_csr_dedups_and_drops_self_loops(ig):
    g = ig.from_edges(4, [0, 0, 1, 2, 2, 3], [1, 1, 0, 2, 3, 3])
    assert g.edges == 4
    assert list(g.nindex) == [0, 1, 2, 3, 4]
    assert list(g.nlist) == [1, 0, 3, 2]


def test_graph_check_detects_violations(ig):
    g = ig.grid(4, 4)
    bad = ig.HostGraph(g.nodes, g.edges, g.nindex, g.nlist[::-1].copy())
    assert bad.check() != 0


def test_egr_roundtrip_and_format(ig, tmp_path):
    g = ig.rmat(8, 8, seed=5, wmin=1, wmax=255)
    p = tmp_path / "g.egr"
    g.write_egr(str(p))
    raw = np.fromfile(p, dtype="<i4")
    # ECLgraph layout: nodes, edges, nindex[nodes+1], nlist[edges], eweight[edges]
    assert raw[0] == g.nodes and raw[1] == g.edges
    assert np.array_equal(raw[2:3 + g.nodes], g.nindex)
    assert np.array_equal(raw[3 + g.nodes:3 + g.nodes + g.edges], g.nlist)
    assert np.array_equal(raw[3 + g.nodes + g.edges:], g.eweight)
    h = ig.read_egr(str(p))
    assert np.array_equal(h.nlist, g.nlist) and np.array_equal(h.eweight, g.eweight)


def test_egr_without_weights_and_truncated(ig, tmp_path):
    g = ig.grid(3, 3)
    p = tmp_path / "nw.egr"
    g.write_egr(str(p))
    assert ig.read_egr(str(p)).eweight is None
    data = open(p, "rb").read()
    q = tmp_path / "trunc.egr"
    q.write_bytes(data[:-4])
    with pytest.raises(ig.IndigoError):
        ig.read_egr(str(q))


def test_golden_egr_readable_by_eclgraph_layout(ig):
    g = ig.read_egr(os.path.join(REPO, "tests", "golden", "grid32x32.egr"))
    assert g.nodes == 1024 and g.edges == 3968 and g.check() == 0


def test_context_fails_cleanly_without_gpu(ig):
    import subprocess
    import sys
    # in a child process, so a machine with a GPU still runs the rest of the suite normally
    code = ("import sys; sys.path.insert(0, %r); import indigo2 as ig, ctypes\n"
            "h = ctypes.c_void_p(); rc = ig.load().indigo_ctx_create(0, ctypes.byref(h))\n"
            "print(rc)") % os.path.join(REPO, "parallelization-artifact-sc23_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    rc = int(out.stdout.strip().splitlines()[-1])
    assert rc in (0, ig._lib.ERR_NO_DEVICE)


def test_batch_structs_match_the_c_layout(ig, tmp_path):
    """ctypes mirrors of indigo_graph_ref / indigo_query against the C compiler."""
    import ctypes
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "indigo_hip.h"\n'
                   'int main(void) { printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(indigo_query), '
                   'offsetof(indigo_query, src), offsetof(indigo_query, d_out), sizeof(indigo_graph_ref), '
                   'offsetof(indigo_query, gang), '
                   'offsetof(indigo_graph_ref, d_eweight)); '
                   'printf("%zu %zu %zu %zu %zu\\n", sizeof(indigo_request), offsetof(indigo_request, d_out), '
                   'offsetof(indigo_request, slot), sizeof(indigo_serve_config), sizeof(indigo_serve_stats)); '
                   'return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, f"-I{REPO}/include", str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    Q, G, R = ig._lib.QueryStruct, ig._lib.GraphRefStruct, ig._lib.RequestStruct
    assert got == [ctypes.sizeof(Q), Q.src.offset, Q.d_out.offset, ctypes.sizeof(G), Q.gang.offset, G.d_eweight.offset,
                   ctypes.sizeof(R), R.d_out.offset, R.slot.offset, ctypes.sizeof(ig._lib.ServeConfigStruct),
                   ctypes.sizeof(ig._lib.ServeStatsStruct)]


def test_autotune_classes_and_lookup(ig):
    """Graph classes from the suite's features, database fallbacks, and the
    gang scaling against the engine's own rule (host-only calls)."""
    from indigo2 import autotune as at

    g = ig.rmat(10, 8, seed=3, wmin=1, wmax=255)
    f = at.features(g, diameter=7)
    assert f["numv"] == g.nodes and f["nume"] == g.edges and f["dMax"] == int(g.degree.max())
    assert abs(f["dAvg"] - g.edges / g.nodes) < 1e-12
    assert at.graph_class(f) == "skewed-small"
    assert at.graph_class(dict(f, diameter=2047)) == "deep-small"
    er = ig.erdos_renyi(1 << 16, 4 << 16, seed=1)
    assert at.graph_class(at.features(er, diameter=12)) == "uniform-medium"

    db = at.Tuning({"single": {"skewed-large": {"bfs": {"hints": {"BLOCKS_PER_CU": 2}}}},
                    "batched": {"uniform-small": {"cc": {"gang_scale": 4.0}}}})
    h = db.hints("bfs", "skewed-medium")  # same shape, other size: falls back
    assert h[ig._lib.HINT_BLOCKS_PER_CU] == 2 and h[ig._lib.HINT_BFS_DIRECTION] == 0
    assert all(v == 0 for v in db.hints("bfs", "deep-large").values())
    # the engine's rule: one workgroup per 2^18 entries, rounded up to a power of two
    from indigo2.kernels import batch_gang
    assert batch_gang("cc", 1 << 20, 7 << 20) == 32
    assert batch_gang("cc", 1000, 8000) == 1
    assert batch_gang("bfs", 1 << 22, 1 << 22) == 16  # deep sparse BFS cap
    assert db.gang("cc", "uniform-small", 1 << 20, 7 << 20) == 128
    assert db.gang("cc", "skewed-small", 1 << 20, 7 << 20) == 0
    assert db.gang("bfs", "uniform-small", 1 << 20, 7 << 20) == 0
    assert at.scaled_gang("pr", 1 << 24, 1 << 28, 4.0) == ig._lib.MAX_GANG


def test_egr_mapping_matches_the_reader(ig, tmp_path):
    """indigo_egr_open gives the same arrays as indigo_read_egr, and follows
    the reference reader's rules for the optional weights
    (include/ECLgraph.h:70-76): no whole weight means none, a partial set is
    an error, trailing bytes after the weights are ignored."""
    import ctypes

    from indigo2 import _lib

    def mapped(path):
        m = _lib.EgrMapStruct()
        _lib.call("indigo_egr_open", str(path).encode(), ctypes.byref(m))
        try:
            n, e = m.nodes, m.edges
            out = (n, e, np.ctypeslib.as_array(m.nindex, shape=(n + 1,)).copy(),
                   np.ctypeslib.as_array(m.nlist, shape=(max(e, 1),))[:e].copy(),
                   np.ctypeslib.as_array(m.eweight, shape=(max(e, 1),))[:e].copy() if m.eweight else None)
        finally:
            _lib.load().indigo_egr_close(ctypes.byref(m))
        return out

    g = ig.rmat(9, 8, seed=4, wmin=1, wmax=99)
    p = tmp_path / "g.egr"
    g.write_egr(str(p))
    n, e, ni, nl, ew = mapped(p)
    h = ig.read_egr(str(p))
    assert (n, e) == (h.nodes, h.edges)
    assert np.array_equal(ni, h.nindex) and np.array_equal(nl, h.nlist) and np.array_equal(ew, h.eweight)
    raw = p.read_bytes()
    base = 4 * (2 + g.nodes + 1 + g.edges)
    cases = {"noweights": raw[:base], "stub": raw[:base + 3], "trailing": raw + b"\x01\x02"}
    for name, data in cases.items():
        q = tmp_path / f"{name}.egr"
        q.write_bytes(data)
        got = mapped(q)
        ref = ig.read_egr(str(q))
        assert (got[4] is None) == (ref.eweight is None), name
        assert name != "trailing" or np.array_equal(got[4], g.eweight)
    for name, data in {"partial": raw[:base + 8], "short": raw[:base - 4], "header": raw[:4]}.items():
        q = tmp_path / f"{name}.egr"
        q.write_bytes(data)
        with pytest.raises(ig.IndigoError):
            mapped(q)
        with pytest.raises(ig.IndigoError):
            ig.read_egr(str(q))


def test_egr_metadata_sidecar(ig, tmp_path):
    g = ig.road(20, 30, keep_p=0.9, diag_p=0.05, seed=11, wmin=1, wmax=1000)
    p = tmp_path / "road.egr"
    g.write_egr(str(p))
    h = ig.read_egr(str(p))
    assert h.meta["generator"] == "road" and h.meta["seed"] == 11 and h.meta["rows"] == 20
    assert h.meta["nodes"] == g.nodes and h.meta["edges"] == g.edges and h.meta["weights"] is True
    # regenerating from the sidecar gives the same graph
    m = h.meta
    again = ig.road(m["rows"], m["cols"], keep_p=m["keep_p"], diag_p=m["diag_p"], seed=m["seed"], wmin=m["wmin"],
                    wmax=m["wmax"])
    assert np.array_equal(again.nlist, h.nlist) and np.array_equal(again.eweight, h.eweight)
