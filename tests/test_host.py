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
    # in a child process, so a GPU box still runs the rest of the suite normally
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
