"""External cross-check (NetworkX) and the GAP .sg export.

validate_networkx checks BFS depths and PageRank of the host library (and,
on the GPU, of gpu_do and gpu_pull_adaptive) against NetworkX on the suite.
gexport's .sg files must equal the ones the reference's own gexport writes
(oracle/_ref, compiled from src/graph/export_sg.cpp), and parse back to the
graph under GAP's layout.
"""
import os
import subprocess

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from graph_analytics_applesilicon_amd import validate_networkx as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEXPORT = os.path.join(ROOT, "graph_analytics_applesilicon_amd", "bin", "gexport")
REF_GEXPORT = os.path.join(ROOT, "oracle", "_ref", "gexport")
SPECS = ["synth:" + s for s in V.SUITE]


def test_host_library_matches_networkx():
    rows = V.validate(SPECS, k=3, gpu=False)
    bad = [r for r in rows if r["status"] != "OK"]
    assert len(rows) > 20 and not bad, bad


@pytest.mark.gpu
def test_gpu_matches_networkx():
    rows = V.validate(SPECS, k=3, gpu=True)
    bad = [r for r in rows if r["status"] != "OK"]
    assert {r["variant"] for r in rows} >= {"gpu_do", "gpu_pull_adaptive"} and not bad, bad


def _read_sg(path):
    b = open(path, "rb").read()
    directed = b[0] != 0
    m, n = np.frombuffer(b[1:17], dtype="<i8")
    off = 17
    rp = np.frombuffer(b[off:off + 8 * (n + 1)], dtype="<i8")
    off += 8 * (n + 1)
    ci = np.frombuffer(b[off:off + 4 * m], dtype="<i4")
    off += 4 * m
    rest = b[off:]
    return directed, int(n), int(m), rp, ci, rest


@pytest.mark.skipif(not os.path.exists(GEXPORT), reason="gexport is not built")
def test_gexport_layout_and_reference(tmp_path):
    for name in ("rmat12_directed", "two_components", "empty1000"):
        g = G.Graph.named(name)
        src = str(tmp_path / f"{name}.bin")
        g.save(src)
        for sym in (False, True):
            out = str(tmp_path / f"{name}{'_sym' if sym else ''}.sg")
            r = subprocess.run([GEXPORT, src, out] + (["--sym"] if sym else []), capture_output=True, text=True)
            assert r.returncode == 0, r.stderr
            directed, n, m, rp, ci, rest = _read_sg(out)
            want = g.symmetrize() if sym else g
            assert directed == (g.directed and not sym) and n == want.n and m == want.m
            assert np.array_equal(rp, want.row_ptr) and np.array_equal(ci.astype(np.uint32), want.col_idx)
            if directed:
                t = g.transpose()
                assert rest == t.row_ptr.astype("<i8").tobytes() + t.col_idx.astype("<i4").tobytes()
            else:
                assert rest == b""
            if os.path.exists(REF_GEXPORT):
                ref = str(tmp_path / "ref.sg")
                r = subprocess.run([REF_GEXPORT, src, ref] + (["--sym"] if sym else []), capture_output=True)
                assert r.returncode == 0 and open(ref, "rb").read() == open(out, "rb").read(), (name, sym)
    assert subprocess.run([GEXPORT, "only-one-arg"], capture_output=True).returncode == 2
    assert subprocess.run([GEXPORT, str(tmp_path / "none.bin"), str(tmp_path / "x.sg")],
                          capture_output=True).returncode == 1
