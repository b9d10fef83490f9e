"""SNAP edge-list ingest (gconvert) and dataset statistics (gstats).

Parity: the .bin written from our ingest must equal, byte for byte, the one
the reference's own convertSnap writes (oracle/_ref, compiled from
src/graph/snap.cpp) on edge lists with comments, blank lines, tabs, CRLF,
weight columns, duplicates, self-loops, sparse raw IDs, a last line without a
newline, gzip, directed and undirected. gstats figures are checked against
numpy / the oracle's BFS, and both CLIs' CSV rows against the reference's
compiled gstats / gconvert programs.
"""
import csv
import gzip
import os
import subprocess

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "graph_analytics_applesilicon_amd",
                   "bin")
REF_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
ref_built = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref is not built (no reference checkout)")

TRICKY = ("# comment line\n% another comment\n\n"
          "  5 7\n7\t5\n5 5\n5 7 0.25 extra\n"        # duplicates both ways, a self-loop, weight columns
          "1000000 3\r\n3 1000000\n"                    # sparse raw IDs, CRLF
          "\t 42 8\n8 42\n0 9")                         # tabs, no trailing newline


def _write(path, text, gz=False):
    if gz:
        with gzip.open(path, "wt") as f:
            f.write(text)
    else:
        with open(path, "w") as f:
            f.write(text)
    return str(path)


def _ours(src, dst, directed):
    g, st = G.Graph.from_snap(src, directed)
    g.save(str(dst))
    return (st.lines, st.raw_max_id, st.n, st.m_before, st.m), g


def _random_edge_list(rng, lines=60000, ids=20000):
    u = rng.integers(0, ids, size=lines) * 3 + 11   # sparse raw IDs
    v = rng.integers(0, ids, size=lines) * 3 + 11
    hub = rng.random(lines) < 0.1
    v[hub] = 11                                      # a hub
    out = ["# random edge list"]
    for a, b in zip(u.tolist(), v.tolist()):
        out.append(f"{a} {b}")
    return "\n".join(out) + "\n"


@ref_built
@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("gz", [False, True])
def test_ingest_equals_reference_converter(tmp_path, directed, gz):
    rng = np.random.default_rng(7)
    for i, text in enumerate((TRICKY, _random_edge_list(rng), "", "# only comments\n")):
        src = _write(tmp_path / f"e{i}.txt{'.gz' if gz else ''}", text, gz)
        mine, theirs = tmp_path / f"m{i}.bin", tmp_path / f"r{i}.bin"
        st, g = _ours(src, mine, directed)
        assert st == O.ref_convert_snap(src, theirs, directed), (i, text[:40])
        assert open(mine, "rb").read() == open(theirs, "rb").read(), i
        assert g.validate() == ""


def test_ingest_tricky_file_by_hand(tmp_path):
    src = _write(tmp_path / "t.txt", TRICKY)
    g, st = G.Graph.from_snap(src, directed=False)
    # raw IDs 0 3 5 7 8 9 42 1000000 -> 0..7
    assert (st.lines, st.raw_max_id, st.n) == (9, 1000000, 8)
    assert st.m_before == 18 and st.m == g.m == 8
    adj = {v: g.col_idx[g.row_ptr[v]:g.row_ptr[v + 1]].tolist() for v in range(g.n)}
    assert adj == {0: [5], 1: [7], 2: [3], 3: [2], 4: [6], 5: [0], 6: [4], 7: [1]}
    d, _ = G.Graph.from_snap(src, directed=True)
    assert d.directed and d.m == 7  # 5->7, 7->5, 1e6->3, 3->1e6, 42->8, 8->42, 0->9


@ref_built
def test_ingest_errors_match_reference(tmp_path):
    for text, what in (("1 2\n3 x\n", "malformed line 2"), ("# c\n1 4294967295\n", "vertex ID too large on line 2"),
                       ("1\n", "malformed line 1")):
        src = _write(tmp_path / "bad.txt", text)
        with pytest.raises(ValueError) as mine:
            G.Graph.from_snap(src, False)
        with pytest.raises(ValueError) as theirs:
            O.ref_convert_snap(src, tmp_path / "bad.bin", False)
        assert what in str(mine.value) and what in str(theirs.value)
    with pytest.raises((OSError, ValueError, RuntimeError)):
        G.Graph.from_snap(str(tmp_path / "missing.txt"), False)


def _splitmix_below(seed, count, bound):
    """SplitMix64 below() (src/util/rng.hpp:8-26), the stream gstats draws its sources from."""
    mask = (1 << 64) - 1
    state = seed
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        yield (z * bound) >> 64


@pytest.mark.parametrize("name", ["rmat12_directed", "two_components", "grid100x100", "empty1000", "star100k"])
def test_stats_match_independent_computation(name):
    g = G.Graph.named(name)
    st = g.stats()
    s = g.symmetrize()
    deg, sdeg = g.degrees().astype(np.int64), s.degrees().astype(np.int64)
    assert (st.n, st.m, st.undirected_edges) == (g.n, g.m, s.m // 2)
    assert (st.min_deg, st.max_deg) == (deg.min(), deg.max()) and st.mean_deg == pytest.approx(g.m / g.n)
    indeg = np.bincount(g.col_idx, minlength=g.n) if g.directed else deg
    assert st.max_in_deg == indeg.max() and st.isolated == int((sdeg == 0).sum())
    # components by BFS from every unlabeled vertex
    label = np.full(g.n, -1)
    sizes = []
    for v in range(g.n):
        if label[v] < 0:
            d = O.bfs_serial(s.row_ptr, s.col_idx, v)
            label[d != O.INF] = len(sizes)
            sizes.append(int((d != O.INF).sum()))
    assert (st.num_cc, st.largest_cc) == (len(sizes), max(sizes))
    # deepest level from the first 10 sources of nonzero degree in the stream
    diam, k, draws = 0, 0, _splitmix_below(1, 10**6, g.n)
    while st.isolated < g.n and k < 10:
        src = next(draws)
        if sdeg[src] == 0:
            continue
        d = O.bfs_serial(s.row_ptr, s.col_idx, int(src))
        diam = max(diam, int(d[d != O.INF].max()))
        k += 1
    assert st.approx_diameter == diam
    for kind, dv in ((0, deg), (1, indeg), (2, sdeg)):
        h = g.degree_histogram(kind)
        want = np.bincount([0 if x == 0 else int(x).bit_length() for x in dv], minlength=1)
        assert h.tolist() == want.tolist(), kind


@ref_built
@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "gstats")), reason="tools not built")
def test_cli_rows_equal_reference_tools(tmp_path):
    rng = np.random.default_rng(3)
    src = _write(tmp_path / "lj_small.txt.gz", _random_edge_list(rng, 30000, 8000), gz=True)
    for directed in (False, True):
        flag = "--directed" if directed else "--undirected"
        rows = {}
        for who, tool in (("mine", os.path.join(BIN, "gconvert")), ("ref", os.path.join(REF_BIN, "gconvert"))):
            out = tmp_path / who / "lj_small.bin"
            out.parent.mkdir(exist_ok=True)
            conv_csv, stats_csv, hist_csv = (tmp_path / who / f"{x}.csv" for x in ("conv", "stats", "hist"))
            r = subprocess.run([tool, src, str(out), flag, "--csv", str(conv_csv)], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr
            gst = os.path.join(BIN if who == "mine" else REF_BIN, "gstats")
            r = subprocess.run([gst, str(out), "--csv", str(stats_csv), "--hist", str(hist_csv)], capture_output=True,
                               text=True)
            assert r.returncode == 0, r.stderr
            conv = list(csv.DictReader(open(conv_csv)))
            rows[who] = ([{k: v for k, v in c.items() if k not in ("seconds", "peak_rss_mb")} for c in conv],
                         list(csv.DictReader(open(stats_csv))), list(csv.DictReader(open(hist_csv))),
                         open(out, "rb").read())
        assert rows["mine"] == rows["ref"], directed
    bad = subprocess.run([os.path.join(BIN, "gconvert"), src, str(tmp_path / "x.bin")], capture_output=True)
    assert bad.returncode == 2
    missing = subprocess.run([os.path.join(BIN, "gstats"), str(tmp_path / "none.bin")], capture_output=True)
    assert missing.returncode == 1
