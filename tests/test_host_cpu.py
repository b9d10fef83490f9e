"""Host graph library (libggpu_host.so) on the CPU.

Mirrors the reference's host-side tests (tests/test_main.cpp:114-131
generators, :133-160 file round trip, transforms, cpu_serial / cpu_omp /
cpu_do and PageRank variants) and pins the generators, transforms and CPU
variants against the golden vectors of the reference's compiled code
(tests/golden/ref_vectors.json) and, where it is built, against that code
directly.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O
from tests.conftest import SUITE, sample_sources

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))
ref_built = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref (reference CPU build) not built")


def _sha(a):
    return hashlib.sha1(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


def test_generators_valid_and_shaped():
    """tests/test_main.cpp:114-131."""
    for name in SUITE:
        assert G.Graph.named(name).validate() == ""
    p = G.Graph.named("path1000")
    assert p.n == 1000 and p.m == 2 * 999
    st = G.Graph.named("star100k")
    d = st.degrees()
    assert st.n == 100001 and d[0] == 100000 and d[5] == 1
    assert G.Graph.named("complete200").m == 200 * 199
    assert G.Graph.named("empty1000").m == 0
    assert G.Graph.named("grid100x100").m == 2 * 2 * 100 * 99
    r = G.Graph.rmat(12, 16, directed=True, seed=43)
    assert r.directed and r.n == 4096 and r.m > 30000


@pytest.mark.parametrize("name", SUITE)
def test_suite_graphs_equal_reference_golden(name):
    g = G.Graph.named(name)
    want = GOLDEN["suite"][name]
    assert (g.n, g.m, g.flags) == (want["n"], want["m"], want["flags"])
    assert _sha(g.row_ptr) == want["row_ptr_sha1"] and _sha(g.col_idx) == want["col_idx_sha1"]
    s = g.symmetrize()
    assert s.m == want["sym_m"] and _sha(s.col_idx) == want["sym_col_idx_sha1"]


def test_rmat_and_permutation_equal_reference_golden():
    for w in GOLDEN["rmat"]:
        g = G.Graph.rmat(w["scale"], w["edgefactor"], directed=w["directed"], seed=w["seed"])
        assert g.m == w["m"] and _sha(g.row_ptr) == w["row_ptr_sha1"] and _sha(g.col_idx) == w["col_idx_sha1"]
    for w in GOLDEN["random_perm"]:
        p = G.random_perm(w["n"], w["seed"])
        assert _sha(p) == w["sha1"] and p[:8].tolist() == w["head"]


def test_workload_graph_shapes():
    """The benchmark's scale-16 graph: RMAT(16, 16, .57/.19/.19), symmetric, ids permuted (seed 1)."""
    g = G.Graph.named("rmat16_c1")
    base = G.Graph.rmat(16, 16, seed=1)
    assert g.n == 65536 and g.m == base.m and g.validate() == ""
    assert g.equal(base.permute(G.random_perm(g.n, 1)))
    assert np.array_equal(np.sort(g.degrees()), np.sort(base.degrees()))


def test_save_load_roundtrip(tmp_path):
    for name in SUITE:
        g = G.Graph.named(name)
        f = str(tmp_path / f"{name}.bin")
        g.save(f)
        h = G.Graph.load(f)
        assert h.equal(g) and h.flags == g.flags
        with open(f, "rb") as fh:
            assert fh.read(8) == b"GGPUCSR1"
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTACSR!" + bytes(64))
    with pytest.raises((OSError, ValueError)):
        G.Graph.load(str(bad))


def test_from_edges_sorts_dedups_and_drops_loops():
    g = G.Graph.from_edges(4, [(0, 1), (1, 0), (0, 1), (2, 2), (3, 0)], directed=True)
    assert g.row_ptr.tolist() == [0, 1, 2, 2, 3] and g.col_idx.tolist() == [1, 0, 0]
    u = G.Graph.from_edges(4, [(0, 1), (1, 0), (0, 1), (2, 2), (3, 0)], directed=False)
    assert u.row_ptr.tolist() == [0, 2, 3, 3, 4] and u.col_idx.tolist() == [1, 3, 0, 0]


def test_transforms_against_numpy():
    g = G.Graph.named("rmat12_directed")
    rp, ci = g.row_ptr.astype(np.int64), g.col_idx.astype(np.int64)
    src = np.repeat(np.arange(g.n), np.diff(rp))
    t = g.transpose()
    tsrc = np.repeat(np.arange(g.n), np.diff(t.row_ptr.astype(np.int64)))
    assert sorted(zip(ci.tolist(), src.tolist())) == list(zip(tsrc.tolist(), t.col_idx.tolist()))
    s = g.symmetrize()
    pairs = set(zip(src.tolist(), ci.tolist())) | set(zip(ci.tolist(), src.tolist()))
    ssrc = np.repeat(np.arange(g.n), np.diff(s.row_ptr.astype(np.int64)))
    assert set(zip(ssrc.tolist(), s.col_idx.tolist())) == pairs and s.m == len(pairs)
    perm = G.random_perm(g.n, 9)
    p = g.permute(perm)
    psrc = np.repeat(np.arange(g.n), np.diff(p.row_ptr.astype(np.int64)))
    assert set(zip(psrc.tolist(), p.col_idx.tolist())) == set(zip(perm[src].tolist(), perm[ci].tolist()))


@ref_built
def test_transforms_equal_compiled_reference():
    for name in ("rmat12_directed", "two_components", "grid100x100"):
        g = G.Graph.named(name)
        rp, ci, fl = g.row_ptr, g.col_idx, g.flags
        for mine, fn in ((g.transpose(), O.ref_transpose), (g.symmetrize(), O.ref_symmetrize)):
            r2, c2, _ = fn(rp, ci, fl)
            assert np.array_equal(mine.row_ptr, r2) and np.array_equal(mine.col_idx, c2)
        perm = O.ref_random_perm(g.n, 3)
        r2, c2, _ = O.ref_permute(rp, ci, fl, perm)
        q = g.permute(perm)
        assert np.array_equal(q.row_ptr, r2) and np.array_equal(q.col_idx, c2)


@ref_built
def test_orderings_equal_compiled_reference():
    """reorder::compute and locality of the compiled reference, bit for bit,
    on the suite (isolated vertices, several components, a hub, a grid) and
    an RMAT graph; orderings with equal degrees everywhere exercise the id
    tie-breaks."""
    for name in SUITE + ["rmat16_c1"]:
        s = G.Graph.named(name).symmetrize()
        for ordering in ("original", "random", "degree", "rcm"):
            for seed in (1, 7):
                mine = s.order_perm(ordering, seed)
                want = O.ref_order_perm(s.row_ptr, s.col_idx, s.flags, ordering, seed)
                assert np.array_equal(mine, want), (name, ordering, seed)
                if ordering in ("degree", "rcm"):
                    break  # no seed
            q = s.permute(mine)
            assert q.locality() == O.ref_locality(q.row_ptr, q.col_idx, q.flags), (name, ordering)


def test_orderings_golden_and_properties():
    for w in GOLDEN["orderings"]:
        s = G.Graph.named(w["graph"]).symmetrize()
        p = s.order_perm(w["ordering"], w["seed"])
        assert _sha(p) == w["sha1"] and p[:8].tolist() == w["head"], w
        assert s.permute(p).locality() == (w["mean_gap"], w["same_block64"]), w
    # closed forms: a path numbered from one end is its own RCM order up to
    # reversal; degree order puts the star's hub first
    path = G.Graph.named("path1000")
    p = path.order_perm("rcm")
    assert sorted(p.tolist()) == list(range(1000))
    assert np.abs(np.diff(p.astype(np.int64))).max() == 1  # neighbours stay neighbours
    star = G.Graph.named("star100k")
    assert star.order_perm("degree")[int(np.argmax(star.degrees()))] == 0
    with pytest.raises(ValueError):
        star.order_perm("spectral")


def test_largest_component_and_sources():
    g = G.Graph.named("two_components")
    lcc, old = g.largest_component()
    assert lcc.n == 300 and old.tolist() == list(range(300))
    r = G.Graph.named("rmat12")
    lcc, old = r.largest_component()
    s = r.pick_sources(64, 1)
    assert len(s) == 64 and len(set(s.tolist())) == 64
    assert np.all(r.degrees()[s] > 0)
    d = O.bfs_serial(r.row_ptr, r.col_idx, int(s[0]))
    reached = set(np.nonzero(d != O.INF)[0].tolist())
    assert set(s.tolist()) <= reached and len(reached) == lcc.n
    assert np.array_equal(r.pick_sources(64, 1), s) and not np.array_equal(r.pick_sources(64, 2), s)


@pytest.mark.parametrize("name", SUITE)
def test_cpu_bfs_variants_match_oracle_and_reference_trace(name):
    """cpu_serial / cpu_omp (all schedules) / cpu_do; cpu_do's trace equals the reference's (golden)."""
    g = G.Graph.named(name).symmetrize()
    rp, ci = g.row_ptr, g.col_idx
    for want in GOLDEN["suite"][name]["bfs"]:
        s = want["source"]
        ref = O.bfs_serial(rp, ci, s)
        d, _ = g.cpu_bfs(s, "cpu_serial")
        assert np.array_equal(d, ref)
        for sched in (0, 1, 2):
            for threads in (1, 4):
                d, _ = g.cpu_bfs(s, "cpu_omp", threads=threads, schedule=sched)
                assert np.array_equal(d, ref)
        d, tr = g.cpu_bfs(s, "cpu_do", threads=4, trace=True)
        assert np.array_equal(d, ref)
        assert [[L.direction.decode(), L.frontier, L.edges] for L in tr] == want["do_levels"]


def test_cpu_pagerank_matches_oracle():
    for name in ("rmat12_directed", "path1000", "star100k"):
        g = G.Graph.named(name)
        t = g.transpose()
        ref, it, _ = O.pagerank(t.row_ptr, t.col_idx, g.degrees(), iters=20)
        for parallel in (False, True):
            r, it2, _ = t.cpu_pagerank(g.degrees(), iters=20, parallel=parallel, threads=4)
            assert it2 == it and np.abs(r - ref).max() < 1e-12
