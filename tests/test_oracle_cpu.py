"""The oracle pinned before it is trusted (CPU only).

Every function of oracle/oracle.c is checked against something it does not
restate: the reference's own known answers (tests/test_main.cpp:268-279 BFS,
:630-652 PageRank), the golden vectors written by tests/golden/make_golden.py
from the reference's compiled CPU code, the compiled reference itself when
oracle/_ref is built, closed forms, invariants, independent libraries
(scipy's shortest paths, networkx PageRank) and, for the MLP, finite
differences. A plausible slip (off-by-one depth, dropped dangling term,
transposed weight) fails at least one of them.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O
from tests.conftest import SUITE, sample_sources

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))


def _sha(a, dtype="<u4"):
    return hashlib.sha1(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def _sym(name):
    return G.Graph.named(name).symmetrize()


# ------------------------------------------------------------------- BFS
def test_bfs_known_answers():
    """tests/test_main.cpp:268-279."""
    g = G.Graph.named("path1000")
    d = O.bfs_serial(g.row_ptr, g.col_idx, 0)
    assert np.array_equal(d, np.arange(1000))
    g = G.Graph.named("star100k")
    d = O.bfs_serial(g.row_ptr, g.col_idx, 7)
    assert d[7] == 0 and d[0] == 1 and d[8] == 2
    g = G.Graph.named("two_components")
    d = O.bfs_serial(g.row_ptr, g.col_idx, 0)
    assert d[299] == 299 and d[300] == O.INF and d[399] == O.INF
    g = G.Graph.named("grid100x100")
    d = O.bfs_serial(g.row_ptr, g.col_idx, 0)
    assert d[99 * 100 + 99] == 198
    r, c = np.divmod(np.arange(10000), 100)
    assert np.array_equal(d, r + c)  # Manhattan distance on the full grid
    g = G.Graph.named("empty1000")
    d = O.bfs_serial(g.row_ptr, g.col_idx, 3)
    assert d[3] == 0 and d[4] == O.INF


@pytest.mark.parametrize("name", SUITE)
def test_bfs_against_golden_and_scipy(name):
    g = _sym(name)
    rp, ci = g.row_ptr, g.col_idx
    gold = GOLDEN["suite"][name]
    assert g.m == gold["sym_m"] and _sha(ci) == gold["sym_col_idx_sha1"]
    A = sp.csr_matrix((np.ones(len(ci)), ci.astype(np.int64), rp.astype(np.int64)), shape=(g.n, g.n))
    for s, want in zip(sample_sources(g.n), gold["bfs"]):
        assert want["source"] == s
        d = O.bfs_serial(rp, ci, s)
        assert _sha(d) == want["dist_sha1"]
        ref = csg.shortest_path(A, unweighted=True, indices=s)
        ref = np.where(np.isinf(ref), O.INF, ref).astype(np.uint32)
        assert np.array_equal(d, ref)


@pytest.mark.parametrize("name", SUITE)
def test_do_levels_against_golden(name):
    """The oracle's direction-optimizing trace equals the reference cpu_do's (bfs_cpu.cpp:227-283)."""
    g = _sym(name)
    for want in GOLDEN["suite"][name]["bfs"]:
        d, levels = O.bfs_do_levels(g.row_ptr, g.col_idx, want["source"])
        assert _sha(d) == want["dist_sha1"]
        assert [list(L) for L in levels] == want["do_levels"]


def test_bfs_invariants_random_graphs():
    """dist[src] = 0, |d(u) - d(v)| <= 1 on every edge, every reached v has a
    parent one level up, the reached set is the component; brute force
    (scipy) on small random graphs."""
    rng = np.random.default_rng(4)
    for trial in range(30):
        n = int(rng.integers(1, 60))
        k = int(rng.integers(0, 3 * n))
        e = rng.integers(0, n, size=(k, 2))
        g = G.Graph.from_edges(n, e, directed=False)
        rp, ci = g.row_ptr, g.col_idx
        s = int(rng.integers(0, n))
        d = O.bfs_serial(rp, ci, s).astype(np.int64)
        assert d[s] == 0
        src = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
        both = (d[src] != O.INF) & (d[ci] != O.INF)
        assert np.all(np.abs(d[src][both] - d[ci][both]) <= 1)
        assert np.all((d[src] == O.INF) == (d[ci] == O.INF))
        for v in range(n):
            if v != s and d[v] != O.INF:
                assert np.any(d[ci[rp[v]:rp[v + 1]]] == d[v] - 1)
        A = sp.csr_matrix((np.ones(len(ci)), ci.astype(np.int64), rp.astype(np.int64)), shape=(n, n))
        _, lab = csg.connected_components(A, directed=False)
        assert np.array_equal(d != O.INF, lab == lab[s])


# -------------------------------------------------------------- PageRank
def _pr(g, **kw):
    t = g.transpose()
    return O.pagerank(t.row_ptr, t.col_idx, g.degrees(), **kw)


def test_pagerank_known_answers():
    """tests/test_main.cpp:630-652 (tol 1e-14, up to 1000 iterations)."""
    kw = dict(damping=0.85, iters=1000, tol=1e-14)
    r, _, _ = _pr(G.Graph.from_edges(2, [(0, 1)], directed=True), **kw)
    assert abs(r[0] - 0.5 / 1.425) < 1e-12 and abs(r[1] - (1 - 0.5 / 1.425)) < 1e-12
    r, _, _ = _pr(G.Graph.from_edges(3, [(0, 1), (1, 2), (2, 0)], directed=True), **kw)
    assert np.all(np.abs(r - 1 / 3) < 1e-12)
    r, _, _ = _pr(G.Graph.named("complete200"), **kw)
    assert np.all(np.abs(r - 1 / 200) < 1e-12)
    r, _, _ = _pr(G.Graph.named("empty1000"), **kw)
    assert np.all(np.abs(r - 1 / 1000) < 1e-12)


@pytest.mark.parametrize("name", SUITE)
def test_pagerank_against_golden(name):
    g = G.Graph.named(name)
    want = GOLDEN["suite"][name]["pagerank20"]
    r, it, delta = _pr(g, damping=0.85, iters=20)
    assert it == want["iters"]
    assert abs(delta - want["last_delta"]) <= 1e-12 * max(1.0, want["last_delta"])
    assert abs(r.sum() - want["sum"]) < 1e-12
    for i, x in want["at"].items():
        assert abs(r[int(i)] - x) <= 1e-14 + 1e-12 * x


def test_pagerank_invariants_and_networkx():
    """Sum 1 with dangling redistribution; the contraction bound
    ||r_k - r*||_1 <= d/(1-d) ||r_k - r_{k-1}||_1; networkx at tight tolerance."""
    import networkx as nx
    g = G.Graph.named("rmat12_directed")
    r20, _, delta = _pr(g, iters=20)
    rstar, _, _ = _pr(g, iters=2000, tol=1e-15)
    assert abs(r20.sum() - 1) < 1e-12
    assert np.abs(r20 - rstar).sum() <= 0.85 / 0.15 * delta * (1 + 1e-9)
    rp, ci = g.row_ptr, g.col_idx
    D = nx.DiGraph()
    D.add_nodes_from(range(g.n))
    D.add_edges_from(zip(np.repeat(np.arange(g.n), np.diff(rp.astype(np.int64))).tolist(), ci.tolist()))
    ref = nx.pagerank(D, alpha=0.85, tol=1e-13 / g.n, max_iter=10000)
    nxr = np.array([ref[i] for i in range(g.n)])
    assert np.abs(nxr - rstar).sum() < 1e-9


def test_ppr_closed_forms():
    """An isolated seed keeps all mass (r = e_s); on a directed n-cycle the
    mass k steps after the seed is (1-d) d^k / (1 - d^n)."""
    g = G.Graph.named("empty1000")
    t = g.transpose()
    r, _, _ = O.ppr(t.row_ptr, t.col_idx, g.degrees(), seed=17, iters=200, tol=1e-15)
    want = np.zeros(1000)
    want[17] = 1
    assert np.abs(r - want).max() < 1e-12
    n, s, d = 7, 3, 0.85
    g = G.Graph.from_edges(n, [(i, (i + 1) % n) for i in range(n)], directed=True)
    t = g.transpose()
    r, _, _ = O.ppr(t.row_ptr, t.col_idx, g.degrees(), seed=s, damping=d, iters=5000, tol=1e-16)
    k = (np.arange(n) - s) % n
    assert np.abs(r - (1 - d) * d ** k / (1 - d ** n)).max() < 1e-12


def test_ppr_matches_networkx():
    """networkx.pagerank with personalization {s: 1} sends restart and dangling
    mass to s (its default dangling vector is the personalization): the FP64
    PPR restatement must agree on a directed RMAT graph with dangling
    vertices, for several seeds, including a dangling one."""
    import networkx as nx
    g = G.Graph.named("rmat12_directed")
    t = g.transpose()
    od = g.degrees()
    D = nx.DiGraph()
    D.add_nodes_from(range(g.n))
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr.astype(np.int64)))
    D.add_edges_from(zip(src.tolist(), g.col_idx.tolist()))
    dangling = np.flatnonzero(od == 0)
    assert len(dangling) > 0
    for s in (0, int(np.argmax(od)), int(dangling[0]), 1234):
        r, _, _ = O.ppr(t.row_ptr, t.col_idx, od, seed=s, iters=10000, tol=1e-14)
        ref = nx.pagerank(D, alpha=0.85, personalization={s: 1.0}, tol=1e-14 / g.n, max_iter=10000)
        want = np.array([ref[v] for v in range(g.n)])
        assert np.abs(r - want).sum() < 1e-9, s


def test_check_ranks_definition():
    """l1, sum and max relative error over the reference's top 1000 (pagerank_cpu.hpp:37-60)."""
    rng = np.random.default_rng(0)
    for n in (5, 999, 1000, 5000):
        ref = rng.permutation(n) + 1.0
        ref = ref / ref.sum()
        got = (ref * (1 + rng.normal(0, 1e-4, n))).astype(np.float32)
        l1, mx, s = O.check_ranks(ref, got)
        g64 = got.astype(np.float64)
        assert abs(l1 - np.abs(ref - g64).sum()) < 1e-15
        assert abs(s - g64.sum()) < 1e-12
        top = np.argsort(-ref)[: min(1000, n)]
        assert abs(mx - (np.abs(ref - g64)[top] / ref[top]).max()) < 1e-15
    assert O.ranks_ok(9e-5, 9e-4, 1 + 9e-5) and not O.ranks_ok(1e-4, 0, 1) and not O.ranks_ok(0, 1e-3, 1)


def _golden_inputs():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(__file__), "golden",
                                                                               "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


def _tie_extremes(ref, got, k=1000):
    """Smallest and largest top-k relative error over every choice of the equal
    reference values that straddle rank k."""
    order = np.argsort(-ref, kind="stable")
    cut = ref[order[k - 1]]
    above = order[ref[order] > cut]
    tied = np.flatnonzero(ref == cut)
    need = k - len(above)
    err = np.abs(ref - got.astype(np.float64)) / ref
    base = err[above].max() if len(above) else 0.0
    e = np.sort(err[tied])
    return max(base, e[need - 1]), max(base, e[-1])


def test_check_ranks_equals_the_reference():
    """or_check_ranks against the reference's compiled checkRanks (pagerank_cpu.hpp:37-60),
    through golden vectors it wrote: L1 and sum bit for bit, the top-1000 error bit for bit
    where no tie straddles rank 1000, and within the extremes of the tie choice where one does
    (std::partial_sort leaves it unspecified; the oracle takes ids in order)."""
    mg = _golden_inputs()
    want = {x["name"]: x for x in GOLDEN["check_ranks"]}
    for name, ref, got in mg.acceptance_inputs():
        l1, mx, s = O.check_ranks(ref, got)
        w = want[name]
        assert float.hex(l1) == w["l1"] and float.hex(s) == w["sum"], name
        if name.startswith("ties"):
            lo, hi = _tie_extremes(ref, got)
            assert lo < hi and lo <= mx <= hi and lo <= float.fromhex(w["max_rel"]) <= hi, (name, lo, mx, hi)
        else:
            assert float.hex(mx) == w["max_rel"], name
        if O.ref_lib() is not None:
            assert O.ref_check_ranks(ref, got)[0] == l1


def test_acceptance_rules_equal_the_reference():
    """ranks_ok and gs_no_worse against the reference's PrCheck::ok and gsNoWorseThanJacobi
    (pagerank_cpu.hpp:31, :80-84) on and beside every threshold (golden booleans)."""
    mg = _golden_inputs()
    pts, jacs = mg.acceptance_points()
    assert [int(O.ranks_ok(*p)) for p in pts] == GOLDEN["ranks_ok"]
    for j, want in zip(jacs, GOLDEN["gs_no_worse"]):
        assert [int(O.gs_no_worse(p, j)) for p in pts] == want, j
    assert 0 < sum(GOLDEN["ranks_ok"]) < len(pts) and all(0 < sum(w) < len(pts) for w in GOLDEN["gs_no_worse"])
    if O.ref_lib() is not None:
        assert [int(O.ref_ranks_ok(*p)) for p in pts] == GOLDEN["ranks_ok"]
        assert [int(O.ref_gs_no_worse(p, jacs[1])) for p in pts] == GOLDEN["gs_no_worse"][1]


# ------------------------------------------------------------------- MLP
def _tiny(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.5, O.lib().or_mlp_nparam())
    x = rng.uniform(-1, 1, (6, 5))
    t = rng.normal(0, 1, (6, 3))
    return p, x, t


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mlp_gradient_finite_differences(seed):
    p, x, t = _tiny(seed)
    mask = (np.random.default_rng(seed).uniform(size=(6, 3)) < 0.7).astype(np.uint8)
    for mk in (None, mask):
        loss, g = O.mlp_loss_grad(p, x, t, mk)
        h = 1e-6
        for i in range(0, len(p), 7):  # every 7th parameter covers all six blocks
            q = p.copy()
            q[i] += h
            lp, _ = O.mlp_loss_grad(q, x, t, mk)
            q[i] -= 2 * h
            lm, _ = O.mlp_loss_grad(q, x, t, mk)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_forward_against_numpy():
    p, x, _ = _tiny(5)
    W1, b1 = p[:160].reshape(32, 5), p[160:192]
    W2, b2 = p[192:1216].reshape(32, 32), p[1216:1248]
    W3, b3 = p[1248:1344].reshape(3, 32), p[1344:1347]
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    np.testing.assert_allclose(O.mlp_forward(p, x), h2 @ W3.T + b3, rtol=1e-13, atol=1e-13)
    loss, _ = O.mlp_loss_grad(p, x, np.zeros((6, 3)))
    assert abs(loss - np.mean((h2 @ W3.T + b3) ** 2)) < 1e-13


# ------------------------------------------------- against the reference build
ref_built = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref (reference CPU build) not built")


@ref_built
@pytest.mark.parametrize("name", ["rmat12", "grid100x100", "two_components"])
def test_oracle_equals_compiled_reference(name):
    g = _sym(name)
    rp, ci = g.row_ptr, g.col_idx
    for s in sample_sources(g.n):
        assert np.array_equal(O.bfs_serial(rp, ci, s), O.ref_bfs_serial(rp, ci, s))
        d1, l1 = O.bfs_do_levels(rp, ci, s)
        d2, l2 = O.ref_bfs_cpu(rp, ci, s, "cpu_do", threads=1)
        assert np.array_equal(d1, d2) and l1 == l2
    gd = G.Graph.named(name)
    t = gd.transpose()
    r1, i1, _ = O.pagerank(t.row_ptr, t.col_idx, gd.degrees(), iters=30)
    r2, i2, _ = O.ref_pagerank(t.row_ptr, t.col_idx, gd.degrees(), iters=30)
    assert i1 == i2 and np.abs(r1 - r2).max() < 1e-15
