"""Parity at the full sizes of the workload graphs (gen_named recipes):

* rmat16_c1 (scale 16, ef 16, permuted, seed 1): gpu_do from all 64 pickSources
  sources bit for bit against cpu_serial, the per-level (direction, frontier)
  of cpu_do for eight of them; PageRank for 20 fixed FP32 iterations (scalar,
  adaptive) against the oracle's 20 FP64 iterations.
* grid_c3 (1024 x 2048 grid, p = 0.7, largest component, 2390 levels):
  gpu_do depths and per-level (direction, frontier) through the host,
  persistent and graph loops and hybrid; gpu_pull_scalar for 20 iterations.
* rmat22_c2 (scale 22, ef 8, permuted): single-source gpu_do, the 64-source
  batch under the learned policy (six sources against the oracle; all 64 rows
  equal to the caller-numbering batch and, like all 256 of a 256-source
  batch, to each source's own sequential gpu_do run), gpu_pull_gs to tol 1e-6
  judged against the converged FP64 ranks.
* rmat23_c5 (scale 23, ef 16, a = .60): Jacobi certified to L1 < 1e-4; the
  64-seed personalized-PageRank block: every column keeps mass 1, two work
  assignments (given numbering, no chunks / degree numbering, 1024-entry
  chunks, lane groups) agree column by column, and one column matches the
  oracle's FP64 PPR after 6 sweeps.

Checkers: the oracle's serial BFS and its cpu_do level restatement; the
FP64 PageRank of the compiled reference (oracle/_ref, OpenMP) when it is
built, else the oracle's. Acceptance is the reference's checkRanks rule
(src/cpu/pagerank_cpu.hpp:37-60) and gsNoWorseThanJacobi (:80-84).
"""
import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _pr64(rp, ci, od, iters, tol):
    if O.ref_lib() is not None:
        return O.ref_pagerank(rp, ci, od, 0.85, iters, tol, parallel=True)[0]
    return O.pagerank(rp, ci, od, 0.85, iters, tol)[0]


def test_rmat16_c1_gpu_do_64_sources_and_pagerank_20():
    g = G.Graph.named("rmat16_c1")
    rp, ci = g.row_ptr, g.col_idx
    bfs = G.GpuBfs(g)
    src = g.pick_sources(64, 1)
    assert len(src) == 64
    for i, s in enumerate(src):
        s = int(s)
        r = bfs.run("gpu_do", s)
        assert np.array_equal(r.dist, O.bfs_serial(rp, ci, s)), s
        if i % 8 == 0:
            _, levels = O.bfs_do_levels(rp, ci, s)
            assert [(L.direction.decode(), L.frontier) for L in r.trace] == [(d, f) for d, f, _ in levels], s
    od = g.degrees()
    ref = O.pagerank(rp, ci, od, 0.85, 20, 0.0)[0]
    pr = G.GpuPageRank(g, od)
    for variant in ("gpu_pull_scalar", "gpu_pull_adaptive"):
        res = pr.run(variant, G.pr_options(iters=20))
        l1, rel, s = O.check_ranks(ref, pr.ranks())
        assert res.iters == 20 and O.ranks_ok(l1, rel, s), (variant, l1, rel, s)


def test_grid_c3_loops_and_scalar_pagerank():
    g = G.Graph.named("grid_c3")
    rp, ci = g.row_ptr, g.col_idx
    bfs = G.GpuBfs(g)
    for src in g.pick_sources(2, 3):
        src = int(src)
        ref = O.bfs_serial(rp, ci, src)
        _, levels = O.bfs_do_levels(rp, ci, src)
        want = [(d, f) for d, f, _ in levels]
        assert len(want) > 1500
        for tag, kw in (("host", dict(loop=G.LOOP_HOST)), ("persistent", dict(loop=G.LOOP_PERSISTENT)),
                        ("graph-k64", dict(loop=G.LOOP_GRAPH, graph_levels=64))):
            r = bfs.run("gpu_do", src, G.bfs_options(**kw))
            assert np.array_equal(r.dist, ref), (tag, src)
            assert [(L.direction.decode(), L.frontier) for L in r.trace] == want, (tag, src)
        r = bfs.run("hybrid", src)
        assert np.array_equal(r.dist, ref), ("hybrid", src)
    od = g.degrees()
    ref = O.pagerank(rp, ci, od, 0.85, 20, 0.0)[0]
    pr = G.GpuPageRank(g, od)
    res = pr.run("gpu_pull_scalar", G.pr_options(iters=20))
    l1, rel, s = O.check_ranks(ref, pr.ranks())
    assert res.iters == 20 and l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4, (l1, rel, s)


def test_rmat22_c2_gpu_do_batch_and_gs():
    g = G.Graph.named("rmat22_c2")
    rp, ci = g.row_ptr, g.col_idx
    bfs = G.GpuBfs(g)
    src = g.pick_sources(64, 2)
    assert len(src) == 64
    checked = [int(s) for s in src[:6]]
    refs = {s: O.bfs_serial(rp, ci, s) for s in checked}
    for s in checked[:3]:
        r = bfs.run("gpu_do", s)
        assert np.array_equal(r.dist, refs[s]), s
    G.install_policies(bfs.ctx)
    dist, res, _ = bfs.run_batch(src, G.bfs_options(policy=G.POLICY_LEARNED))
    for i, s in enumerate(checked):
        assert np.array_equal(dist[i], refs[s]), (i, s)
    # all 64 rows: the default run (degree numbering, depths un-permuted) equals the
    # batch traversed in the caller's numbering, which has neither step
    given, _, _ = bfs.run_batch(src, G.bfs_options(policy=G.POLICY_LEARNED, batch_order=G.BATCH_ORDER_GIVEN))
    assert np.array_equal(dist, given)
    del given
    # every source of the 64 and of a 256-source batch bit-exact against its own
    # sequential heuristic gpu_do run
    seq = {}
    for i, s in enumerate(src):
        seq[int(s)] = bfs.run("gpu_do", int(s), trace=False).dist
        assert np.array_equal(dist[i], seq[int(s)]), (i, int(s))
    del dist
    src256 = g.pick_sources(256, 2)
    dist256, _, _ = bfs.run_batch(src256, G.bfs_options(policy=G.POLICY_LEARNED))
    for i, s in enumerate(src256):
        want = seq[int(s)] if int(s) in seq else bfs.run("gpu_do", int(s), trace=False).dist
        assert np.array_equal(dist256[i], want), (i, int(s))
    del dist256, seq
    od = g.degrees()
    conv = _pr64(rp, ci, od, 1000, 1e-10)
    jac = _pr64(rp, ci, od, 1000, 1e-6)
    jl1 = float(np.abs(conv - jac).sum())
    top = np.argsort(-conv, kind="stable")[:1000]
    jerr = (jl1, float((np.abs(conv[top] - jac[top]) / conv[top]).max()), float(jac.sum()))
    pr = G.GpuPageRank(g, od)
    pr.run("gpu_pull_gs", G.pr_options(tol=1e-6, iters=1000))
    c = O.check_ranks(conv, pr.ranks())
    assert O.ranks_ok(*c) and O.gs_no_worse(c, jerr), (c, jerr)


def test_rmat23_c5_certified_jacobi():
    if O.ref_lib() is None:
        pytest.skip("the converged FP64 ranks of 251M arcs need the compiled reference's OpenMP PageRank")
    g = G.Graph.named("rmat23_c5")
    rp, ci = g.row_ptr, g.col_idx
    od = g.degrees()
    conv = _pr64(rp, ci, od, 1000, 1e-10)
    pr = G.GpuPageRank(g, od)
    tol_cert = 1e-4 * 0.15 / 0.85  # ||r_k - r*||_1 <= d / (1 - d) * delta_k
    for layout in (G.PR_LAYOUT_HOT, G.PR_LAYOUT_STREAM):
        res = pr.run("gpu_pull_adaptive", G.pr_options(tol=tol_cert, iters=1000, order=G.PR_ORDER_DEGREE,
                                                       layout=layout))
        l1, rel, s = O.check_ranks(conv, pr.ranks())
        assert not res.floor_stop and res.last_delta < tol_cert, (layout, res.last_delta)
        assert l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4, (layout, l1, rel, s)


def test_rmat23_c5_ppr_block():
    g = G.Graph.named("rmat23_c5")
    od = g.degrees()
    B = 64
    seeds = np.asarray(g.pick_sources(B, 5), dtype=np.uint32)
    pr = G.GpuPageRank(g, od)
    sweeps = 6
    pr.run_ppr(seeds, G.pr_options(iters=sweeps, order=G.PR_ORDER_DEGREE, large_min=1024, tg=1024))
    a = pr.ppr_block()
    sums = a.sum(axis=0, dtype=np.float64)
    assert np.abs(sums - 1).max() < 1e-4, np.abs(sums - 1).max()
    ref, _, _ = O.ppr(g.row_ptr, g.col_idx, od, int(seeds[0]), 0.85, sweeps)
    err = float(np.abs(a[:, 0].astype(np.float64) - ref).sum())
    assert err < 1e-5, err
    pr.run_ppr(seeds, G.pr_options(iters=sweeps, order=G.PR_ORDER_GIVEN, large_min=0xFFFFFFFF, tg=256))
    b = pr.ppr_block()
    diff = np.abs(a.astype(np.float64) - b).sum(axis=0)
    assert diff.max() < 1e-5, diff.max()
