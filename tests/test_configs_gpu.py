"""Parity at the full sizes of the workload graphs (gen_named recipes):

* rmat16_c1 (scale 16, ef 16, permuted, seed 1): gpu_do from all 64 pickSources
  sources bit for bit against cpu_serial, the per-level (direction, frontier)
  of cpu_do for eight of them; PageRank for 20 fixed FP32 iterations (scalar,
  adaptive) against the oracle's 20 FP64 iterations.
* grid_c3 (1024 x 2048 grid, p = 0.7, largest component, 2390 levels):
  gpu_do depths and per-level (direction, frontier) through the host,
  persistent and graph loops and hybrid; gpu_pull_scalar for 20 iterations.
* rmat22_c2 (scale 22, ef 8, permuted): for every one of the 256 pickSources
  sources, the row of a 256-source batch against the oracle's serial BFS, and
  for the first 64 also single-source gpu_do and the row of the 64-source
  batch (learned policy; the 64 rows also equal the batch traversed in the
  caller's numbering); gpu_pull_gs to tol 1e-6 judged against the converged
  FP64 ranks.
* rmat23_c5 (scale 23, ef 16, a = .60): Jacobi certified to L1 < 1e-4, and
  in-place Gauss-Seidel (gpu_pull_gs) to L1 < 1e-4 and no worse than Jacobi
  at the same tolerance; the 64-seed personalized-PageRank block: every
  column keeps mass 1, two work assignments (given numbering, no chunks /
  degree numbering, 1024-entry chunks, lane groups) agree column by column,
  and eight columns match the oracle's FP64 PPR after 6 sweeps.

The oracle runs at these sizes take seconds each (a serial BFS of rmat22:
about 1 s, a 6-sweep PPR column of rmat23: about 11 s), so they run on a pool
of host threads (the C calls release the interpreter lock).

Checkers: the oracle's serial BFS and its cpu_do level restatement; the
FP64 PageRank of the compiled reference (oracle/_ref, OpenMP) when it is
built, else the oracle's. Acceptance is the reference's checkRanks rule
(src/cpu/pagerank_cpu.hpp:37-60) and gsNoWorseThanJacobi (:80-84).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _pr64(rp, ci, od, iters, tol):
    if O.ref_lib() is not None:
        return O.ref_pagerank(rp, ci, od, 0.85, iters, tol, parallel=True)[0]
    return O.pagerank(rp, ci, od, 0.85, iters, tol)[0]


def _workers(cap=16):
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cap, cores))


def _oracle_rows(rp, ci, sources):
    """(i, source, the oracle's serial BFS depths) for every source, in order,
    computed on a pool of host threads."""
    with ThreadPoolExecutor(_workers()) as ex:
        yield from zip(range(len(sources)), sources, ex.map(lambda s: O.bfs_serial(rp, ci, s), sources))


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
    src256 = [int(s) for s in g.pick_sources(256, 2)]
    src = [int(s) for s in g.pick_sources(64, 2)]
    assert len(src256) == 256 and src == src256[:64]
    G.install_policies(bfs.ctx)
    learned = G.bfs_options(policy=G.POLICY_LEARNED)
    dist, _, _ = bfs.run_batch(np.array(src, dtype=np.uint32), learned)
    # the default run (degree numbering, depths un-permuted) equals the batch
    # traversed in the caller's numbering, which has neither step
    given, _, _ = bfs.run_batch(np.array(src, dtype=np.uint32),
                                G.bfs_options(policy=G.POLICY_LEARNED, batch_order=G.BATCH_ORDER_GIVEN))
    assert np.array_equal(dist, given)
    del given
    dist256, _, _ = bfs.run_batch(np.array(src256, dtype=np.uint32), learned)
    for i, s, ref in _oracle_rows(rp, ci, src256):
        assert np.array_equal(dist256[i], ref), ("256-source batch", i, s)
        if i < 64:
            assert np.array_equal(dist[i], ref), ("64-source batch", i, s)
            assert np.array_equal(bfs.run("gpu_do", s, trace=False).dist, ref), ("gpu_do", s)
    del dist, dist256
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
    for layout, lmin in ((G.PR_LAYOUT_HOT, 256), (G.PR_LAYOUT_STREAM, 256), (G.PR_LAYOUT_HOT_TAIL, 64),
                         (G.PR_LAYOUT_HOT_TAIL, 256)):
        res = pr.run("gpu_pull_adaptive", G.pr_options(tol=tol_cert, iters=1000, order=G.PR_ORDER_DEGREE,
                                                       layout=layout, large_min=lmin))
        l1, rel, s = O.check_ranks(conv, pr.ranks())
        assert not res.floor_stop and res.last_delta < tol_cert, (layout, lmin, res.last_delta)
        assert l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4, (layout, lmin, l1, rel, s)
    # in-place Gauss-Seidel to tol 1e-6: within the reference's acceptance of the
    # converged ranks, and no worse than FP64 Jacobi stopped at the same tolerance
    for order in (G.PR_ORDER_DEGREE, G.PR_ORDER_GIVEN):
        res = pr.run("gpu_pull_gs", G.pr_options(tol=1e-6, iters=1000, order=order))
        c = O.check_ranks(conv, pr.ranks())
        assert not res.floor_stop and c[0] < 1e-4 and O.ranks_ok(*c), (order, res.iters, c)
    jac = _pr64(rp, ci, od, 1000, 1e-6)
    top = np.argsort(-conv, kind="stable")[:1000]
    jerr = (float(np.abs(conv - jac).sum()), float((np.abs(conv[top] - jac[top]) / conv[top]).max()),
            float(jac.sum()))
    assert O.gs_no_worse(c, jerr), (c, jerr)


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
    cols = list(range(0, B, B // 8))  # eight columns spread over the block
    with ThreadPoolExecutor(_workers(8)) as ex:
        refs = ex.map(lambda j: O.ppr(g.row_ptr, g.col_idx, od, int(seeds[j]), 0.85, sweeps)[0], cols)
        for j, ref in zip(cols, refs):
            err = float(np.abs(a[:, j].astype(np.float64) - ref).sum())
            assert err < 1e-5, (j, int(seeds[j]), err)
    pr.run_ppr(seeds, G.pr_options(iters=sweeps, order=G.PR_ORDER_GIVEN, large_min=0xFFFFFFFF, tg=256))
    b = pr.ppr_block()
    diff = np.abs(a.astype(np.float64) - b).sum(axis=0)
    assert diff.max() < 1e-5, diff.max()
