"""GPU PageRank parity against the FP64 oracle.

Mirrors tests/test_main.cpp:630-783: every pull variant and option extreme
on every suite graph, fixed 20 sweeps and tol 1e-6, accepted by the
reference's rule (L1 < 1e-4, top-1000 relative error < 1e-3, |sum-1| < 1e-4;
src/cpu/pagerank_cpu.hpp:37-60) with the iteration count within 3 of the
oracle's; gpu_pull_gs judged against the converged oracle
(gsNoWorseThanJacobi, pagerank_cpu.hpp:80-84). Tolerances are the
reference's acceptance thresholds, written in the assertions below.
"""
import ctypes

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O
from tests.conftest import SUITE

pytestmark = pytest.mark.gpu


def _in_graph(name):
    g = G.Graph.named(name)
    t = g.transpose() if g.directed else g
    return g, t, g.degrees()


B = G.PR_LAYOUT_BINS
CFGS = [("gpu_pull_scalar", {}), ("gpu_pull_vector", {}),
        ("gpu_pull_adaptive", {}),                                          # stream layout (default)
        ("gpu_pull_adaptive", dict(order=G.PR_ORDER_DEGREE)),               # stream on the degree-ordered copy
        ("gpu_pull_adaptive", dict(order=G.PR_ORDER_DEGREE, layout=1)),
        ("gpu_pull_scalar", dict(order=G.PR_ORDER_DEGREE)),
        ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_HOT)),                # LDS-cached contributions
        ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_HOT, order=G.PR_ORDER_DEGREE)),
        ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_TILED, order=G.PR_ORDER_DEGREE, large_min=4)),  # source tiles
        ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_TILED, order=G.PR_ORDER_DEGREE, large_min=128)),
        ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_TILED, large_min=1)),            # every row with arcs tiled
        ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_TILED, large_min=1024)),         # almost no row tiled
        ("gpu_pull_adaptive", dict(layout=B)),                              # the reference's 16 / 256 bins
        ("gpu_pull_adaptive", dict(layout=B, small_max=1 << 30, large_min=1 << 31)),  # all lane-per-vertex
        ("gpu_pull_adaptive", dict(layout=B, small_max=0, large_min=0)),              # all workgroup-per-vertex
        ("gpu_pull_adaptive", dict(layout=B, small_max=0, large_min=1 << 30)),        # all medium
        ("gpu_pull_adaptive", dict(layout=B, small_max=2, large_min=3)),
        ("gpu_pull_adaptive", dict(layout=B, small_max=4, large_min=64, medium_lanes=8)),
        ("gpu_pull_adaptive", dict(layout=B, small_max=8, large_min=512, medium_lanes=16)),
        ("gpu_push", {}),                                                   # atomics along out-arcs
        ("gpu_push", dict(order=G.PR_ORDER_DEGREE))]


@pytest.mark.parametrize("name", SUITE)
def test_pagerank_variants_match_oracle(name):
    """On empty1000 (no arcs) the TILED sweep's empty lanes read the placeholder
    arc 0 and gather the contribution it names: the placeholder must hold a
    valid source id (it once held whatever the allocation held before, and a
    stale value sent that gather out of bounds)."""
    g, t, od = _in_graph(name)
    pr = G.GpuPageRank(t, od)
    for tol in (0.0, 1e-6):
        iters = 1000 if tol > 0 else 20
        ref, ref_iters, _ = O.pagerank(t.row_ptr, t.col_idx, od, 0.85, iters, tol)
        for variant, kw in CFGS:
            for tg in (64, 256, 1024):
                res = pr.run(variant, G.pr_options(iters=iters, tol=tol, tg=tg, **kw))
                l1, rel, s = O.check_ranks(ref, pr.ranks())
                assert l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4, \
                    f"{name} {variant} {kw} tol {tol} tg {tg}: L1 {l1:.3g} rel {rel:.3g} sum {s:.8f}"
                if tol > 0 and not res.floor_stop:
                    assert abs(res.iters - ref_iters) <= 3, f"{name} {variant}: {res.iters} vs {ref_iters}"
                if tol == 0:
                    assert res.iters == 20
                    assert abs(res.sum - 1.0) < 1e-4  # FP64 sum of the FP32 ranks


def test_pagerank_known_answers_on_gpu():
    """Closed forms of tests/test_main.cpp:630-652 on the GPU (FP32, to 1e-6)."""
    cases = [(2, [(0, 1)], [0.5 / 1.425, 1 - 0.5 / 1.425]),
             (3, [(0, 1), (1, 2), (2, 0)], [1 / 3] * 3)]
    for n, edges, want in cases:
        g = G.Graph.from_edges(n, edges, directed=True)
        t = g.transpose()
        pr = G.GpuPageRank(t, g.degrees())
        for v in ("gpu_pull_scalar", "gpu_pull_vector", "gpu_pull_adaptive", "gpu_pull_gs", "gpu_push"):
            pr.run(v, G.pr_options(iters=200))
            assert np.allclose(pr.ranks(), want, atol=1e-6), v
    for name, n in (("complete200", 200), ("empty1000", 1000)):
        g = G.Graph.named(name)
        pr = G.GpuPageRank(g, g.degrees())
        pr.run("gpu_pull_adaptive", G.pr_options(iters=100))
        assert np.allclose(pr.ranks(), 1.0 / n, rtol=1e-5)


def test_pagerank_gs_converges_to_oracle():
    colored = natural = deferred_graphs = 0
    failures = []
    for name in SUITE:
        g, t, od = _in_graph(name)
        if g.n == 0:
            continue
        conv, _, _ = O.pagerank(t.row_ptr, t.col_idx, od, 0.85, 1000, 1e-10)
        jac, _, _ = O.pagerank(t.row_ptr, t.col_idx, od, 0.85, 1000, 1e-6)
        jerr = _check64(conv, jac)  # the Jacobi reference's own distance from convergence
        for K in (0, 2):
            pr = G.GpuPageRank(t, od)
            pr.prepare_gs(K)
            is_col, sets, dfr = pr.gs_info()
            colored += is_col
            natural += not is_col
            deferred_graphs += dfr > 0
            for tg in (64, 256, 1024):
                for what, tol, tol_rel, iters, kw in (("20 sweeps", 0.0, 0.0, 20, {}),
                                                      ("200 sweeps", 0.0, 0.0, 200, {}),
                                                      ("200 sweeps bins", 0.0, 0.0, 200, dict(layout=1)),
                                                      ("200 sweeps degree", 0.0, 0.0, 200, dict(order=1)),
                                                      ("tol 2.5e-7", 2.5e-7, 0.0, 1000, {}),
                                                      ("tol 2.5e-7 bins", 2.5e-7, 0.0, 1000, dict(layout=1)),
                                                      ("tol 1e-6 + rel", 1e-6, 2.5e-5, 1000, {})):
                    res = pr.run("gpu_pull_gs", G.pr_options(iters=iters, tol=tol, tol_rel=tol_rel, tg=tg, **kw))
                    c = O.check_ranks(conv, pr.ranks())
                    if iters == 20:
                        ok = abs(c[2] - 1.0) < 1e-4
                    elif iters == 200:
                        ok = O.ranks_ok(*c)
                    else:
                        ok = O.gs_no_worse(c, jerr)
                    if not ok:
                        failures.append((name, K, tg, what, c, res.iters, jerr))
    assert not failures, failures[:5]
    assert colored > 0 and natural > 0 and deferred_graphs > 0


def _check64(ref, got):
    l1 = float(np.abs(ref - got).sum())
    top = np.argsort(-ref, kind="stable")[: min(1000, len(ref))]
    rel = float((np.abs(ref[top] - got[top]) / ref[top]).max()) if len(top) else 0.0
    return l1, rel, float(got.sum())


@pytest.mark.parametrize("name", ["rmat12_directed", "grid100x100", "star100k", "two_components"])
def test_inplace_jacobi_variants_converge(name):
    g, t, od = _in_graph(name)
    conv, _, _ = O.pagerank(t.row_ptr, t.col_idx, od, 0.85, 1000, 1e-10)
    pr = G.GpuPageRank(t, od)
    for v, kw in (("gpu_pull_scalar", {}), ("gpu_pull_vector", {}), ("gpu_pull_adaptive", {}),
                  ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_BINS)),
                  ("gpu_pull_adaptive", dict(layout=G.PR_LAYOUT_HOT))):  # in place: row blocks
        pr.run(v, G.pr_options(iters=200, in_place=1, **kw))
        c = O.check_ranks(conv, pr.ranks())
        assert O.ranks_ok(*c), (v, kw, c)


def test_pr_errors_are_reported():
    g, t, od = _in_graph("rmat12_directed")
    pr = G.GpuPageRank(t, od)
    with pytest.raises(RuntimeError):
        pr.ranks()  # no run yet
    with pytest.raises(ValueError):
        pr.run("gpu_pull_adaptive", G.pr_options(tg=100))
    with pytest.raises(ValueError):
        pr.run("gpu_pull_adaptive", G.pr_options(medium_lanes=3))
    with pytest.raises(ValueError):
        pr.run("gpu_pull_adaptive", G.pr_options(layout=G.PR_LAYOUT_HOT_TAIL + 1))
    with pytest.raises(ValueError):
        pr.run("gpu_pull_adaptive", G.pr_options(layout=G.PR_LAYOUT_TILED, large_min=0))
    with pytest.raises(ValueError):
        pr.run("gpu_pull_adaptive", G.pr_options(layout=G.PR_LAYOUT_TILED, large_min=1025))
    with pytest.raises(ValueError):
        pr.run("gpu_push", G.pr_options(in_place=1))


# PPR work assignments: the defaults (heavy in-lists >= 256 entries in chunks,
# a wave per light vertex), every list of >= 64 entries chunked with lane
# groups of 1 and 16, no chunking at all, and the degree-ordered numbering.
PPR_ASSIGN = [dict(), dict(large_min=64, medium_lanes=1, tg=64), dict(large_min=64, medium_lanes=16, tg=1024),
              dict(large_min=0xFFFFFFFF, medium_lanes=64, tg=256),
              dict(order=G.PR_ORDER_DEGREE, large_min=64, medium_lanes=4, tg=256)]


def _ppr_check(tag, blk, refs, tol=2e-5):
    for b, ref in enumerate(refs):
        col = blk[:, b].astype(np.float64)
        err = float(np.abs(col - ref).sum())
        assert err < tol, f"{tag} col {b}: L1 {err:.3g}"
        assert abs(float(col.sum()) - 1.0) < tol, f"{tag} col {b}: sum {col.sum()}"


@pytest.mark.parametrize("name", ["rmat12_directed", "grid100x100", "two_components", "path1000"])
def test_ppr_block_matches_oracle(name):
    g, t, od = _in_graph(name)
    pr = G.GpuPageRank(t, od)
    rng = np.random.default_rng(3)
    for B in (1, 5, 64, 100):
        seeds = rng.integers(0, g.n, size=B).astype(np.uint32)
        refs = [O.ppr(t.row_ptr, t.col_idx, od, int(s), 0.85, 60)[0] for s in seeds]
        for kw in PPR_ASSIGN:
            pr.run_ppr(seeds, G.pr_options(iters=60, **kw))
            blk = pr.ppr_block()
            assert blk.shape == (g.n, B)
            _ppr_check(f"{name} B={B} {kw}", blk, refs)


def test_ppr_hub_chunks():
    """The star's hub has in-degree 100000: 98 chunks of 1024 entries (the last
    one short) combined in chunk order, against a single wave walking the whole
    list; 3 and 130 columns (one and three column passes), seeds on the hub
    and on leaves."""
    g, t, od = _in_graph("star100k")
    pr = G.GpuPageRank(t, od)
    hub = int(np.argmax(np.diff(t.row_ptr)))
    for B in (3, 130):
        seeds = np.array([hub] + [1 + 997 * i % (g.n - 1) for i in range(B - 1)], dtype=np.uint32)
        refs = [O.ppr(t.row_ptr, t.col_idx, od, int(s), 0.85, 30)[0] for s in seeds]
        for kw in (dict(large_min=1024), dict(large_min=1000, medium_lanes=4), dict(large_min=0xFFFFFFFF),
                   dict(large_min=1024, order=G.PR_ORDER_DEGREE)):
            pr.run_ppr(seeds, G.pr_options(iters=30, **kw))
            _ppr_check(f"star B={B} {kw}", pr.ppr_block(), refs)


def test_ppr_autotune_and_errors():
    g, t, od = _in_graph("rmat12_directed")
    pr = G.GpuPageRank(t, od)
    seeds = np.arange(0, 8 * 97, 97, dtype=np.uint32)
    o = pr.autotune_ppr(seeds, iters=2, opts=G.pr_options(iters=40))
    assert o.large_min in (64, 128, 256, 1024) and o.medium_lanes in (8, 16, 32, 64) and o.tg in (256, 1024)
    assert o.order in (G.PR_ORDER_GIVEN, G.PR_ORDER_DEGREE)
    assert o.iters == 40
    pr.run_ppr(seeds, o)
    refs = [O.ppr(t.row_ptr, t.col_idx, od, int(s), 0.85, 40)[0] for s in seeds]
    _ppr_check(f"autotuned {o.large_min} {o.medium_lanes} {o.tg}", pr.ppr_block(), refs)
    for kw in (dict(large_min=63), dict(medium_lanes=3), dict(medium_lanes=128), dict(tg=96), dict(order=2)):
        with pytest.raises(ValueError):
            pr.run_ppr(seeds, G.pr_options(**kw))
    with pytest.raises(ValueError):
        pr.run_ppr(np.zeros(257, dtype=np.uint32))
    with pytest.raises(ValueError):
        pr.run_ppr(np.array([g.n], dtype=np.uint32))


def test_ppr_isolated_seed_is_one_hot():
    g = G.Graph.named("two_components")
    iso = G.Graph.from_edges(5, [(0, 1), (1, 2)], directed=False)  # 3, 4 isolated
    pr = G.GpuPageRank(iso, iso.degrees())
    pr.run_ppr([3, 4, 0], G.pr_options(iters=30))
    blk = pr.ppr_block()
    assert np.allclose(blk[:, 0], [0, 0, 0, 1, 0]) and np.allclose(blk[:, 1], [0, 0, 0, 0, 1])
    assert abs(blk[:, 2].sum() - 1) < 1e-6
    del g


def test_stream_layout_long_rows_and_autotune():
    """In-lists longer than a row block are split into chunks and recombined in
    chunk order: the star's hub (in-degree 100000 = 13 chunks) and a 2-chunk
    row must match the oracle; the autotuner's pick must leave ranks valid."""
    for name in ("star100k", "rmat12_directed"):
        g, t, od = _in_graph(name)
        ref, _, _ = O.pagerank(t.row_ptr, t.col_idx, od, 0.85, 20, 0.0)
        pr = G.GpuPageRank(t, od)
        for tg in (64, 256, 1024):
            pr.run("gpu_pull_adaptive", G.pr_options(iters=20, tg=tg))
            assert O.ranks_ok(*O.check_ranks(ref, pr.ranks())), (name, tg)
        for order in (G.PR_ORDER_GIVEN, G.PR_ORDER_DEGREE):  # hub: 196 wave chunks
            pr.run("gpu_pull_adaptive", G.pr_options(iters=20, layout=G.PR_LAYOUT_HOT, order=order))
            assert O.ranks_ok(*O.check_ranks(ref, pr.ranks())), (name, "hot", order)
        o = pr.autotune(iters=3)
        pr.run("gpu_pull_adaptive", G.pr_options(iters=20, layout=o.layout, order=o.order, small_max=o.small_max,
                                                 large_min=o.large_min, medium_lanes=o.medium_lanes))
        assert O.ranks_ok(*O.check_ranks(ref, pr.ranks())), (name, o.layout, o.small_max, o.large_min)
    # a hub with 10000 in-neighbours: 2 chunks of the long-row path
    edges = [(i, 0) for i in range(1, 10001)] + [(0, 1)]
    g = G.Graph.from_edges(10001, edges, directed=True)
    t = g.transpose()
    ref, _, _ = O.pagerank(t.row_ptr, t.col_idx, g.degrees(), 0.85, 50, 0.0)
    pr = G.GpuPageRank(t, g.degrees())
    for layout in (G.PR_LAYOUT_STREAM, G.PR_LAYOUT_HOT):
        pr.run("gpu_pull_adaptive", G.pr_options(iters=50, layout=layout))
        assert O.ranks_ok(*O.check_ranks(ref, pr.ranks())), layout


def test_hot_layout_cache_boundary():
    """More vertices than the LDS cache holds (65536 > 30720): gathers below
    the boundary come from LDS, the rest from HBM; both orders, fixed sweeps
    and tol mode, against the FP64 oracle with the acceptance rule above."""
    g = G.Graph.named("rmat16_c1")
    od = g.degrees()
    pr = G.GpuPageRank(g, od)
    for tol, iters in ((0.0, 20), (1e-6, 1000)):
        ref, ref_iters, _ = O.pagerank(g.row_ptr, g.col_idx, od, 0.85, iters, tol)
        for order in (G.PR_ORDER_GIVEN, G.PR_ORDER_DEGREE):
            res = pr.run("gpu_pull_adaptive", G.pr_options(iters=iters, tol=tol, layout=G.PR_LAYOUT_HOT, order=order))
            l1, rel, s = O.check_ranks(ref, pr.ranks())
            assert l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4, (tol, order, l1, rel, s)
            if tol > 0 and not res.floor_stop:
                assert abs(res.iters - ref_iters) <= 3


def _tail_geometry(pr, tail_begin, tile, band_end=0, band_tile=0):
    pr.ctx.check(pr.ctx.lib.ggpu_test_pr_tail_geometry(pr._h, tail_begin, tile, band_end, band_tile))


def _tail_info(pr, order):
    """(split rows, tiles, tail arcs, items, rows merged by a wave, head chunks, band tiles, band tasks)
    of the last HOT_TAIL layout."""
    st = (ctypes.c_uint64 * 8)()
    pr.ctx.check(pr.ctx.lib.ggpu_test_pr_tail_info(pr._h, order, st))
    return [int(x) for x in st]


# first tail source, sources per tail tile, split-row in-degree (large_min), band end, sources per band tile
TAIL = [(64, 100, 1, 0, 0), (64, 100, 4, 0, 0), (16, 7, 0, 0, 0), (1024, 512, 2, 0, 0),
        (64, 300, 1, 1000, 100), (16, 7, 2, 500, 37), (0, 1000, 1, 64, 64)]


@pytest.mark.parametrize("name", SUITE)
def test_hot_tail_layout_matches_oracle(name):
    """HOT_TAIL with small first-tail ids and tiles (the suite graphs lie far
    below the default geometry): the split rows' heads summed by the HOT sweep
    as partials, their band arcs per LDS tile by the band kernel, their tails
    per tile by the tail kernel, each split row merged from head, tile and
    overflow partials; with and without a band, both orders, several split
    thresholds, fixed sweeps and tol mode, against the FP64 oracle with the
    acceptance rule above (L1 < 1e-4, top-1000 relative < 1e-3, |sum-1| < 1e-4)."""
    g, t, od = _in_graph(name)
    pr = G.GpuPageRank(t, od)
    split = tasks = 0
    for tol, iters in ((0.0, 20), (1e-6, 1000)):
        ref, ref_iters, _ = O.pagerank(t.row_ptr, t.col_idx, od, 0.85, iters, tol)
        for tail_begin, tile, lmin, band_end, band_tile in TAIL:
            _tail_geometry(pr, tail_begin, tile, band_end, band_tile)
            for order in (G.PR_ORDER_GIVEN, G.PR_ORDER_DEGREE):
                res = pr.run("gpu_pull_adaptive", G.pr_options(iters=iters, tol=tol, layout=G.PR_LAYOUT_HOT_TAIL,
                                                               order=order, large_min=lmin))
                l1, rel, s = O.check_ranks(ref, pr.ranks())
                assert l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4, \
                    f"{name} tol {tol} tail {tail_begin}/{tile}/{lmin}/{band_end}/{band_tile} order {order}: " \
                    f"L1 {l1:.3g} rel {rel:.3g} sum {s}"
                if tol > 0 and not res.floor_stop:
                    assert abs(res.iters - ref_iters) <= 3, (name, res.iters, ref_iters)
                if tol == 0:
                    assert res.iters == 20 and abs(res.sum - 1.0) < 1e-4
                info = _tail_info(pr, order)
                split += info[0]
                tasks += info[7]
                if not band_tile:
                    assert info[6] == 0 and info[7] == 0
    if g.n > 64 and g.m:
        assert split > 0, "no configuration split a row"
        assert tasks > 0, "no configuration ran the band kernel"


def test_hot_tail_hub_segments_and_wave_merge():
    """rmat16 (65,536 vertices) in the degree order with 4096-source tiles past
    id 4096, with and without a band of 4096- or 16384-source LDS tiles, and the
    default geometry (three band tiles past id 30720): split hub rows whose
    tile segments exceed one item (overflow partials), rows merged by a whole
    wave, rows merged by a lane; HOT_TAIL equals the FP64 oracle by the
    acceptance rule and agrees with HOT sweep for sweep (both compensated FP32,
    summed in different orders: within 5e-6 relative on every rank)."""
    g = G.Graph.named("rmat16_c1")
    od = g.degrees()
    pr = G.GpuPageRank(g, od)
    ref, _, _ = O.pagerank(g.row_ptr, g.col_idx, od, 0.85, 20, 0.0)
    pr.run("gpu_pull_adaptive", G.pr_options(iters=20, layout=G.PR_LAYOUT_HOT, order=G.PR_ORDER_DEGREE))
    hot = pr.ranks()
    geoms = [(4096, 4096, 64, 0, 0), (4096, 4096, 256, 0, 0), (1024, 1000, 1, 0, 0), (30720, 8192, 64, 0, 0),
             (4096, 4096, 64, 4096 + 4 * 4096, 4096), (1024, 8192, 1, 1024 + 3 * 16384, 16384), (0, 0, 256, 0, 0)]
    for tail_begin, tile, lmin, band_end, band_tile in geoms:
        _tail_geometry(pr, tail_begin, tile, band_end, band_tile)
        for tg in (64, 1024):
            pr.run("gpu_pull_adaptive", G.pr_options(iters=20, tg=tg, layout=G.PR_LAYOUT_HOT_TAIL,
                                                     order=G.PR_ORDER_DEGREE, large_min=lmin))
            got = pr.ranks()
            assert O.ranks_ok(*O.check_ranks(ref, got)), (tail_begin, tile, lmin)
            assert np.allclose(got, hot, rtol=5e-6, atol=0), (tail_begin, tile, lmin, np.max(np.abs(got / hot - 1)))
        R, T, tail, items, big, heads, band_tiles, tasks = _tail_info(pr, G.PR_ORDER_DEGREE)
        assert R > 0 and T > 1 and tail > 0 and items >= T, (R, T, tail, items)
        if band_tile:
            assert band_tiles == (band_end - tail_begin + band_tile - 1) // band_tile and tasks >= 1
        if (tail_begin, tile) == (0, 0):  # the default geometry: 16384-source band tiles from 30720
            assert band_tiles == 3 and tasks >= 1, (band_tiles, tasks)
        if (tail_begin, tile, lmin) == (4096, 4096, 64):
            assert big > 0  # hub rows with many head chunks or overflow partials


def test_hot_tail_without_tail_sources_is_hot():
    """rmat12 (4096 vertices) under the default geometry: no source reaches
    id 30720, so no row is split and HOT_TAIL sums exactly as HOT does."""
    g = G.Graph.named("rmat12")
    pr = G.GpuPageRank(g, g.degrees())
    pr.run("gpu_pull_adaptive", G.pr_options(iters=20, layout=G.PR_LAYOUT_HOT, order=G.PR_ORDER_DEGREE))
    hot = pr.ranks()
    pr.run("gpu_pull_adaptive", G.pr_options(iters=20, layout=G.PR_LAYOUT_HOT_TAIL, order=G.PR_ORDER_DEGREE))
    assert _tail_info(pr, G.PR_ORDER_DEGREE)[0] == 0
    assert np.array_equal(pr.ranks(), hot)


def test_hot_tail_geometry_hook_rejects_bad_bands():
    g = G.Graph.named("rmat12")
    pr = G.GpuPageRank(g, g.degrees())
    for args in ((64, 0, 0, 0), (64, 100, 1000, 16385), (64, 100, 64, 16), (64, 100, 10, 16)):
        assert pr.ctx.lib.ggpu_test_pr_tail_geometry(pr._h, *args) != 0, args
