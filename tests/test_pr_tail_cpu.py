"""The HOT_TAIL PageRank work plan on the CPU (csrc/host/pr_tail_plan.cpp).

The planner's output is executed piece by piece in FP64 by the library's own
emulator (split rows: head sum, then the tile partials in tile order, then the
overflow partials, as the merge kernel adds them; other rows: their whole
in-list) and compared with the direct in-list sums. The emulator also checks
the plan's invariants: every tail arc of a split row in exactly one piece, each
piece inside one tile, items of one tile in tile order within their arc and
piece limits, every partial written at most once, every overflow partial
written, each split at the row's first tail arc. Small first-tail ids and tile
sizes make the suite graphs exercise many tiles, overflow pieces and items cut
by either limit, with and without a band of tiles of another size.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import graph_analytics_applesilicon_amd as G


def _emulate(g, contrib, params):
    lib = C.CDLL(G.HOST_LIB)
    f = lib.ggpu_test_pr_tail_emulate
    u32p = C.POINTER(C.c_uint32)
    f.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, u32p, C.POINTER(C.c_float), C.POINTER(C.c_double),
                  C.POINTER(C.c_uint64)]
    lib.ggpu_host_last_error.restype = C.c_char_p
    rp = np.ascontiguousarray(g.row_ptr, np.uint32)
    ci = np.ascontiguousarray(g.col_idx if len(g.col_idx) else np.zeros(1, np.uint32), np.uint32)
    c = np.ascontiguousarray(contrib, np.float32)
    sums = np.zeros(max(g.n, 1), np.float64)
    prm = np.ascontiguousarray(tuple(params) + (0, 0) * (len(params) == 5), np.uint32)
    stats = np.zeros(8, np.uint64)
    rc = f(g.n, g.m, rp.ctypes.data_as(u32p), ci.ctypes.data_as(u32p), prm.ctypes.data_as(u32p),
           c.ctypes.data_as(C.POINTER(C.c_float)), sums.ctypes.data_as(C.POINTER(C.c_double)),
           stats.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert rc == 0, lib.ggpu_host_last_error().decode()
    return sums[: g.n], stats


def _direct(g, contrib):
    A = sp.csr_matrix((np.ones(g.m), g.col_idx.astype(np.int64), g.row_ptr.astype(np.int64)), shape=(g.n, g.n))
    return A @ contrib.astype(np.float64)


def _tail_arcs(g, min_indeg, tail_begin):
    """Arcs into rows of in-degree >= min_indeg from sources >= tail_begin (rows are sorted)."""
    deg = np.diff(g.row_ptr.astype(np.int64))
    rows = np.repeat(np.arange(g.n), deg)
    return int(((deg[rows] >= max(min_indeg, 1)) & (g.col_idx >= tail_begin)).sum())


# min in-degree, first tail source, sources per tail tile, arcs per piece and item, pieces per item,
# band end, sources per band tile
SMALL = [(4, 16, 64, 64, 8, 0, 0), (1, 0, 7, 64, 1, 0, 0), (64, 512, 256, 64, 64, 0, 0), (8, 100, 1000, 512, 512, 0, 0),
         (4, 16, 300, 64, 8, 1000, 64), (1, 0, 7, 64, 1, 50, 13), (8, 100, 1000, 512, 512, 90000, 16384)]


@pytest.mark.parametrize("name", ["rmat12", "rmat12_directed", "star100k", "grid100x100", "complete200",
                                  "two_components", "path1000", "empty1000"])
@pytest.mark.parametrize("params", SMALL, ids=["small", "tiny-tiles", "hub-rows", "wide", "band", "tiny-band",
                                                "band-to-90000"])
def test_plan_covers_every_arc_once(name, params):
    g0 = G.Graph.named(name)
    t = g0.transpose()
    rng = np.random.default_rng(3)
    contrib = rng.random(max(t.n, 1)).astype(np.float32)[: t.n]
    sums, stats = _emulate(t, contrib, params)
    assert np.allclose(sums, _direct(t, contrib), rtol=1e-12, atol=1e-12)
    R, T, tail, head, pieces, items, ovf, band = map(int, stats)
    assert tail == _tail_arcs(t, params[0], params[1])
    if tail:
        assert R > 0 and T > 0 and items > 0 and pieces >= items
    if T:
        assert (band, T) == _tiles(t.n, params), (band, T)


def _tiles(n, params):
    """(band tiles, all tiles) of the geometry: the band [tail_begin, band_end) in band tiles,
    then [band_end, n) in tail tiles (without a band, [tail_begin, n) in tail tiles)."""
    tail_begin, tile, band_end, band_tile = params[1], params[2], params[5], params[6]
    if not band_tile:
        return 0, -(-(n - tail_begin) // tile)
    band = -(-(min(n, band_end) - tail_begin) // band_tile)
    return band, band + (-(-(n - band_end) // tile) if n > band_end else 0)


def test_no_tail_when_every_source_is_below_the_first_tail_id():
    g = G.Graph.named("rmat12").transpose()
    contrib = np.random.default_rng(1).random(g.n).astype(np.float32)
    sums, stats = _emulate(g, contrib, (1, g.n, 4096, 512, 512))
    assert np.allclose(sums, _direct(g, contrib), rtol=1e-12, atol=1e-12)
    assert int(stats[0]) == 0 and int(stats[2]) == 0


def test_unsorted_in_list_is_not_split():
    # row 0's in-list is unsorted: its tail is not a suffix, so it stays whole
    rp = np.array([0, 4, 6, 6, 6, 6, 6, 6, 6], np.uint32)
    ci = np.array([7, 1, 6, 2, 5, 7], np.uint32)
    g = G.Graph.from_arrays(rp, ci)
    contrib = np.arange(1, 9, dtype=np.float32)
    sums, stats = _emulate(g, contrib, (1, 4, 2, 64, 64))
    assert np.allclose(sums, _direct(g, contrib))
    assert int(stats[0]) == 1 and int(stats[2]) == 2  # only row 1 (sources 5, 7) is split


def test_plan_on_a_degree_ordered_rmat_graph():
    """rmat16 (the 65,536-vertex workload graph) renumbered by descending degree,
    as the HOT_TAIL layout runs it: split hub rows, several tiles, overflow
    pieces of hub segments longer than one item."""
    g = G.Graph.named("rmat16_c1")
    deg = g.degrees()
    order = np.argsort(-deg.astype(np.int64), kind="stable")
    perm = np.empty(g.n, np.uint32)
    perm[order] = np.arange(g.n, dtype=np.uint32)
    h = g.permute(perm)
    contrib = np.random.default_rng(5).random(h.n).astype(np.float32)
    params = (64, 4096, 4096, 512, 512)
    sums, stats = _emulate(h, contrib, params)
    assert np.allclose(sums, _direct(h, contrib), rtol=1e-12, atol=1e-12)
    R, T, tail, head, pieces, items, ovf, _ = map(int, stats)
    assert R > 0 and T == (h.n - 4096 + 4095) // 4096 and ovf > 0 and items >= T
    assert tail == _tail_arcs(h, 64, 4096) and tail + head == int(np.diff(h.row_ptr.astype(np.int64))[
        np.diff(h.row_ptr.astype(np.int64)) >= 64].sum()) - _untouched_heads(h, 64, 4096)


def _untouched_heads(g, min_indeg, tail_begin):
    """Arcs of rows of in-degree >= min_indeg that have no tail arc (not split)."""
    rp = g.row_ptr.astype(np.int64)
    deg = np.diff(rp)
    total = 0
    for v in np.nonzero(deg >= min_indeg)[0]:
        if g.col_idx[rp[v + 1] - 1] < tail_begin:
            total += int(deg[v])
    return total
