"""The TILED PageRank work plan on the CPU (csrc/host/pr_tiling.cpp).

The planner's output is executed item by item in FP64 by the library's own
emulator (the same decomposition the GPU takes: runs of short segments,
chunks of long ones combined per task, the cold pseudo-tile, the merge over
tiles, the untiled rows) and compared with the direct in-list sums; the
emulator also checks the plan's invariants (every item in exactly one task,
every task on exactly one workgroup, every partial written at most once,
every row stored exactly once, item and slot limits). Small tile sizes make
the suite graphs exercise many tiles, dense and cold ones, chunked segments
and split tasks.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import graph_analytics_applesilicon_amd as G


def _emulate(g, contrib, params):
    lib = C.CDLL(G.HOST_LIB)
    f = lib.ggpu_test_pr_tiled_emulate
    u32p = C.POINTER(C.c_uint32)
    f.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, u32p, C.POINTER(C.c_float), C.POINTER(C.c_double),
                  C.POINTER(C.c_uint64)]
    lib.ggpu_host_last_error.restype = C.c_char_p
    rp = np.ascontiguousarray(g.row_ptr, np.uint32)
    ci = np.ascontiguousarray(g.col_idx if len(g.col_idx) else np.zeros(1, np.uint32), np.uint32)
    prm = np.ascontiguousarray(params, np.uint32)
    c = np.ascontiguousarray(contrib, np.float32)
    sums = np.zeros(max(g.n, 1), np.float64)
    stats = np.zeros(8, np.uint64)
    rc = f(g.n, g.m, rp.ctypes.data_as(u32p), ci.ctypes.data_as(u32p), prm.ctypes.data_as(u32p),
           c.ctypes.data_as(C.POINTER(C.c_float)), sums.ctypes.data_as(C.POINTER(C.c_double)),
           stats.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert rc == 0, lib.ggpu_host_last_error().decode()
    return sums[: g.n], stats


def _direct(g, contrib):
    A = sp.csr_matrix((np.ones(g.m), g.col_idx.astype(np.int64), g.row_ptr.astype(np.int64)), shape=(g.n, g.n))
    return A @ contrib.astype(np.float64)


# min in-degree, tile, arcs per item, segments per run, slots, dense min arcs, arcs per task, items per task, groups
SMALL = [(8, 256, 64, 8, 16, 500, 2048, 32, 7), (1, 64, 64, 4, 64, 1, 256, 8, 3), (100, 4096, 1024, 64, 512, 0,
                                                                                   131072, 2048, 256)]


@pytest.mark.parametrize("name", ["rmat12", "rmat12_directed", "star100k", "grid100x100", "complete200",
                                  "two_components", "path1000", "empty1000"])
@pytest.mark.parametrize("params", SMALL, ids=["small", "tiny", "default-like"])
def test_plan_covers_every_arc_once(name, params):
    g0 = G.Graph.named(name)
    t = g0.transpose()
    rng = np.random.default_rng(3)
    contrib = rng.random(max(t.n, 1)).astype(np.float32)[: t.n]
    sums, stats = _emulate(t, contrib, params)
    want = _direct(t, contrib)
    assert np.allclose(sums, want, rtol=1e-12, atol=1e-12)
    assert stats[2] + stats[3] + stats[4] == t.m  # dense + cold + untiled arcs


def test_plan_on_a_degree_ordered_rmat_graph():
    """rmat16 (the 65,536-vertex workload graph) renumbered by descending degree,
    as the TILED layout runs it: hub rows tiled, long segments chunked."""
    g = G.Graph.named("rmat16_c1")
    deg = g.degrees()
    order = np.argsort(-deg.astype(np.int64), kind="stable")
    perm = np.empty(g.n, np.uint32)
    perm[order] = np.arange(g.n, dtype=np.uint32)
    h = g.permute(perm)
    contrib = np.random.default_rng(5).random(h.n).astype(np.float32)
    params = (64, 2048, 256, 64, 64, 4096, 32768, 512, 32)
    sums, stats = _emulate(h, contrib, params)
    assert np.allclose(sums, _direct(h, contrib), rtol=1e-12, atol=1e-12)
    R, K, dense, cold, untiled, items, tasks, segs = map(int, stats)
    assert R > 0 and K > 1 and dense > untiled and cold > 0 and segs > 0 and tasks > 32
