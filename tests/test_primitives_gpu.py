"""The level-loop primitives on their own, at the boundary sizes of the
reference's primitive tests (tests/test_main.cpp:355-460): the in-place
exclusive scan the edge-balanced top-down runs over frontier degrees, and the
host loop's frontier conversions (depths -> frontier bitmap -> queue with
degrees), through the test hooks of include/ggpu_hip_testing.h.

Expected values come from numpy on the same seeded inputs: the scan is an
exclusive prefix sum in wrapping uint32 arithmetic; the bitmap has bit v
exactly where dist[v] == level (none past n) and the queue holds exactly those
vertices, each with its degree.
"""
import ctypes as C

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G

pytestmark = pytest.mark.gpu

# empty, one element, a wave, a 256-thread block, a 4096-value tile, several
# tiles, the reference's two- and three-level scan sizes, a large odd size
PRIM_SIZES = [0, 1, 2, 31, 32, 33, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 4095, 4096, 4097, 100003,
              1048575, 1048576, 1048577, 3000017]


def _u32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _scan(ctx, v, max_blocks):
    n = len(v)
    out = np.full(max(n, 1), 0xDEADBEEF, np.uint32)
    ctx.check(ctx.lib.ggpu_test_scan(ctx._h, n, _u32(v) if n else None, _u32(out), max_blocks))
    return out[:n]


def _exclusive(v):
    want = np.zeros(len(v), np.uint32)
    if len(v) > 1:
        want[1:] = np.cumsum(v[:-1], dtype=np.uint32)
    return want


@pytest.mark.parametrize("max_blocks", [0, 3])
def test_exclusive_scan_at_boundary_sizes(max_blocks):
    """max_blocks 3: each block loops over many tiles, and the carry of a tile
    sums tiles scanned by other blocks."""
    ctx = G.default_context()
    for n in PRIM_SIZES:
        v = np.random.default_rng(200 + n).integers(0, 1000, n, dtype=np.uint32)
        assert np.array_equal(_scan(ctx, v, max_blocks), _exclusive(v)), n


def test_exclusive_scan_wraps_like_uint32():
    """Values over the whole 32-bit range: the running sums wrap on both sides."""
    ctx = G.default_context()
    for n in (1, 4097, 100003):
        v = np.random.default_rng(300 + n).integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(_scan(ctx, v, 0), _exclusive(v)), n


def _random_graph(n, seed):
    rng = np.random.default_rng(seed)
    m = 2 * n
    edges = np.stack([rng.integers(0, n, m), rng.integers(0, n, m)], axis=1)
    return G.Graph.from_edges(n, edges)


@pytest.mark.parametrize("n", [1, 31, 32, 33, 63, 64, 65, 1000, 100003])
def test_frontier_bitmap_and_queue_round_trip(n):
    g = _random_graph(n, 400 + n)
    bfs = G.GpuBfs(g)
    rng = np.random.default_rng(500 + n)
    level = 5
    # about a third at `level`, the rest at neighbouring depths or unreached
    dist = rng.choice(np.array([level, level - 1, level + 1, G.INF], dtype=np.uint32), n, p=[0.34, 0.22, 0.22, 0.22])
    dist = dist.astype(np.uint32)
    members = np.flatnonzero(dist == level)
    nw = (n + 63) // 64
    for tg in (64, 256, 1024):
        bitmap = np.zeros(nw, np.uint64)
        queue = np.zeros(n, np.uint32)
        deg = np.zeros(n, np.uint32)
        count = C.c_uint32()
        bfs.ctx.check(bfs.ctx.lib.ggpu_test_frontier_convert(
            bfs._h, _u32(dist), level, tg, bitmap.ctypes.data_as(C.POINTER(C.c_uint64)), _u32(queue), _u32(deg),
            C.byref(count)))
        bits = np.unpackbits(bitmap.view(np.uint8), bitorder="little")
        assert np.array_equal(np.flatnonzero(bits[:n]), members), (n, tg)
        assert not bits[n:].any(), (n, tg)  # no bits past n
        c = count.value
        assert c == len(members), (n, tg)
        order = np.argsort(queue[:c], kind="stable")
        assert np.array_equal(queue[:c][order], members), (n, tg)
        assert np.array_equal(deg[:c][order], g.degrees()[members]), (n, tg)
