"""Batched multi-source BFS parity: every source's depths equal the oracle's
single-source BFS from it, bit for bit, whatever the batch size, the
duplicates in the batch, the per-level action sequence and the state left
by earlier batches on the same object.

The batched BFS has no counterpart in the reference; its oracle is the
reference's serial BFS (src/cpu/bfs_cpu.cpp:10-29, restated in
oracle/oracle.c), run once per source.
"""
import ctypes

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O
from tests.conftest import SUITE, hub, sample_sources

pytestmark = pytest.mark.gpu

_GRAPHS = {}


def _graph(name):
    if name not in _GRAPHS:
        g = G.Graph.named(name).symmetrize()
        _GRAPHS[name] = (g, G.GpuBfs(g))
    return _GRAPHS[name]


def _batch(g, k, seed=7):
    """k sources: the sample sources and the hub first, then seeded random ones
    (with repeats when k exceeds n or by chance)."""
    base = list(dict.fromkeys(sample_sources(g.n) + [hub(g)]))
    rng = np.random.default_rng(seed)
    extra = rng.integers(0, g.n, size=max(0, k - len(base))).tolist()
    return np.array((base + extra)[:k], dtype=np.uint32)


def _check(g, bfs, src, opts, tag):
    dist, res, trace = bfs.run_batch(src, opts, trace=True)
    assert dist.shape == (len(src), g.n)
    depth = 0
    for i, s in enumerate(src):
        ref = O.bfs_serial(g.row_ptr, g.col_idx, int(s))
        assert np.array_equal(dist[i], ref), f"{tag}: source #{i} ({s})"
        fin = ref[ref != O.INF]
        depth = max(depth, int(fin.max()) + 1)
    assert res.levels == depth, tag
    assert sum(res.level_actions) == depth, tag
    assert trace[0].frontier == len(set(int(s) for s in src)), tag
    return res, trace


MODES = {
    "heuristic": dict(),
    "all-bottom-up": dict(alpha=1e12, beta=1e12),
    "all-top-down": dict(alpha=1e-9),
}
# traversal numbering: renumbered by degree (depths are mapped back to the
# caller's ids), the caller's own, and the default choice between them
ORDERS = {"degree": G.BATCH_ORDER_DEGREE, "given": G.BATCH_ORDER_GIVEN, "auto": G.BATCH_ORDER_AUTO}


@pytest.mark.parametrize("order", list(ORDERS))
@pytest.mark.parametrize("name", SUITE)
@pytest.mark.parametrize("k", [1, 5, 64, 100, 256])
def test_batched_matches_serial_per_source(name, k, order):
    g, bfs = _graph(name)
    src = _batch(g, k)
    for mode, kw in MODES.items():
        _check(g, bfs, src, G.bfs_options(batch_order=ORDERS[order], **kw), f"{name} k={k} {mode} {order}")


def _route(bfs, min_edges):
    """Sets the routed top-down's minimum frontier arcs (0: the default); returns
    the routed levels of the last run."""
    last = ctypes.c_uint32()
    bfs.ctx.check(bfs.ctx.lib.ggpu_test_msbfs_route(bfs._h, min_edges, ctypes.byref(last)))
    return last.value


@pytest.mark.parametrize("order", list(ORDERS))
@pytest.mark.parametrize("name", SUITE)
@pytest.mark.parametrize("k", [1, 64, 100, 256])
def test_batched_routed_top_down(name, k, order):
    """Every top-down level whose frontier has at most 4096 vertices routed
    through the per-range LDS bins (queue, binned records, apply and settle),
    under the heuristic and all top-down: every source's depths bit-exact.
    Batches of more than 128 sources (three or four words per vertex) are
    never routed."""
    g, bfs = _graph(name)
    src = _batch(g, k)
    routed = 0
    try:
        _route(bfs, 1)
        for mode in ("heuristic", "all-top-down"):
            _check(g, bfs, src, G.bfs_options(batch_order=ORDERS[order], **MODES[mode]),
                   f"{name} k={k} {mode} {order} routed")
            routed += _route(bfs, 1)
    finally:
        _route(bfs, 0)
    if k > 128:
        assert routed == 0, f"{name}: a {k}-source batch was routed"
    elif g.m:
        assert routed > 0, f"{name}: no level was routed"


@pytest.mark.parametrize("name", ["rmat12", "grid100x100", "path1000", "star100k"])
def test_batched_scripted_actions(name):
    """Any per-level mix of the three actions gives the same depths."""
    g, bfs = _graph(name)
    rng = np.random.default_rng(11)
    for k in (3, 64, 130):
        src = _batch(g, k, seed=k)
        for trial in range(4):
            actions = rng.integers(0, 3, size=int(rng.integers(1, 9))).tolist()
            order = list(ORDERS)[trial % 3]
            res, trace = _check(g, bfs, src, G.bfs_options(policy=G.POLICY_SCRIPTED, actions=actions,
                                                           batch_order=ORDERS[order]),
                                f"{name} k={k} actions={actions} {order}")
            for L in trace[: len(trace)]:
                assert L.action == actions[L.level % len(actions)]


def test_batched_all_same_source_and_state_reuse():
    """Duplicates, shrinking and growing word counts on one object."""
    g, bfs = _graph("rmat12")
    ref = O.bfs_serial(g.row_ptr, g.col_idx, 17)
    for k in (256, 1, 64, 65, 2):
        dist, res, _ = bfs.run_batch(np.full(k, 17, dtype=np.uint32))
        assert all(np.array_equal(dist[i], ref) for i in range(k))
    src = _batch(g, 200)
    d1, _, _ = bfs.run_batch(src)
    d2, _, _ = bfs.run_batch(src)
    assert np.array_equal(d1, d2)
    # switching the numbering back and forth on one object rebuilds its arrays
    for order in ("given", "degree", "auto", "given", "degree"):
        d3, _, _ = bfs.run_batch(src, G.bfs_options(batch_order=ORDERS[order]))
        assert np.array_equal(d1, d3), order


def test_batched_larger_graph():
    """A 65,536-vertex RMAT graph (rmat(16, 16, .57, .19, .19), permuted): more
    vertices than the bottom-up LDS cache holds (16384 / W), so arcs to cached
    and to uncached vertices mix; big top-down levels take the unmarked push
    and the full collect, small ones the marked path; 64 and 256 sources."""
    g = G.Graph.named("rmat16_c1")
    bfs = G.GpuBfs(g)
    for k in (64, 256):
        src = np.array(g.pick_sources(k, 1), dtype=np.uint32)
        for mode, kw in MODES.items():
            for order, o in ORDERS.items():
                _check(g, bfs, src, G.bfs_options(batch_order=o, **kw), f"rmat16_c1 k={k} {mode} {order}")


def test_batched_errors():
    g, _ = _graph("rmat12")
    fresh = G.GpuBfs(g)
    with pytest.raises(ValueError):
        fresh.run_batch(np.array([], dtype=np.uint32))
    with pytest.raises(ValueError):
        fresh.run_batch(np.zeros(257, dtype=np.uint32))
    with pytest.raises(ValueError):
        fresh.run_batch(np.array([g.n], dtype=np.uint32))
    with pytest.raises(ValueError):
        fresh.run_batch(np.array([0], dtype=np.uint32), G.bfs_options(policy=G.POLICY_SCRIPTED))
    with pytest.raises(ValueError):
        fresh.run_batch(np.array([0], dtype=np.uint32), G.bfs_options(batch_order=3))
    # nothing has run on `fresh`: fetching is a state error
    out = np.zeros(g.n, dtype=np.uint32)
    rc = fresh.ctx.lib.ggpu_msbfs_fetch_dist(fresh._h, out.ctypes.data_as(G.C.POINTER(G.C.c_uint32)))
    assert rc == 5
