"""Device-driven BFS level loops (persistent kernel, hipGraph of K levels).

The same contract as the host loop (tests/test_main.cpp:467-623 of the
reference applied to GpuBfs): depths equal the oracle's serial BFS bit for
bit; the per-level (direction, frontier) sequence of gpu_do equals the
oracle's restatement of cpu_do in every forced mode; top-down levels report
exactly the frontier's edges, bottom-up levels at least the oracle's examined
count. The graph loop replays one captured graph across sources and runs, so
every replay is checked against the oracle too.
"""
import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O
from tests.conftest import SUITE, hub, sample_sources

pytestmark = pytest.mark.gpu

MODES = [("default", 15.0, 18.0), ("force-bu", 1e9, 1e9), ("never-bu", 1e-9, 18.0), ("flip-flop", 2.0, 2.0)]
LOOPS = [("persistent", dict(loop=G.LOOP_PERSISTENT)), ("graph-k1", dict(loop=G.LOOP_GRAPH, graph_levels=1)),
         ("graph-k3", dict(loop=G.LOOP_GRAPH, graph_levels=3)), ("graph-k8", dict(loop=G.LOOP_GRAPH, graph_levels=8))]
_CACHE = {}


def _bfs(name):
    if name not in _CACHE:
        g = G.Graph.named(name).symmetrize()
        _CACHE[name] = (g, G.GpuBfs(g))
    return _CACHE[name]


def _sources(g):
    return sorted(set(sample_sources(g.n) + [hub(g)]))


@pytest.mark.parametrize("name", SUITE)
@pytest.mark.parametrize("loop", [x[0] for x in LOOPS])
def test_device_loops_match_oracle_and_decisions(name, loop):
    kw = dict(LOOPS)[loop]
    g, bfs = _bfs(name)
    rp, ci = g.row_ptr, g.col_idx
    for src in _sources(g):
        ref = O.bfs_serial(rp, ci, src)
        reached = int((ref != O.INF).sum())
        for variant in ("gpu_edge", "gpu_binned"):
            r = bfs.run(variant, src, G.bfs_options(**kw))
            assert np.array_equal(r.dist, ref), f"{name} {loop} {variant} src {src}"
            assert all(L.direction == b"T" for L in r.trace)
            assert sum(L.frontier for L in r.trace) == reached
        for what, a, b in MODES:
            _, levels = O.bfs_do_levels(rp, ci, src, a, b)
            r = bfs.run("gpu_do", src, G.bfs_options(alpha=a, beta=b, **kw))
            assert np.array_equal(r.dist, ref), f"{name} {loop} src {src} {what}"
            got = [(L.direction.decode(), L.frontier) for L in r.trace]
            want = [(d, f) for d, f, _ in levels]
            assert got == want, f"{name} {loop} src {src} {what}: {got[:6]} vs {want[:6]}"
            for L, (d, _, e) in zip(r.trace, levels):
                assert (L.edges == e) if d == "T" else (L.edges >= e), (name, loop, src, what, L.level)
            assert r.levels == len(levels)


@pytest.mark.parametrize("name", ["rmat12", "grid100x100", "star100k", "path1000", "two_components"])
@pytest.mark.parametrize("loop", ["persistent", "graph-k3"])
def test_device_loops_any_action_sequence(name, loop):
    kw = dict(LOOPS)[loop]
    g, bfs = _bfs(name)
    rng = np.random.default_rng(17)
    for src in _sources(g)[:3]:
        ref = O.bfs_serial(g.row_ptr, g.col_idx, src)
        for trial in range(5):
            acts = [int(x) for x in rng.integers(0, 3, size=int(rng.integers(1, 9)))]
            r = bfs.run("gpu_do", src, G.bfs_options(policy=G.POLICY_SCRIPTED, actions=acts, **kw))
            assert np.array_equal(r.dist, ref), f"{name} {loop} src {src} actions {acts}"
            assert [L.action for L in r.trace] == [acts[i % len(acts)] for i in range(len(r.trace))]


def test_device_loops_learned_policy_and_switching():
    """Random MLP weights; loops alternate on one object (state reuse), results exact."""
    g, bfs = _bfs("rmat12")
    ctx = bfs.ctx
    p = G.mlp_init(3)
    G.policy_set(p, ctx, G.POLICY_SINGLE)
    try:
        for src in _sources(g):
            ref = O.bfs_serial(g.row_ptr, g.col_idx, src)
            for kw in (dict(loop=G.LOOP_PERSISTENT), dict(loop=G.LOOP_GRAPH, graph_levels=2),
                       dict(loop=G.LOOP_HOST)):
                r = bfs.run("gpu_do", src, G.bfs_options(policy=G.POLICY_LEARNED, **kw))
                assert np.array_equal(r.dist, ref)
                r = bfs.run("gpu_do", src, G.bfs_options(**kw), trace=False)
                assert np.array_equal(r.dist, ref)
    finally:
        G.policy_clear(ctx, G.POLICY_SINGLE)


def test_device_loops_larger_graph():
    g = G.Graph.named("rmat16_c1")
    bfs = G.GpuBfs(g)
    for src in g.pick_sources(4, 1):
        ref = O.bfs_serial(g.row_ptr, g.col_idx, int(src))
        _, levels = O.bfs_do_levels(g.row_ptr, g.col_idx, int(src))
        for kw in (dict(loop=G.LOOP_PERSISTENT), dict(loop=G.LOOP_GRAPH, graph_levels=4)):
            r = bfs.run("gpu_do", int(src), G.bfs_options(**kw))
            assert np.array_equal(r.dist, ref)
            assert [(L.direction.decode(), L.frontier) for L in r.trace] == [(d, f) for d, f, _ in levels]


def test_device_loop_errors():
    g, bfs = _bfs("path1000")
    with pytest.raises(ValueError):
        bfs.run("hybrid", 0, G.bfs_options(loop=G.LOOP_PERSISTENT))
    with pytest.raises(ValueError):
        bfs.run("gpu_do", 0, G.bfs_options(loop=G.LOOP_GRAPH, graph_levels=0))
    with pytest.raises(ValueError):
        bfs.run("gpu_do", 0, G.bfs_options(loop=7))
