"""GPU BFS parity: every variant against the oracle, bit for bit.

Mirrors tests/test_main.cpp:467-623 of the reference: all suite graphs
(symmetrized), the sample sources plus the highest-degree vertex, several
workgroup sizes and option extremes; dist must equal the oracle exactly, the
trace must account for every reached vertex once, gpu_do must take the same
per-level (direction, frontier) sequence as the oracle's restatement of
cpu_do, and hybrid must do so at every CPU/GPU threshold.
"""
import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O
from tests.conftest import SUITE, hub, sample_sources

pytestmark = pytest.mark.gpu


def _sym(name):
    return G.Graph.named(name).symmetrize()


def _sources(g):
    return sorted(set(sample_sources(g.n) + [hub(g)]))


def _reached_stats(g, ref):
    reached = ref != O.INF
    return int(reached.sum()), int(g.degrees()[reached].astype(np.int64).sum())


@pytest.mark.parametrize("name", SUITE)
def test_gpu_edge_and_binned_match_oracle(name):
    g = _sym(name)
    bfs = G.GpuBfs(g)
    rp, ci = g.row_ptr, g.col_idx
    cfgs = [("gpu_edge", dict(ept=1)), ("gpu_edge", dict(ept=2)), ("gpu_edge", dict(ept=4)),
            ("gpu_edge", dict(ept=16)),
            ("gpu_binned", dict(small_max=16, large_min=256)),
            ("gpu_binned", dict(small_max=1 << 30, large_min=(1 << 31))),  # all small
            ("gpu_binned", dict(small_max=0, large_min=0)),                # all large
            ("gpu_binned", dict(small_max=1, large_min=1 << 30)),          # medium
            ("gpu_binned", dict(small_max=2, large_min=3))]
    for src in _sources(g):
        ref = O.bfs_serial(rp, ci, src)
        want_frontier, want_edges = _reached_stats(g, ref)
        for variant, kw in cfgs:
            for tg in (64, 256, 1024):
                r = bfs.run(variant, src, G.bfs_options(tg=tg, **kw))
                assert np.array_equal(r.dist, ref), f"{name} src {src} {variant} {kw} tg {tg}"
                assert sum(L.frontier for L in r.trace) == want_frontier
                assert sum(L.edges for L in r.trace) == want_edges


@pytest.mark.parametrize("name", SUITE)
def test_gpu_vertex_and_naive_match_oracle(name):
    """runVertex / runNaive: top-down levels until nothing is discovered; the
    trace holds one level per depth (the last one discovers nothing), each
    with the number of vertices at that depth."""
    g = _sym(name)
    bfs = G.GpuBfs(g)
    for src in _sources(g):
        ref = O.bfs_serial(g.row_ptr, g.col_idx, src)
        hist = [int(x) for x in np.bincount(ref[ref != O.INF])]
        want_frontier, want_edges = _reached_stats(g, ref)
        for variant in ("gpu_vertex", "gpu_naive"):
            for tg in (64, 256, 1024):
                r = bfs.run(variant, src, G.bfs_options(tg=tg))
                assert np.array_equal(r.dist, ref), f"{name} src {src} {variant} tg {tg}"
                assert [L.frontier for L in r.trace] == hist
                assert {L.direction for L in r.trace} == {b"T"}
                assert sum(L.edges for L in r.trace) == want_edges and sum(hist) == want_frontier


MODES = [("default", 15.0, 18.0), ("force-bu", 1e9, 1e9), ("never-bu", 1e-9, 18.0), ("flip-flop", 2.0, 2.0)]


@pytest.mark.parametrize("name", SUITE)
def test_gpu_do_matches_oracle_and_decisions(name):
    g = _sym(name)
    bfs = G.GpuBfs(g)
    rp, ci = g.row_ptr, g.col_idx
    for src in _sources(g):
        ref = O.bfs_serial(rp, ci, src)
        reached, _ = _reached_stats(g, ref)
        for what, a, b in MODES:
            _, levels = O.bfs_do_levels(rp, ci, src, a, b)
            for tg in (64, 256):
                r = bfs.run("gpu_do", src, G.bfs_options(tg=tg, alpha=a, beta=b))
                assert np.array_equal(r.dist, ref), f"{name} src {src} {what} tg {tg}"
                got = [(L.direction.decode(), L.frontier) for L in r.trace]
                want = [(dd, f) for dd, f, _ in levels]
                assert got == want, f"{name} src {src} {what}: {got[:6]} vs {want[:6]}"
                # top-down levels examine exactly the frontier's lists; bottom-up
                # scans at least the entries up to each vertex's first parent
                # (lists split into chunks are scanned in parallel)
                for L, (dd, _, e) in zip(r.trace, levels):
                    assert (L.edges == e) if dd == "T" else (L.edges >= e), (name, src, what, L.level)
                assert sum(L.frontier for L in r.trace) == reached
                if a > 1e6 and reached > 1:
                    assert any(L.direction == b"B" for L in r.trace)
                if a < 1e-6:
                    assert all(L.direction == b"T" for L in r.trace)


@pytest.mark.parametrize("name", SUITE)
def test_hybrid_matches_oracle_at_every_threshold(name):
    g = _sym(name)
    bfs = G.GpuBfs(g)
    rp, ci = g.row_ptr, g.col_idx
    degs = g.degrees()
    Ts = [0, 1, 64, 1000, 100000, 2**64 - 1]
    for src in _sources(g)[:4]:
        ref = O.bfs_serial(rp, ci, src)
        for what, a, b in [MODES[0], MODES[1], MODES[3]]:
            _, levels = O.bfs_do_levels(rp, ci, src, a, b)
            want = [(d, f) for d, f, _ in levels]
            for T in Ts:
                for threads in (1, 8):
                    r = bfs.run("hybrid", src, G.bfs_options(alpha=a, beta=b, hybrid_t=T, cpu_threads=threads))
                    assert np.array_equal(r.dist, ref), f"{name} src {src} {what} T {T} threads {threads}"
                    assert [(L.direction.decode(), L.frontier) for L in r.trace] == want
                    devs = {L.device for L in r.trace}
                    if T == 0:
                        assert devs == {b"G"}
                    if T == 2**64 - 1:
                        assert devs == {b"C"}
                    if r.trace[0].direction == b"T":
                        assert (r.trace[0].device == b"C") == (int(degs[src]) < T)


@pytest.mark.parametrize("name", ["rmat12", "grid100x100", "star100k", "two_components", "path1000"])
def test_any_action_sequence_is_bit_exact(name):
    """Direction choices only change cost: random per-level action scripts
    (edge top-down / binned top-down / bottom-up, switching formats at any
    level) must give the oracle's depths."""
    g = _sym(name)
    bfs = G.GpuBfs(g)
    rng = np.random.default_rng(7)
    for src in _sources(g)[:3]:
        ref = O.bfs_serial(g.row_ptr, g.col_idx, src)
        for trial in range(6):
            acts = [int(x) for x in rng.integers(0, 3, size=int(rng.integers(1, 9)))]
            r = bfs.run("gpu_do", src, G.bfs_options(policy=G.POLICY_SCRIPTED, actions=acts))
            assert np.array_equal(r.dist, ref), f"{name} src {src} actions {acts}"
            assert [L.action for L in r.trace] == [acts[i % len(acts)] for i in range(len(r.trace))]


def test_errors_are_reported():
    g = _sym("path1000")
    bfs = G.GpuBfs(g)
    with pytest.raises(ValueError):
        bfs.run("gpu_do", 0, G.bfs_options(tg=96))
    with pytest.raises(ValueError):
        bfs.run("gpu_edge", 0, G.bfs_options(ept=0))
    with pytest.raises(ValueError):
        bfs.run("gpu_do", g.n, G.bfs_options())
    for variant in ("gpu_vertex", "gpu_naive"):
        with pytest.raises(ValueError):
            bfs.run(variant, 0, G.bfs_options(loop=G.LOOP_PERSISTENT))
        with pytest.raises(ValueError):
            bfs.run(variant, 0, G.bfs_options(policy=G.POLICY_SCRIPTED, actions=[2]))
