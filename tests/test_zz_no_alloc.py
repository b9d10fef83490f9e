"""ggpu_bfs_run_noalloc against ggpu_bfs_run (gpu_do, persistent loop), bit for bit.

The no-allocation entry point keeps every working buffer in the library's
fixed device array; the caller brings the CSR and the distance buffer in
device memory. Here the caller's buffers come from hipMalloc through ctypes
(the test is the caller). Two calls in a row on one GPU, the second on a
smaller graph, must each give the distances and the level count of the
allocating entry point (and of the oracle's serial BFS), so the second call
must not see anything the first left in the shared array.
"""
import ctypes as C

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O
from tests.conftest import hub

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    G.hip_lib()  # loads the HIP runtime the backend links
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


class _DeviceArray:
    """A uint32 buffer in device memory owned by the caller."""

    def __init__(self, rt, count):
        self.rt, self.count = rt, count
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(count, 1) * 4) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        assert a.size == self.count
        assert self.rt.hipMemcpy(self.p, a.ctypes.data, a.nbytes, _H2D) == 0

    def download(self):
        out = np.empty(self.count, dtype=np.uint32)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        if self.p:
            self.rt.hipFree(self.p)
            self.p = C.c_void_p()


def _graphs():
    big = G.Graph.rmat(16, 16, seed=5).symmetrize()        # hubs: every vertex on the light paths
    small = G.Graph.grid(48, 64, keep=0.7, seed=3).symmetrize()
    return [("rmat16", big), ("grid48x64", small)]


def _run_noalloc(rt, g, src, opts=None):
    rp, ci, dist = _DeviceArray(rt, g.n + 1), _DeviceArray(rt, g.m), _DeviceArray(rt, g.n)
    try:
        rp.upload(g.row_ptr)
        ci.upload(g.col_idx)
        dist.upload(np.full(g.n, 0x5A5A5A5A, dtype=np.uint32))  # every entry must be overwritten
        levels = G.bfs_run_noalloc(g.n, g.m, rp.p.value, ci.p.value, src, dist.p.value, opts)
        return dist.download(), levels
    finally:
        for b in (rp, ci, dist):
            b.free()


def test_noalloc_equals_allocating_entry_point_on_two_calls():
    rt = _hip()
    persistent = G.bfs_options(loop=G.LOOP_PERSISTENT)
    for name, g in _graphs():  # the second graph is the smaller one
        bfs = G.GpuBfs(g)
        for src in sorted({0, g.n // 2, g.n - 1, hub(g)}):
            print(f"noalloc {name} n {g.n} m {g.m} source {src}", flush=True)
            want = bfs.run("gpu_do", src, persistent)
            got, levels = _run_noalloc(rt, g, src)
            assert np.array_equal(got, want.dist), f"{name} source {src}: distances differ"
            assert levels == want.levels, f"{name} source {src}: {levels} levels, want {want.levels}"
            assert np.array_equal(got, O.bfs_serial(g.row_ptr, g.col_idx, src)), f"{name} source {src}"


def test_noalloc_rejects_what_it_cannot_hold():
    rt = _hip()
    g = G.Graph.rmat(8, 4, seed=1).symmetrize()
    d = _DeviceArray(rt, 1)
    try:
        with pytest.raises(ValueError, match="outside 1"):
            G.bfs_run_noalloc(G.NOALLOC_MAX_N + 1, 0, d.p.value, d.p.value, 0, d.p.value)
        with pytest.raises(ValueError, match="outside 1"):
            G.bfs_run_noalloc(0, 0, d.p.value, d.p.value, 0, d.p.value)
        with pytest.raises(ValueError, match="source out of range"):
            G.bfs_run_noalloc(g.n, g.m, d.p.value, d.p.value, g.n, d.p.value)
        with pytest.raises(ValueError, match="scripted"):
            G.bfs_run_noalloc(g.n, g.m, d.p.value, d.p.value, 0, d.p.value,
                              G.bfs_options(policy=G.POLICY_SCRIPTED, actions=[0]))
    finally:
        d.free()
