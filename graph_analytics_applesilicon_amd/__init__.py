"""MI355X (gfx950) BFS and PageRank backend -- Python host mirror of the C-ABI.

The product is two shared libraries built in-tree by ``csrc/Makefile``:

* ``lib/libggpu_host.so`` -- the GGPUCSR1 graph format, generators, transforms
  and the CPU variants (``include/ggpu_graph.h``);
* ``lib/libggpu_hip.so``  -- the HIP backend (``include/ggpu_hip.h``).

This module binds both with ctypes and mirrors the reference's host classes:
``GpuBfs`` (src/metal/bfs_gpu.hpp:35-105) and ``GpuPageRank``
(src/metal/pagerank_gpu.hpp:46-127), with the same option names and the
same error behaviour (a bad option raises ``ValueError``; a device failure
raises ``RuntimeError``). There is no CPU fallback: constructing a GPU object
without the HIP library raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libggpu_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libggpu_hip.so")
GBENCH = os.path.join(_PKG, "bin", "gbench")

INF = 0xFFFFFFFF
OK, EINVAL, EDEVICE, EIO, ENOMEM, ESTATE, ETIMEOUT = range(7)
FLAG_DIRECTED, FLAG_SORTED, FLAG_DEDUP = 1, 2, 4

BFS_EDGE, BFS_BINNED, BFS_DO, BFS_HYBRID, BFS_VERTEX, BFS_NAIVE = 0, 1, 2, 3, 4, 5
BFS_VARIANTS = {"gpu_edge": BFS_EDGE, "gpu_binned": BFS_BINNED, "gpu_do": BFS_DO, "hybrid": BFS_HYBRID,
                "gpu_vertex": BFS_VERTEX, "gpu_naive": BFS_NAIVE}
LOOP_HOST, LOOP_GRAPH, LOOP_PERSISTENT = 0, 1, 2
POLICY_HEURISTIC, POLICY_LEARNED, POLICY_SCRIPTED = 0, 1, 2
# batched BFS traversal numbering (depths in caller ids); AUTO: DEGREE when max degree >= 8x mean
BATCH_ORDER_GIVEN, BATCH_ORDER_DEGREE, BATCH_ORDER_AUTO = 0, 1, 2
ACTION_TD_EDGE, ACTION_TD_BINNED, ACTION_BU = 0, 1, 2
PR_SCALAR, PR_VECTOR, PR_ADAPTIVE, PR_GS, PR_PUSH = 1, 2, 3, 4, 5
PR_LAYOUT_STREAM, PR_LAYOUT_BINS, PR_LAYOUT_HOT, PR_LAYOUT_TILED, PR_LAYOUT_HOT_TAIL = 0, 1, 2, 3, 4
PR_ORDER_GIVEN, PR_ORDER_DEGREE = 0, 1
PR_VARIANTS = {"gpu_pull_scalar": PR_SCALAR, "gpu_pull_vector": PR_VECTOR,
               "gpu_pull_adaptive": PR_ADAPTIVE, "gpu_pull_gs": PR_GS, "gpu_push": PR_PUSH}
MLP_PARAMS = 1347
POLICY_SINGLE, POLICY_BATCHED = 0, 1


# ------------------------------------------------------------------ structs
class Level(C.Structure):
    """ggpu_level (include/ggpu_types.h) = util::BfsLevel + policy inputs."""
    _fields_ = [("level", C.c_uint32), ("action", C.c_uint32), ("frontier", C.c_uint64),
                ("edges", C.c_uint64), ("frontier_edges", C.c_uint64), ("unvisited_edges", C.c_uint64),
                ("direction", C.c_char), ("device", C.c_char), ("pad", C.c_char * 6),
                ("time_ms", C.c_double)]


class CpuBfsOpts(C.Structure):
    _fields_ = [("threads", C.c_int), ("schedule", C.c_int), ("alpha", C.c_double), ("beta", C.c_double)]


class BfsOpts(C.Structure):
    """ggpu_bfs_opts = gpu::GpuBfsOptions (src/metal/bfs_gpu.hpp:23-33) + loop/policy."""
    _fields_ = [("tg", C.c_uint32), ("small_max", C.c_uint32), ("large_min", C.c_uint32),
                ("ept", C.c_uint32), ("alpha", C.c_double), ("beta", C.c_double),
                ("hybrid_t", C.c_uint64), ("cpu_threads", C.c_int), ("loop", C.c_uint32),
                ("policy", C.c_uint32), ("actions", C.POINTER(C.c_uint8)), ("n_actions", C.c_uint32),
                ("graph_levels", C.c_uint32), ("time_levels", C.c_uint32), ("batch_order", C.c_uint32)]


class MsbfsResult(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("pad", C.c_uint32), ("gpu_ms", C.c_double),
                ("level_actions", C.c_uint64 * 3), ("setup_ms", C.c_double), ("levels_ms", C.c_double),
                ("finish_ms", C.c_double)]


class PrOpts(C.Structure):
    """ggpu_pr_opts = gpu::GpuPrOptions (src/metal/pagerank_gpu.hpp:27-36) + medium_lanes."""
    _fields_ = [("damping", C.c_float), ("iters", C.c_int), ("tol", C.c_double), ("tol_rel", C.c_double),
                ("tg", C.c_uint32), ("small_max", C.c_uint32), ("large_min", C.c_uint32),
                ("in_place", C.c_int), ("medium_lanes", C.c_uint32), ("layout", C.c_uint32),
                ("order", C.c_uint32)]


class PrResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("floor_stop", C.c_int), ("last_delta", C.c_double),
                ("last_max_rel", C.c_double), ("gpu_ms", C.c_double), ("dangling", C.c_double),
                ("sum", C.c_double)]


# ------------------------------------------------------------------ loading
_host = None
_hip = None
_u32p = C.POINTER(C.c_uint32)


def _ptr(a: np.ndarray, ctype=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(ctype))


def host_lib() -> C.CDLL:
    """libggpu_host.so (graph format, generators, CPU variants)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise ImportError(f"{HOST_LIB} is not built: run __graft_entry__.build() or make -C csrc")
        lib = C.CDLL(HOST_LIB)
        vp = C.c_void_p
        lib.ggpu_host_last_error.restype = C.c_char_p
        lib.ggpu_graph_n.restype = C.c_uint64
        lib.ggpu_graph_n.argtypes = [vp]
        lib.ggpu_graph_m.restype = C.c_uint64
        lib.ggpu_graph_m.argtypes = [vp]
        lib.ggpu_graph_flags.restype = C.c_uint32
        lib.ggpu_graph_flags.argtypes = [vp]
        lib.ggpu_graph_row_ptr.restype = _u32p
        lib.ggpu_graph_row_ptr.argtypes = [vp]
        lib.ggpu_graph_col_idx.restype = _u32p
        lib.ggpu_graph_col_idx.argtypes = [vp]
        lib.ggpu_graph_free.argtypes = [vp]
        lib.ggpu_graph_free.restype = None
        lib.ggpu_graph_equal.argtypes = [vp, vp]
        lib.ggpu_gen_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.c_uint64, C.POINTER(vp)]
        lib.ggpu_gen_grid.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, C.POINTER(vp)]
        lib.ggpu_random_perm.argtypes = [C.c_uint64, C.c_uint64, _u32p]
        lib.ggpu_order_perm.argtypes = [vp, C.c_char_p, C.c_uint64, _u32p]
        lib.ggpu_graph_from_snap.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp), C.POINTER(SnapStats)]
        lib.ggpu_graph_stats_compute.argtypes = [vp, C.POINTER(GraphStats)]
        lib.ggpu_degree_histogram.argtypes = [vp, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        lib.ggpu_graph_locality.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.ggpu_pick_sources.argtypes = [vp, C.c_uint32, C.c_uint64, C.c_int, _u32p, _u32p]
        lib.ggpu_graph_from_arrays.argtypes = [C.c_uint64, C.c_uint64, _u32p, _u32p, C.c_uint32, C.POINTER(vp)]
        lib.ggpu_graph_from_edges.argtypes = [C.c_uint64, C.c_uint64, _u32p, _u32p, C.c_int, C.POINTER(vp)]
        lib.ggpu_graph_largest_component.argtypes = [vp, C.POINTER(vp), _u32p, C.POINTER(C.c_uint64)]
        lib.ggpu_cpu_bfs.argtypes = [vp, C.c_char_p, C.c_uint32, C.POINTER(CpuBfsOpts), _u32p,
                                     C.POINTER(Level), C.c_size_t, C.POINTER(C.c_size_t)]
        lib.ggpu_cpu_pagerank.argtypes = [vp, _u32p, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                          C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        lib.ggpu_check_ranks.argtypes = [C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_float),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.ggpu_check_ranks.restype = None
        _host = lib
    return _host


def hip_lib() -> C.CDLL:
    """libggpu_hip.so (the HIP backend). Raises ImportError if it is not built."""
    global _hip
    if _hip is None:
        host_lib()
        if not os.path.exists(HIP_LIB):
            raise ImportError(f"{HIP_LIB} is not built: the GPU path has no fallback")
        lib = C.CDLL(HIP_LIB)
        vp = C.c_void_p
        lib.ggpu_last_error.restype = C.c_char_p
        lib.ggpu_last_error.argtypes = [vp]
        lib.ggpu_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        lib.ggpu_ctx_destroy.argtypes = [vp]
        lib.ggpu_ctx_destroy.restype = None
        lib.ggpu_ctx_sync.argtypes = [vp]
        lib.ggpu_device_info.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
        lib.ggpu_bfs_opts_default.argtypes = [C.POINTER(BfsOpts)]
        lib.ggpu_bfs_opts_default.restype = None
        lib.ggpu_bfs_create.argtypes = [vp, C.c_uint64, C.c_uint64, _u32p, _u32p, C.POINTER(vp)]
        lib.ggpu_bfs_destroy.argtypes = [vp]
        lib.ggpu_bfs_destroy.restype = None
        lib.ggpu_bfs_run.argtypes = [vp, C.c_int, C.c_uint32, C.POINTER(BfsOpts), C.POINTER(Level), C.c_size_t,
                                     C.POINTER(C.c_size_t)]
        lib.ggpu_bfs_fetch_dist.argtypes = [vp, _u32p]
        lib.ggpu_bfs_last_gpu_ms.argtypes = [vp]
        lib.ggpu_bfs_last_gpu_ms.restype = C.c_double
        lib.ggpu_bfs_last_levels.argtypes = [vp]
        lib.ggpu_bfs_last_levels.restype = C.c_uint32
        for name, args in _optional_hip_symbols().items():
            fn = getattr(lib, name, None)
            if fn is not None:
                fn.argtypes = args[0]
                fn.restype = args[1]
        _hip = lib
    return _hip


def _optional_hip_symbols():
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    return {
        "ggpu_msbfs_run": ([vp, _u32p, C.c_uint32, C.POINTER(BfsOpts), C.POINTER(MsbfsResult), C.POINTER(Level),
                            C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
        "ggpu_msbfs_fetch_dist": ([vp, _u32p], C.c_int),
        "ggpu_test_scan": ([vp, C.c_uint32, _u32p, _u32p, C.c_uint32], C.c_int),
        "ggpu_test_frontier_convert": ([vp, _u32p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), _u32p, _u32p,
                                        C.POINTER(C.c_uint32)], C.c_int),
        "ggpu_test_msbfs_route": ([vp, C.c_uint64, C.POINTER(C.c_uint32)], C.c_int),
        "ggpu_test_pr_tail_geometry": ([vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32], C.c_int),
        "ggpu_test_pr_tail_info": ([vp, C.c_uint32, C.POINTER(C.c_uint64)], C.c_int),
        "ggpu_bfs_run_noalloc": ([vp, C.c_uint64, C.c_uint64, vp, vp, C.c_uint32, C.POINTER(BfsOpts), vp,
                                  C.POINTER(C.c_uint32)], C.c_int),
        "ggpu_pr_opts_default": ([C.POINTER(PrOpts)], None),
        "ggpu_pr_create": ([vp, C.c_uint64, C.c_uint64, _u32p, _u32p, _u32p, C.POINTER(vp)], C.c_int),
        "ggpu_pr_destroy": ([vp], None),
        "ggpu_pr_prepare_adaptive": ([vp, C.c_uint32, C.c_uint32], C.c_int),
        "ggpu_pr_prepare_gs": ([vp, C.c_uint32], C.c_int),
        "ggpu_pr_gs_info": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
        "ggpu_pr_run": ([vp, C.c_int, C.POINTER(PrOpts), C.POINTER(PrResult)], C.c_int),
        "ggpu_pr_fetch_ranks": ([vp, fp], C.c_int),
        "ggpu_pr_autotune": ([vp, C.c_int, C.POINTER(PrOpts)], C.c_int),
        "ggpu_ppr_run": ([vp, _u32p, C.c_uint32, C.POINTER(PrOpts), C.POINTER(PrResult)], C.c_int),
        "ggpu_ppr_fetch": ([vp, fp], C.c_int),
        "ggpu_ppr_autotune": ([vp, _u32p, C.c_uint32, C.c_int, C.POINTER(PrOpts)], C.c_int),
        "ggpu_policy_set": ([vp, C.c_int, fp, C.c_uint32], C.c_int),
        "ggpu_policy_clear": ([vp, C.c_int], C.c_int),
        "ggpu_policy_get": ([vp, C.c_int, fp, C.c_uint32, C.POINTER(C.c_int)], C.c_int),
        "ggpu_policy_features": ([C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, fp], None),
        "ggpu_mlp_forward": ([vp, fp, fp, C.c_uint64, fp], C.c_int),
        "ggpu_mlp_loss_grad": ([vp, fp, fp, fp, C.POINTER(C.c_uint8), C.c_uint64, fp, fp], C.c_int),
        "ggpu_mlp_train": ([vp, fp, fp, fp, C.POINTER(C.c_uint8), C.c_uint64, C.c_uint32, C.c_float, fp],
                           C.c_int),
        "ggpu_mlp_train_sgd": ([vp, fp, fp, fp, C.POINTER(C.c_uint8), C.c_uint64, C.c_uint32, C.c_float, C.c_float,
                                fp], C.c_int),
        "ggpu_bfs_record_levels": ([vp, C.c_uint32, C.POINTER(BfsOpts), fp, fp, C.POINTER(Level), C.c_size_t,
                                    C.POINTER(C.c_size_t)], C.c_int),
        "ggpu_msbfs_record_levels": ([vp, _u32p, C.c_uint32, C.POINTER(BfsOpts), fp, fp, C.POINTER(Level),
                                      C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
    }


def _host_check(rc: int):
    if rc != OK:
        msg = host_lib().ggpu_host_last_error().decode()
        if rc in (EINVAL,):
            raise ValueError(msg)
        if rc == EIO:
            raise OSError(msg)
        raise RuntimeError(msg)


# ------------------------------------------------------------------ graphs
class Graph:
    """A host CSR graph (graph::CSR, src/graph/csr.hpp:29-45) owned by libggpu_host."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def __del__(self):
        if getattr(self, "_h", None) and _host is not None:
            _host.ggpu_graph_free(self._h)
            self._h = None

    @staticmethod
    def _make(fn, *args) -> "Graph":
        out = C.c_void_p()
        _host_check(fn(*args, C.byref(out)))
        return Graph(out)

    # constructors
    @staticmethod
    def named(name: str) -> "Graph":
        """gen::suite() names plus the workload graphs rmat16_c1, rmat22_c2, grid_c3, rmat23_c5."""
        return Graph._make(host_lib().ggpu_gen_named, name.encode())

    @staticmethod
    def rmat(scale: int, edgefactor: int, a=0.57, b=0.19, c=0.19, directed=False, seed=1) -> "Graph":
        return Graph._make(host_lib().ggpu_gen_rmat, scale, edgefactor, a, b, c, int(directed), seed)

    @staticmethod
    def grid(rows: int, cols: int, keep: float = 1.0, seed: int = 0) -> "Graph":
        return Graph._make(host_lib().ggpu_gen_grid, rows, cols, keep, seed)

    @staticmethod
    def load(path: str) -> "Graph":
        return Graph._make(host_lib().ggpu_graph_load, path.encode())

    @staticmethod
    def from_snap(path: str, directed: bool):
        """graph::convertSnap (src/graph/snap.cpp:85-148) without the file:
        (Graph, SnapStats) from a plain or gzip SNAP edge list."""
        out = C.c_void_p()
        st = SnapStats()
        _host_check(host_lib().ggpu_graph_from_snap(str(path).encode(), int(directed), C.byref(out), C.byref(st)))
        return Graph(out), st

    def stats(self) -> "GraphStats":
        """The figures gstats reports (src/graph/stats.cpp:40-120)."""
        st = GraphStats()
        _host_check(host_lib().ggpu_graph_stats_compute(self._h, C.byref(st)))
        return st

    def degree_histogram(self, kind: int = 0) -> np.ndarray:
        """log2 buckets (0: degree 0, k: [2^(k-1), 2^k)) of stored (0), in (1)
        or symmetrized (2) degrees."""
        h = (C.c_uint64 * 34)()
        k = C.c_uint32()
        _host_check(host_lib().ggpu_degree_histogram(self._h, kind, h, C.byref(k)))
        return np.array(h[: k.value], dtype=np.uint64)

    @staticmethod
    def from_arrays(row_ptr, col_idx, flags=FLAG_SORTED | FLAG_DEDUP) -> "Graph":
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        ci = np.ascontiguousarray(col_idx, dtype=np.uint32)
        return Graph._make(host_lib().ggpu_graph_from_arrays, len(rp) - 1, len(ci), _ptr(rp), _ptr(ci), flags)

    @staticmethod
    def from_edges(n: int, edges: Sequence, directed=False) -> "Graph":
        e = np.asarray(edges, dtype=np.uint32).reshape(-1, 2)
        src = np.ascontiguousarray(e[:, 0])
        dst = np.ascontiguousarray(e[:, 1])
        return Graph._make(host_lib().ggpu_graph_from_edges, n, len(e), _ptr(src), _ptr(dst), int(directed))

    # properties
    @property
    def n(self) -> int:
        return int(host_lib().ggpu_graph_n(self._h))

    @property
    def m(self) -> int:
        return int(host_lib().ggpu_graph_m(self._h))

    @property
    def flags(self) -> int:
        return int(host_lib().ggpu_graph_flags(self._h))

    @property
    def directed(self) -> bool:
        return bool(self.flags & FLAG_DIRECTED)

    @property
    def row_ptr(self) -> np.ndarray:
        p = host_lib().ggpu_graph_row_ptr(self._h)
        return np.ctypeslib.as_array(p, shape=(self.n + 1,)).copy()

    @property
    def col_idx(self) -> np.ndarray:
        if self.m == 0:
            return np.zeros(0, dtype=np.uint32)
        p = host_lib().ggpu_graph_col_idx(self._h)
        return np.ctypeslib.as_array(p, shape=(self.m,)).copy()

    def raw_pointers(self):
        lib = host_lib()
        return lib.ggpu_graph_row_ptr(self._h), lib.ggpu_graph_col_idx(self._h)

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr.astype(np.int64)).astype(np.uint32)

    # transforms
    def save(self, path: str):
        _host_check(host_lib().ggpu_graph_save(self._h, path.encode()))

    def symmetrize(self) -> "Graph":
        return Graph._make(host_lib().ggpu_graph_symmetrize, self._h)

    def transpose(self) -> "Graph":
        return Graph._make(host_lib().ggpu_graph_transpose, self._h)

    def permute(self, perm) -> "Graph":
        p = np.ascontiguousarray(perm, dtype=np.uint32)
        return Graph._make(host_lib().ggpu_graph_permute, self._h, _ptr(p))

    def order_perm(self, ordering: str, seed: int = 1) -> np.ndarray:
        """reorder::compute (src/graph/reorder.cpp:97-103) on this (symmetric)
        graph: perm[old] = new for "original" | "random" | "degree" | "rcm"."""
        p = np.zeros(max(self.n, 1), dtype=np.uint32)
        _host_check(host_lib().ggpu_order_perm(self._h, ordering.encode(), seed, _ptr(p)))
        return p[: self.n]

    def locality(self):
        """(mean |u - v|, share of arcs within one 64-vertex block)."""
        g, s = C.c_double(), C.c_double()
        _host_check(host_lib().ggpu_graph_locality(self._h, C.byref(g), C.byref(s)))
        return g.value, s.value

    def largest_component(self):
        out = C.c_void_p()
        old = np.zeros(max(self.n, 1), dtype=np.uint32)
        k = C.c_uint64()
        _host_check(host_lib().ggpu_graph_largest_component(self._h, C.byref(out), _ptr(old), C.byref(k)))
        return Graph(out), old[: k.value].copy()

    def validate(self) -> str:
        buf = C.create_string_buffer(512)
        host_lib().ggpu_graph_validate(self._h, buf, 512)
        return buf.value.decode()

    def equal(self, other: "Graph") -> bool:
        return bool(host_lib().ggpu_graph_equal(self._h, other._h))

    def pick_sources(self, count: int = 64, seed: int = 1, any_component: bool = False) -> np.ndarray:
        out = np.zeros(max(count, 1), dtype=np.uint32)
        k = C.c_uint32()
        _host_check(host_lib().ggpu_pick_sources(self._h, count, seed, int(any_component), _ptr(out), C.byref(k)))
        return out[: k.value].copy()

    # CPU variants (gbench cpu_serial / cpu_omp / cpu_do)
    def cpu_bfs(self, source: int, variant="cpu_serial", threads=0, schedule=1, alpha=15.0, beta=18.0,
                trace=False):
        o = CpuBfsOpts(threads, schedule, alpha, beta)
        dist = np.zeros(max(self.n, 1), dtype=np.uint32)
        cap = self.n + 2 if trace else 0
        tr = (Level * max(cap, 1))()
        ln = C.c_size_t()
        _host_check(host_lib().ggpu_cpu_bfs(self._h, variant.encode(), source, C.byref(o), _ptr(dist),
                                            tr if trace else None, cap, C.byref(ln)))
        levels = [tr[i] for i in range(min(ln.value, cap))] if trace else None
        return dist[: self.n], levels

    def cpu_pagerank(self, outdeg, damping=0.85, iters=20, tol=0.0, parallel=False, threads=0):
        """self must be the in-CSR (transpose)."""
        od = np.ascontiguousarray(outdeg, dtype=np.uint32)
        r = np.zeros(max(self.n, 1), dtype=np.float64)
        it = C.c_int()
        dl = C.c_double()
        _host_check(host_lib().ggpu_cpu_pagerank(self._h, _ptr(od), damping, iters, tol, int(parallel), threads,
                                                 _ptr(r, C.c_double), C.byref(it), C.byref(dl)))
        return r[: self.n], it.value, dl.value


class SnapStats(C.Structure):
    _fields_ = [("lines", C.c_uint64), ("raw_max_id", C.c_uint64), ("n", C.c_uint64), ("m_before", C.c_uint64),
                ("m", C.c_uint64)]


class GraphStats(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64), ("undirected_edges", C.c_uint64), ("min_deg", C.c_uint32),
                ("max_deg", C.c_uint32), ("max_in_deg", C.c_uint32), ("approx_diameter", C.c_uint32),
                ("mean_deg", C.c_double), ("isolated", C.c_uint64), ("num_cc", C.c_uint64),
                ("largest_cc", C.c_uint64)]


def random_perm(n: int, seed: int) -> np.ndarray:
    p = np.zeros(max(n, 1), dtype=np.uint32)
    _host_check(host_lib().ggpu_random_perm(n, seed, _ptr(p)))
    return p[:n]


def out_degrees(g: Graph) -> np.ndarray:
    return g.degrees()


# ------------------------------------------------------------------ context
class Context:
    """gpu::MetalContext (src/metal/context.hpp:21-59): one device, one stream."""

    def __init__(self, device: int = 0):
        lib = hip_lib()
        h = C.c_void_p()
        rc = lib.ggpu_ctx_create(device, C.byref(h))
        if rc != OK:
            raise RuntimeError(f"ggpu_ctx_create failed with status {rc} (is a gfx950 GPU visible?)")
        self._h = h
        self.lib = lib

    def __del__(self):
        if getattr(self, "_h", None) and _hip is not None:
            _hip.ggpu_ctx_destroy(self._h)
            self._h = None

    def check(self, rc: int):
        if rc != OK:
            msg = self.lib.ggpu_last_error(self._h).decode()
            if rc == EINVAL:
                raise ValueError(msg)
            if rc == ESTATE:
                raise RuntimeError(msg)
            raise RuntimeError(f"status {rc}: {msg}")

    def info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int()
        hbm = C.c_uint64()
        self.check(self.lib.ggpu_device_info(self._h, name, 256, C.byref(cus), C.byref(hbm)))
        return name.value.decode(), cus.value, hbm.value

    def sync(self):
        self.check(self.lib.ggpu_ctx_sync(self._h))


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def bfs_options(**kw) -> BfsOpts:
    o = BfsOpts()
    hip_lib().ggpu_bfs_opts_default(C.byref(o))
    actions = kw.pop("actions", None)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise ValueError(f"unknown BFS option {k}")
        setattr(o, k, v)
    if actions is not None:
        arr = (C.c_uint8 * len(actions))(*actions)
        o.actions = C.cast(arr, C.POINTER(C.c_uint8))
        o.n_actions = len(actions)
        o._keep = arr  # keep the buffer alive with the struct
    return o


@dataclass
class BfsRun:
    dist: np.ndarray
    trace: List[Level]
    gpu_ms: float
    levels: int


class GpuBfs:
    """gpu::GpuBfs (src/metal/bfs_gpu.hpp:35-105): the symmetric CSR lives in HBM;
    runs leave dist on the device until fetched."""

    def __init__(self, g: Graph, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.graph = g
        self.n = g.n
        rp, ci = g.raw_pointers()
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.ggpu_bfs_create(self.ctx._h, g.n, g.m, rp, ci, C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _hip is not None:
            _hip.ggpu_bfs_destroy(self._h)
            self._h = None

    def run(self, variant, source: int, opts: Optional[BfsOpts] = None, trace: bool = True,
            fetch: bool = True) -> BfsRun:
        v = BFS_VARIANTS[variant] if isinstance(variant, str) else int(variant)
        o = opts or bfs_options()
        cap = self.n + 2 if trace else 0
        tr = (Level * max(cap, 1))()
        ln = C.c_size_t()
        self.ctx.check(self.ctx.lib.ggpu_bfs_run(self._h, v, source, C.byref(o), tr if trace else None, cap,
                                                 C.byref(ln)))
        dist = self.fetch_dist() if fetch else None
        levels = [tr[i] for i in range(min(ln.value, cap))] if trace else []
        return BfsRun(dist, levels, self.last_gpu_ms(), self.last_levels())

    # the reference's per-variant entry points
    def run_edge(self, source, opts=None, trace=True):
        return self.run(BFS_EDGE, source, opts, trace)

    def run_binned(self, source, opts=None, trace=True):
        return self.run(BFS_BINNED, source, opts, trace)

    def run_do(self, source, opts=None, trace=True):
        return self.run(BFS_DO, source, opts, trace)

    def run_hybrid(self, source, opts=None, trace=True):
        return self.run(BFS_HYBRID, source, opts, trace)

    def fetch_dist(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), dtype=np.uint32)
        self.ctx.check(self.ctx.lib.ggpu_bfs_fetch_dist(self._h, _ptr(out)))
        return out[: self.n]

    def last_gpu_ms(self) -> float:
        return float(self.ctx.lib.ggpu_bfs_last_gpu_ms(self._h))

    def last_levels(self) -> int:
        return int(self.ctx.lib.ggpu_bfs_last_levels(self._h))

    # batched multi-source BFS
    def run_batch(self, sources, opts: Optional[BfsOpts] = None, trace: bool = False, fetch: bool = True):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        o = opts or bfs_options()
        res = MsbfsResult()
        cap = 4096 if trace else 0
        tr = (Level * max(cap, 1))()
        ln = C.c_size_t()
        self.ctx.check(self.ctx.lib.ggpu_msbfs_run(self._h, _ptr(src), len(src), C.byref(o), C.byref(res),
                                                   tr if trace else None, cap, C.byref(ln)))
        dist = None
        if fetch:
            dist = np.zeros(len(src) * max(self.n, 1), dtype=np.uint32)
            self.ctx.check(self.ctx.lib.ggpu_msbfs_fetch_dist(self._h, _ptr(dist)))
            dist = dist[: len(src) * self.n].reshape(len(src), self.n)
        levels = [tr[i] for i in range(min(ln.value, cap))] if trace else []
        return dist, res, levels

    # policy training data
    def _record(self, fn, args, opts, cap):
        o = opts or bfs_options()
        feat = np.zeros((cap, 5), dtype=np.float32)
        times = np.zeros((cap, 3), dtype=np.float32)
        tr = (Level * cap)()
        ln = C.c_size_t()
        self.ctx.check(fn(self._h, *args, C.byref(o), _ptr(feat, C.c_float), _ptr(times, C.c_float), tr, cap,
                          C.byref(ln)))
        k = min(ln.value, cap)
        return feat[:k], times[:k], [tr[i] for i in range(k)]

    def record_levels(self, source: int, opts: Optional[BfsOpts] = None, cap: int = 4096):
        """gpu_do from `source` once per action: (features[L,5], times_ms[L,3], trace)."""
        return self._record(self.ctx.lib.ggpu_bfs_record_levels, (int(source),), opts, cap)

    def record_batch_levels(self, sources, opts: Optional[BfsOpts] = None, cap: int = 4096):
        """Batched BFS once per action: (features[L,5], times_ms[L,3], trace)."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        return self._record(self.ctx.lib.ggpu_msbfs_record_levels, (_ptr(src), len(src)), opts, cap)


def pr_options(**kw) -> PrOpts:
    o = PrOpts()
    hip_lib().ggpu_pr_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise ValueError(f"unknown PageRank option {k}")
        setattr(o, k, v)
    return o


class GpuPageRank:
    """gpu::GpuPageRank (src/metal/pagerank_gpu.hpp:46-127), pull variants:
    in-CSR (transpose) + out-degrees in HBM."""

    def __init__(self, in_graph: Graph, outdeg, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.n = in_graph.n
        self.outdeg = np.ascontiguousarray(outdeg, dtype=np.uint32)
        rp, ci = in_graph.raw_pointers()
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.ggpu_pr_create(self.ctx._h, in_graph.n, in_graph.m, rp, ci, _ptr(self.outdeg),
                                                   C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _hip is not None:
            _hip.ggpu_pr_destroy(self._h)
            self._h = None

    def prepare_adaptive(self, small_max=16, large_min=256):
        self.ctx.check(self.ctx.lib.ggpu_pr_prepare_adaptive(self._h, small_max, large_min))

    def prepare_gs(self, pair_max_outdeg=2):
        self.ctx.check(self.ctx.lib.ggpu_pr_prepare_gs(self._h, pair_max_outdeg))

    def gs_info(self):
        col = C.c_int()
        sets = C.c_uint64()
        dfr = C.c_uint64()
        self.ctx.check(self.ctx.lib.ggpu_pr_gs_info(self._h, C.byref(col), C.byref(sets), C.byref(dfr)))
        return bool(col.value), sets.value, dfr.value

    def run(self, variant, opts: Optional[PrOpts] = None) -> PrResult:
        v = PR_VARIANTS[variant] if isinstance(variant, str) else int(variant)
        res = PrResult()
        self.ctx.check(self.ctx.lib.ggpu_pr_run(self._h, v, C.byref(opts or pr_options()), C.byref(res)))
        return res

    def ranks(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), dtype=np.float32)
        self.ctx.check(self.ctx.lib.ggpu_pr_fetch_ranks(self._h, _ptr(out, C.c_float)))
        return out[: self.n]

    def autotune(self, iters=5, opts: Optional[PrOpts] = None) -> PrOpts:
        o = opts or pr_options()
        self.ctx.check(self.ctx.lib.ggpu_pr_autotune(self._h, iters, C.byref(o)))
        return o

    def run_ppr(self, seeds, opts: Optional[PrOpts] = None) -> PrResult:
        s = np.ascontiguousarray(seeds, dtype=np.uint32)
        res = PrResult()
        self._ppr_b = len(s)
        self.ctx.check(self.ctx.lib.ggpu_ppr_run(self._h, _ptr(s), len(s), C.byref(opts or pr_options()),
                                                 C.byref(res)))
        return res

    def autotune_ppr(self, seeds, iters=3, opts: Optional[PrOpts] = None) -> PrOpts:
        """Fastest (large_min, medium_lanes, tg) of the PPR block for these
        seeds; the block then holds the last candidate's ranks."""
        s = np.ascontiguousarray(seeds, dtype=np.uint32)
        o = opts or pr_options()
        self._ppr_b = len(s)
        self.ctx.check(self.ctx.lib.ggpu_ppr_autotune(self._h, _ptr(s), len(s), iters, C.byref(o)))
        return o

    def ppr_block(self) -> np.ndarray:
        b = self._ppr_b
        out = np.zeros(max(self.n * b, 1), dtype=np.float32)
        self.ctx.check(self.ctx.lib.ggpu_ppr_fetch(self._h, _ptr(out, C.c_float)))
        return out[: self.n * b].reshape(self.n, b)


# ------------------------------------------------------------ learned policy
def policy_features(n_f, m_f, m_u, level, mean_degree) -> np.ndarray:
    f = np.zeros(5, dtype=np.float32)
    hip_lib().ggpu_policy_features(n_f, m_f, m_u, level, mean_degree, _ptr(f, C.c_float))
    return f


def policy_set(params, ctx: Optional[Context] = None, slot: int = POLICY_BATCHED):
    """Installs MLP parameters for single-source (POLICY_SINGLE) or batched
    (POLICY_BATCHED) traversals run with policy=POLICY_LEARNED."""
    c = ctx or default_context()
    p = np.ascontiguousarray(params, dtype=np.float32)
    c.check(c.lib.ggpu_policy_set(c._h, slot, _ptr(p, C.c_float), len(p)))


def policy_clear(ctx: Optional[Context] = None, slot: int = POLICY_BATCHED):
    c = ctx or default_context()
    c.check(c.lib.ggpu_policy_clear(c._h, slot))


def policy_get(ctx: Optional[Context] = None, slot: int = POLICY_BATCHED) -> Optional[np.ndarray]:
    c = ctx or default_context()
    p = np.zeros(MLP_PARAMS, dtype=np.float32)
    is_set = C.c_int()
    c.check(c.lib.ggpu_policy_get(c._h, slot, _ptr(p, C.c_float), MLP_PARAMS, C.byref(is_set)))
    return p if is_set.value else None


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-initialised parameters (biases zero) in the layout W1 b1 W2 b2 W3 b3."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / 5), 32 * 5), np.zeros(32), rng.normal(0, np.sqrt(2 / 32), 32 * 32),
             np.zeros(32), rng.normal(0, np.sqrt(2 / 32), 3 * 32), np.zeros(3)]
    return np.concatenate(parts).astype(np.float32)


def mlp_forward(params, x, ctx: Optional[Context] = None) -> np.ndarray:
    c = ctx or default_context()
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
    y = np.zeros((len(xx), 3), dtype=np.float32)
    c.check(c.lib.ggpu_mlp_forward(c._h, _ptr(p, C.c_float), _ptr(xx, C.c_float), len(xx), _ptr(y, C.c_float)))
    return y


def _mlp_data(x, t, mask):
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 5)
    tt = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 3)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1, 3)
    return xx, tt, mk


def mlp_loss_grad(params, x, t, mask=None, ctx: Optional[Context] = None):
    """GPU loss and gradient of the masked mean squared error."""
    c = ctx or default_context()
    p = np.ascontiguousarray(params, dtype=np.float32)
    xx, tt, mk = _mlp_data(x, t, mask)
    g = np.zeros(MLP_PARAMS, dtype=np.float32)
    loss = C.c_float()
    c.check(c.lib.ggpu_mlp_loss_grad(c._h, _ptr(p, C.c_float), _ptr(xx, C.c_float), _ptr(tt, C.c_float),
                                     None if mk is None else _ptr(mk, C.c_uint8), len(xx), _ptr(g, C.c_float),
                                     C.byref(loss)))
    return float(loss.value), g


def mlp_train(params, x, t, mask=None, epochs: int = 1000, lr: float = 1e-2, ctx: Optional[Context] = None):
    """Full-batch Adam on the GPU; returns (params, final loss)."""
    c = ctx or default_context()
    p = np.array(params, dtype=np.float32)
    xx, tt, mk = _mlp_data(x, t, mask)
    loss = C.c_float()
    c.check(c.lib.ggpu_mlp_train(c._h, _ptr(p, C.c_float), _ptr(xx, C.c_float), _ptr(tt, C.c_float),
                                 None if mk is None else _ptr(mk, C.c_uint8), len(xx), epochs, lr,
                                 C.byref(loss)))
    return p, float(loss.value)


def mlp_train_sgd(params, x, t, mask=None, epochs: int = 1000, lr: float = 1e-2, momentum: float = 0.0,
                  ctx: Optional[Context] = None):
    """Full-batch SGD with momentum on the GPU; returns (params, final loss)."""
    c = ctx or default_context()
    p = np.array(params, dtype=np.float32)
    xx, tt, mk = _mlp_data(x, t, mask)
    loss = C.c_float()
    c.check(c.lib.ggpu_mlp_train_sgd(c._h, _ptr(p, C.c_float), _ptr(xx, C.c_float), _ptr(tt, C.c_float),
                                     None if mk is None else _ptr(mk, C.c_uint8), len(xx), epochs, lr, momentum,
                                     C.byref(loss)))
    return p, float(loss.value)


POLICY_DIR = os.path.join(_PKG, "policy")
POLICY_FILES = {POLICY_SINGLE: os.path.join(POLICY_DIR, "single_source.f32"),
                POLICY_BATCHED: os.path.join(POLICY_DIR, "batched.f32")}


def load_policy_file(slot: int = POLICY_BATCHED, path: Optional[str] = None) -> Optional[np.ndarray]:
    path = path or POLICY_FILES[slot]
    if not os.path.exists(path):
        return None
    p = np.fromfile(path, dtype=np.float32)
    return p if len(p) == MLP_PARAMS else None


def install_policies(ctx: Optional[Context] = None) -> dict:
    """Sets every slot whose trained parameter file exists; returns {slot: installed}."""
    done = {}
    for slot in (POLICY_SINGLE, POLICY_BATCHED):
        p = load_policy_file(slot)
        if p is not None:
            policy_set(p, ctx, slot)
        done[slot] = p is not None
    return done


NOALLOC_MAX_N = 1 << 24


def bfs_run_noalloc(n: int, m: int, row_ptr: int, col_idx: int, source: int, dist: int,
                    opts: Optional[BfsOpts] = None, ctx: Optional[Context] = None) -> int:
    """ggpu_bfs_run_noalloc: one gpu_do traversal with the persistent loop over a
    CSR already in device memory (row_ptr, col_idx, dist are device addresses,
    e.g. a torch tensor's data_ptr()), using only the library's fixed working
    array; returns the number of levels. No counterpart in the reference."""
    ctx = ctx or default_context()
    levels = C.c_uint32()
    o = opts or bfs_options()
    ctx.check(ctx.lib.ggpu_bfs_run_noalloc(ctx._h, n, m, C.c_void_p(row_ptr), C.c_void_p(col_idx), source,
                                           C.byref(o), C.c_void_p(dist), C.byref(levels)))
    return int(levels.value)
