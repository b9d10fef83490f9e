"""ctypes binding of the oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; the
product package never imports it. Two libraries:

* ``liboracle.so`` -- the C restatement in oracle.c (always built);
* ``_ref/libggpu_ref.so`` -- the reference's own CPU code compiled by
  oracle/Makefile from the reference checkout (present when it was built).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libggpu_ref.so")
INF = 0xFFFFFFFF

_or = None
_ref = None
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)


def _p(a, t=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _or
    if _or is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} not built (make -C oracle)")
        L = C.CDLL(ORACLE_LIB)
        L.or_bfs_serial.argtypes = [C.c_uint64, u32p, u32p, C.c_uint32, u32p]
        L.or_bfs_serial.restype = None
        L.or_bfs_do_levels.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, C.c_uint32, C.c_double, C.c_double, u32p,
                                       C.c_char_p, u64p, u64p, C.c_uint64]
        L.or_bfs_do_levels.restype = C.c_uint64
        L.or_pagerank.argtypes = [C.c_uint64, u32p, u32p, u32p, C.c_double, C.c_int, C.c_double, dp,
                                  C.POINTER(C.c_int), dp]
        L.or_pagerank.restype = None
        L.or_ppr.argtypes = [C.c_uint64, u32p, u32p, u32p, C.c_double, C.c_int, C.c_double, C.c_uint32, dp,
                             C.POINTER(C.c_int), dp]
        L.or_ppr.restype = None
        L.or_check_ranks.argtypes = [C.c_uint64, dp, fp, dp, dp, dp]
        L.or_check_ranks.restype = None
        L.or_mlp_nparam.restype = C.c_int
        L.or_mlp_forward.argtypes = [dp, dp, C.c_uint64, dp]
        L.or_mlp_forward.restype = None
        L.or_mlp_loss_grad.argtypes = [dp, dp, dp, C.POINTER(C.c_uint8), C.c_uint64, dp]
        L.or_mlp_loss_grad.restype = C.c_double
        _or = L
    return _or


def ref_lib() -> Optional[C.CDLL]:
    """The compiled reference CPU code, or None when it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        L = C.CDLL(REF_LIB)
        pp = C.POINTER(u32p)
        L.ref_gen_suite.argtypes = [C.c_char_p, u64p, u64p, u32p, pp, pp]
        L.ref_gen_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_uint64, u64p, u64p, u32p, pp, pp]
        L.ref_symmetrize.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, u32p, u32p, u64p, u64p, u32p, pp, pp]
        L.ref_transpose.argtypes = L.ref_symmetrize.argtypes
        L.ref_random_perm.argtypes = [C.c_uint64, C.c_uint64, u32p]
        L.ref_order_perm.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, u32p, u32p, C.c_char_p, C.c_uint64, u32p]
        L.ref_convert_snap.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_uint64), C.c_char_p,
                                       C.c_size_t]
        L.ref_locality.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, u32p, u32p, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        L.ref_locality.restype = None
        L.ref_permute.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, u32p, u32p, u32p, u64p, u64p, u32p, pp, pp]
        L.ref_bfs_serial.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, C.c_uint32, u32p]
        L.ref_bfs_serial.restype = None
        L.ref_bfs_cpu.argtypes = [C.c_int, C.c_uint64, C.c_uint64, u32p, u32p, C.c_uint32, C.c_double, C.c_double,
                                  C.c_int, u32p, C.c_char_p, u64p, u64p, C.c_uint64]
        L.ref_bfs_cpu.restype = C.c_uint64
        L.ref_pagerank.argtypes = [C.c_uint64, C.c_uint64, u32p, u32p, u32p, C.c_double, C.c_int, C.c_double,
                                   C.c_int, C.c_int, dp, dp]
        L.ref_pagerank.restype = C.c_int
        L.ref_check_ranks.argtypes = [C.c_uint64, dp, fp, dp, dp, dp]
        L.ref_check_ranks.restype = None
        L.ref_gs_no_worse.argtypes = [C.c_double] * 6
        L.ref_gs_no_worse.restype = C.c_int
        L.ref_ranks_ok.argtypes = [C.c_double] * 3
        L.ref_ranks_ok.restype = C.c_int
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_free.restype = None
        _ref = L
    return _ref


# ------------------------------------------------------------------ oracle
def bfs_serial(rp: np.ndarray, ci: np.ndarray, src: int) -> np.ndarray:
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.uint32)
    ci = np.ascontiguousarray(ci if len(ci) else np.zeros(1, np.uint32), np.uint32)
    d = np.zeros(max(n, 1), np.uint32)
    lib().or_bfs_serial(n, _p(rp), _p(ci), src, _p(d))
    return d[:n]


def bfs_do_levels(rp, ci, src, alpha=15.0, beta=18.0):
    """dist, [(direction, frontier, edges)] of the GAP direction-optimizing BFS."""
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.uint32)
    ci = np.ascontiguousarray(ci if len(ci) else np.zeros(1, np.uint32), np.uint32)
    d = np.zeros(max(n, 1), np.uint32)
    cap = n + 2
    dirs = C.create_string_buffer(cap)
    fr = np.zeros(cap, np.uint64)
    ed = np.zeros(cap, np.uint64)
    k = lib().or_bfs_do_levels(n, int(rp[-1]), _p(rp), _p(ci), src, alpha, beta, _p(d), dirs, _p(fr, C.c_uint64),
                               _p(ed, C.c_uint64), cap)
    levels = [(dirs.raw[i:i + 1].decode(), int(fr[i]), int(ed[i])) for i in range(min(k, cap))]
    return d[:n], levels


def pagerank(in_rp, in_ci, outdeg, damping=0.85, iters=20, tol=0.0):
    n = len(in_rp) - 1
    in_rp = np.ascontiguousarray(in_rp, np.uint32)
    in_ci = np.ascontiguousarray(in_ci if len(in_ci) else np.zeros(1, np.uint32), np.uint32)
    od = np.ascontiguousarray(outdeg, np.uint32)
    r = np.zeros(max(n, 1), np.float64)
    it = C.c_int()
    dl = C.c_double()
    lib().or_pagerank(n, _p(in_rp), _p(in_ci), _p(od), damping, iters, tol, _p(r, C.c_double), C.byref(it),
                      C.byref(dl))
    return r[:n], it.value, dl.value


def ppr(in_rp, in_ci, outdeg, seed, damping=0.85, iters=100, tol=0.0):
    n = len(in_rp) - 1
    in_rp = np.ascontiguousarray(in_rp, np.uint32)
    in_ci = np.ascontiguousarray(in_ci if len(in_ci) else np.zeros(1, np.uint32), np.uint32)
    od = np.ascontiguousarray(outdeg, np.uint32)
    r = np.zeros(max(n, 1), np.float64)
    it = C.c_int()
    dl = C.c_double()
    lib().or_ppr(n, _p(in_rp), _p(in_ci), _p(od), damping, iters, tol, seed, _p(r, C.c_double), C.byref(it),
                 C.byref(dl))
    return r[:n], it.value, dl.value


def check_ranks(ref: np.ndarray, got: np.ndarray):
    """(l1, max_rel_top1000, sum) as cpu::checkRanks."""
    ref = np.ascontiguousarray(ref, np.float64)
    got = np.ascontiguousarray(got, np.float32)
    l1, mx, s = C.c_double(), C.c_double(), C.c_double()
    lib().or_check_ranks(len(ref), _p(ref, C.c_double), _p(got, C.c_float), C.byref(l1), C.byref(mx), C.byref(s))
    return l1.value, mx.value, s.value


def ranks_ok(l1, max_rel, s) -> bool:
    """PrCheck::ok (src/cpu/pagerank_cpu.hpp:31)."""
    return l1 < 1e-4 and max_rel < 1e-3 and abs(s - 1.0) < 1e-4


def gs_no_worse(gs, jac) -> bool:
    """cpu::gsNoWorseThanJacobi (src/cpu/pagerank_cpu.hpp:80-84); gs/jac = (l1, max_rel, sum)."""
    budget = max(1e-4, jac[0])
    return gs[0] <= budget and abs(gs[2] - 1.0) <= budget and gs[1] <= max(1e-3, jac[1] * (1 + 1e-4))


def mlp_forward(params, x):
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 5)
    y = np.zeros((len(x), 3), np.float64)
    lib().or_mlp_forward(_p(p, C.c_double), _p(x, C.c_double), len(x), _p(y, C.c_double))
    return y


def mlp_loss_grad(params, x, t, mask=None):
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 5)
    t = np.ascontiguousarray(t, np.float64).reshape(-1, 3)
    g = np.zeros(lib().or_mlp_nparam(), np.float64)
    mk = None
    if mask is not None:
        mk = np.ascontiguousarray(mask, np.uint8).reshape(-1, 3)
    loss = lib().or_mlp_loss_grad(_p(p, C.c_double), _p(x, C.c_double), _p(t, C.c_double),
                                  _p(mk, C.c_uint8) if mk is not None else None, len(x), _p(g, C.c_double))
    return loss, g


# ------------------------------------------------------------- reference
def _take_csr(L, n, m, fl, rp, ci):
    r = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
    c = np.ctypeslib.as_array(ci, shape=(max(m.value, 1),)).copy()[: m.value]
    L.ref_free(C.cast(rp, C.c_void_p))
    L.ref_free(C.cast(ci, C.c_void_p))
    return r, c, fl.value


def ref_suite(name: str):
    L = ref_lib()
    n, m, fl = C.c_uint64(), C.c_uint64(), C.c_uint32()
    rp, ci = u32p(), u32p()
    if L.ref_gen_suite(name.encode(), C.byref(n), C.byref(m), C.byref(fl), C.byref(rp), C.byref(ci)) != 0:
        raise KeyError(name)
    return _take_csr(L, n, m, fl, rp, ci)


def ref_rmat(scale, ef, directed, seed):
    L = ref_lib()
    n, m, fl = C.c_uint64(), C.c_uint64(), C.c_uint32()
    rp, ci = u32p(), u32p()
    if L.ref_gen_rmat(scale, ef, int(directed), seed, C.byref(n), C.byref(m), C.byref(fl), C.byref(rp),
                      C.byref(ci)) != 0:
        raise RuntimeError("ref rmat failed")
    return _take_csr(L, n, m, fl, rp, ci)


def _ref_unary(fn, rp, ci, flags, *extra):
    L = ref_lib()
    rp = np.ascontiguousarray(rp, np.uint32)
    ci = np.ascontiguousarray(ci if len(ci) else np.zeros(1, np.uint32), np.uint32)
    n, m, fl = C.c_uint64(), C.c_uint64(), C.c_uint32()
    orp, oci = u32p(), u32p()
    getattr(L, fn)(len(rp) - 1, int(rp[-1]), flags, _p(rp), _p(ci), *extra, C.byref(n), C.byref(m), C.byref(fl),
                   C.byref(orp), C.byref(oci))
    return _take_csr(L, n, m, fl, orp, oci)


def ref_symmetrize(rp, ci, flags):
    return _ref_unary("ref_symmetrize", rp, ci, flags)


def ref_transpose(rp, ci, flags):
    return _ref_unary("ref_transpose", rp, ci, flags)


def ref_permute(rp, ci, flags, perm):
    p = np.ascontiguousarray(perm, np.uint32)
    return _ref_unary("ref_permute", rp, ci, flags, _p(p))


def ref_random_perm(n, seed):
    p = np.zeros(max(n, 1), np.uint32)
    ref_lib().ref_random_perm(n, seed, _p(p))
    return p[:n]


def ref_order_perm(rp, ci, flags, ordering, seed=1):
    """graph::reorder::compute of the compiled reference."""
    rp = np.ascontiguousarray(rp, np.uint32)
    ci = np.ascontiguousarray(ci, np.uint32)
    n = len(rp) - 1
    p = np.zeros(max(n, 1), np.uint32)
    if ref_lib().ref_order_perm(n, len(ci), flags, _p(rp), _p(ci), ordering.encode(), seed, _p(p)):
        raise ValueError(ordering)
    return p[:n]


def ref_convert_snap(src, dst, directed):
    """graph::convertSnap of the compiled reference: writes dst (.bin); returns
    (lines, raw_max_id, n, m_before, m), or raises ValueError with its message."""
    st = (C.c_uint64 * 5)()
    err = C.create_string_buffer(512)
    if ref_lib().ref_convert_snap(str(src).encode(), str(dst).encode(), int(directed), st, err, 512):
        raise ValueError(err.value.decode())
    return tuple(int(x) for x in st)


def ref_locality(rp, ci, flags):
    rp = np.ascontiguousarray(rp, np.uint32)
    ci = np.ascontiguousarray(ci, np.uint32)
    g, s = C.c_double(), C.c_double()
    ref_lib().ref_locality(len(rp) - 1, len(ci), flags, _p(rp), _p(ci), C.byref(g), C.byref(s))
    return g.value, s.value


def ref_bfs_serial(rp, ci, src):
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.uint32)
    ci = np.ascontiguousarray(ci if len(ci) else np.zeros(1, np.uint32), np.uint32)
    d = np.zeros(max(n, 1), np.uint32)
    ref_lib().ref_bfs_serial(n, int(rp[-1]), _p(rp), _p(ci), src, _p(d))
    return d[:n]


def ref_bfs_cpu(rp, ci, src, variant="cpu_do", alpha=15.0, beta=18.0, threads=0, trace=True):
    """The reference's cpu_do / cpu_omp; returns dist and [(dir, frontier, edges)]."""
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.uint32)
    ci = np.ascontiguousarray(ci if len(ci) else np.zeros(1, np.uint32), np.uint32)
    d = np.zeros(max(n, 1), np.uint32)
    cap = n + 2 if trace else 0
    dirs = C.create_string_buffer(max(cap, 1))
    fr = np.zeros(max(cap, 1), np.uint64)
    ed = np.zeros(max(cap, 1), np.uint64)
    k = ref_lib().ref_bfs_cpu(2 if variant == "cpu_do" else 1, n, int(rp[-1]), _p(rp), _p(ci), src, alpha, beta,
                              threads, _p(d), dirs if trace else None, _p(fr, C.c_uint64), _p(ed, C.c_uint64), cap)
    levels = [(dirs.raw[i:i + 1].decode(), int(fr[i]), int(ed[i])) for i in range(min(k, cap))] if trace else []
    return d[:n], levels


def ref_pagerank(in_rp, in_ci, outdeg, damping=0.85, iters=20, tol=0.0, parallel=False, threads=0):
    n = len(in_rp) - 1
    in_rp = np.ascontiguousarray(in_rp, np.uint32)
    in_ci = np.ascontiguousarray(in_ci if len(in_ci) else np.zeros(1, np.uint32), np.uint32)
    od = np.ascontiguousarray(outdeg, np.uint32)
    r = np.zeros(max(n, 1), np.float64)
    dl = C.c_double()
    it = ref_lib().ref_pagerank(n, int(in_rp[-1]), _p(in_rp), _p(in_ci), _p(od), damping, iters, tol,
                                int(parallel), threads, _p(r, C.c_double), C.byref(dl))
    return r[:n], it, dl.value


def ref_check_ranks(ref, got):
    ref = np.ascontiguousarray(ref, np.float64)
    got = np.ascontiguousarray(got, np.float32)
    l1, mx, s = C.c_double(), C.c_double(), C.c_double()
    ref_lib().ref_check_ranks(len(ref), _p(ref, C.c_double), _p(got, C.c_float), C.byref(l1), C.byref(mx),
                              C.byref(s))
    return l1.value, mx.value, s.value


def ref_ranks_ok(l1, max_rel, s) -> bool:
    """The reference's PrCheck::ok (compiled)."""
    return bool(ref_lib().ref_ranks_ok(l1, max_rel, s))


def ref_gs_no_worse(gs, jac) -> bool:
    """The reference's gsNoWorseThanJacobi (compiled); gs/jac = (l1, max_rel, sum)."""
    return bool(ref_lib().ref_gs_no_worse(gs[0], gs[1], gs[2], jac[0], jac[1], jac[2]))
