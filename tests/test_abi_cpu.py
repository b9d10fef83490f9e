"""The C-ABI boundary on the CPU: both libraries load, every function the
headers declare is exported, the pure host entry points behave, and the GPU
path fails loudly (no fallback) where no GPU is present. No compute call
touches a GPU here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import graph_analytics_applesilicon_amd as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^[A-Za-z_][\w \*]*?\b(ggpu_\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


@pytest.mark.parametrize("header,lib", [("ggpu_graph.h", G.HOST_LIB), ("ggpu_hip.h", G.HIP_LIB),
                                        ("ggpu_testing.h", G.HOST_LIB), ("ggpu_hip_testing.h", G.HIP_LIB)])
def test_every_declared_function_is_exported(header, lib):
    names = declared(header)
    assert len(names) > (10 if "testing" not in header else 0)
    so = C.CDLL(lib)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{os.path.basename(lib)} lacks {missing}"


def test_types_header_compiles_standalone(tmp_path):
    """The headers are plain C: a C compiler accepts them on their own."""
    src = tmp_path / "t.c"
    src.write_text('#include "ggpu_types.h"\n#include "ggpu_graph.h"\n#include "ggpu_hip.h"\n'
                   "int main(void) { ggpu_bfs_opts o; ggpu_pr_opts p; (void)o; (void)p; return 0; }\n")
    rc = os.system(f"gcc -std=c99 -Wall -Werror -I{ROOT}/include -c {src} -o {tmp_path}/t.o")
    assert rc == 0


def test_option_defaults_follow_the_reference():
    """GpuBfsOptions (src/metal/bfs_gpu.hpp:23-33) and GpuPrOptions (pagerank_gpu.hpp:27-36) defaults."""
    o = G.bfs_options()
    assert (o.tg, o.small_max, o.large_min, o.ept) == (256, 16, 256, 2)
    assert (o.alpha, o.beta, o.hybrid_t) == (15.0, 18.0, 100000)
    assert o.policy == G.POLICY_HEURISTIC and o.loop == G.LOOP_HOST
    p = G.pr_options()
    assert abs(p.damping - 0.85) < 1e-7 and p.iters == 20 and p.tol == 0.0 and p.tg == 256
    assert (p.small_max, p.large_min) == (16, 256)
    with pytest.raises(ValueError):
        G.bfs_options(no_such_option=1)


def test_variant_names_parse():
    lib = G.hip_lib()
    v = C.c_int()
    for name, want in G.BFS_VARIANTS.items():
        assert lib.ggpu_bfs_parse_variant(name.encode(), C.byref(v)) == 0 and v.value == want
    assert lib.ggpu_bfs_parse_variant(b"gpu_nonsense", C.byref(v)) != 0


def test_policy_features_definition():
    f = G.policy_features(1000, 5000, 10**6, 3, 12.5)
    want = [np.log2(1001) / 32, np.log2(5001) / 32, np.log2(10**6 + 1) / 32, 3 / 64, np.log2(13.5) / 16]
    np.testing.assert_allclose(f, want, rtol=1e-6)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_gpu_path_fails_loudly_without_a_gpu():
    with pytest.raises(RuntimeError):
        G.Context(0)
