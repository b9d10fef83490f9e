"""gbench, the reference harness's command line over the two C-ABIs
(src/bench/main.cpp): exit codes, the CSV schema, --validate for every
variant, and --ordering (relabel, run, map back, validate).

The CPU cases run everywhere; the GPU cases (marked, one gbench process per
case) run every GPU variant of the binary with --validate, which compares
against the reference's own rules (exact dist for BFS, checkRanks for
PageRank). The device level loops are checked under the original numbering
only: an ordering is a host-side relabel that the other variants already
cover. Each case starts a fresh process, so this module is collected after
the in-process GPU tests.
"""
import csv
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GBENCH = os.path.join(ROOT, "graph_analytics_applesilicon_amd", "bin", "gbench")
ORDERINGS = ("original", "random", "degree", "rcm")
HEADER = ("timestamp,git_commit,chip,dataset,n,m,ordering,algo,variant,params,source,trial,"
          "time_ms_e2e,time_ms_gpu,gteps,iters,validated,load1")

needs_binary = pytest.mark.skipif(not os.path.exists(GBENCH), reason="gbench is not built")


def gbench(*args, timeout=300):
    return subprocess.run([GBENCH, *args], capture_output=True, text=True, timeout=timeout)


@needs_binary
def test_exit_codes():
    assert gbench("--graph", "synth:rmat12", "--algo", "bfs", "--variant", "gpu_nope").returncode == 2
    assert gbench("--graph", "synth:rmat12", "--algo", "pr", "--variant", "gpu_nope").returncode == 2
    assert gbench("--graph", "synth:rmat12", "--algo", "sssp").returncode == 2
    assert gbench("--graph", "synth:rmat12", "--algo", "msbfs", "--batch-order", "bogus").returncode == 2
    assert gbench("--graph", "synth:rmat12", "--algo", "bfs", "--source", "4096").returncode == 2
    assert gbench("--graph", "synth:rmat12", "--algo", "bfs", "--ordering", "spectral").returncode == 1


@needs_binary
@pytest.mark.parametrize("ordering", ORDERINGS)
def test_cpu_variants_validate_under_every_ordering(tmp_path, ordering):
    out = tmp_path / "r.csv"
    loc = tmp_path / "loc.csv"
    common = ["--warmup", "0", "--trials", "1", "--validate", "--csv", str(out), "--ordering", ordering,
              "--locality-csv", str(loc)]
    for variant in ("cpu_serial", "cpu_omp", "cpu_do"):
        r = gbench("--graph", "synth:two_components", "--algo", "bfs", "--variant", variant, "--sources", "4", *common)
        assert r.returncode == 0 and "validated OK" in r.stdout, r.stdout + r.stderr
    for variant in ("cpu_serial", "cpu_omp"):
        r = gbench("--graph", "synth:rmat12_directed", "--algo", "pr", "--variant", variant, *common)
        assert r.returncode == 0 and "validated OK" in r.stdout, r.stdout + r.stderr
    rows = list(csv.DictReader(open(out)))
    assert open(out).readline().strip() == HEADER
    assert {r["algo"] for r in rows} == {"bfs", "pr"} and {r["ordering"] for r in rows} == {ordering}
    assert sum(r["algo"] == "pr" for r in rows) == 2 and all(r["validated"] == "1" for r in rows)
    lrows = list(csv.DictReader(open(loc)))
    assert {r["ordering"] for r in lrows} == {ordering} and all(float(r["mean_gap"]) >= 0 for r in lrows)


BFS_GRAPHS = ("synth:rmat12", "synth:grid100x100", "synth:star100k")
PR_GRAPHS = ("synth:rmat12_directed", "synth:star100k")
# one gbench run per case: (ordering, graph, algo, what, extra arguments)
GPU_CASES = (
    [(o, g, "bfs", v, ["--variant", v, "--sources", "3"]) for o in ORDERINGS for g in BFS_GRAPHS
     for v in ("gpu_vertex", "gpu_naive", "gpu_binned", "gpu_edge", "gpu_do", "hybrid")]
    # the device level loops under the original numbering (an ordering is a host-side relabel)
    + [("original", g, "bfs", "gpu_do-" + loop, ["--variant", "gpu_do", "--loop", loop, "--sources", "3"])
       for g in BFS_GRAPHS for loop in ("graph", "persistent")]
    + [(o, g, "msbfs", "batch-" + b, ["--sources", "70", "--batch", "64", "--batch-order", b]) for o in ORDERINGS
       for g in BFS_GRAPHS for b in ("auto", "degree", "given")]
    + [(o, g, "pr", v, ["--variant", v]) for o in ORDERINGS for g in PR_GRAPHS
       for v in ("gpu_push", "gpu_pull_scalar", "gpu_pull_vector", "gpu_pull_adaptive")]
    + [(o, g, "pr", "gpu_pull_gs", ["--variant", "gpu_pull_gs", "--tol", "1e-6", "--ref-tol", "1e-10"])
       for o in ORDERINGS for g in PR_GRAPHS])


@pytest.mark.gpu
@needs_binary
@pytest.mark.parametrize("ordering,graph,algo,what,extra", GPU_CASES,
                         ids=["-".join((c[0], c[1][6:], c[2], c[3])) for c in GPU_CASES])
def test_gpu_variant_validates(tmp_path, ordering, graph, algo, what, extra):
    out = tmp_path / "r.csv"
    print(f"gbench {graph} {algo} {what} --ordering {ordering}", flush=True)
    r = gbench("--graph", graph, "--algo", algo, *extra, "--warmup", "1", "--trials", "2", "--validate", "--csv",
               str(out), "--ordering", ordering)
    assert r.returncode == 0 and "validated OK" in r.stdout, r.stdout + r.stderr
    rows = list(csv.DictReader(open(out)))
    assert rows and {x["ordering"] for x in rows} == {ordering}
    assert all(x["validated"] == "1" for x in rows)
