"""Cross-check against NetworkX, an implementation independent of both this
package and the reference (the reference's tools/validate_networkx.py plays
this part for its CPU code).

BFS: depths from several sources on the symmetrized graph must equal
networkx.single_source_shortest_path_length exactly, for the host library's
cpu_serial and, with --gpu, for the HIP gpu_do.
PageRank (d = 0.85, on the directed graph): networkx.pagerank with
tol = 1e-10 / n (an L1 change below 1e-10) is the fixed point. The host
library's FP64 cpu_serial run to tol 1e-10 must be within L1 1e-6 and
top-1000 relative error 1e-4 of it (both sides FP64); with --gpu the FP32
gpu_pull_adaptive, run until its certified distance to the fixed point is
below 1e-4, must pass the reference's acceptance rule against it (L1 < 1e-4,
top-1000 relative error < 1e-3, |sum - 1| < 1e-4; src/cpu/pagerank_cpu.hpp).

  python -m graph_analytics_applesilicon_amd.validate_networkx [--graphs synth:NAME | FILE.bin ...]
         [--sources 4] [--gpu] [--csv validate_networkx.csv]

Exit status 0 when every check passes, 1 otherwise.
"""
import argparse
import csv
import sys

import numpy as np

import graph_analytics_applesilicon_amd as G

UNREACHED = 0xFFFFFFFF
SUITE = ["path1000", "star100k", "complete200", "two_components", "empty1000", "grid100x100", "rmat12",
         "rmat12_directed"]


def load(spec):
    return G.Graph.named(spec[6:]) if spec.startswith("synth:") else G.Graph.load(spec)


def sources_of(g, k, seed=7):
    """Vertex 0, the highest-degree vertex, then seeded random vertices."""
    if g.n == 0:
        return []
    rng = np.random.default_rng(seed)
    out = [0, int(np.argmax(g.degrees()))]
    while len(out) < min(k, g.n):
        out.append(int(rng.integers(g.n)))
    return list(dict.fromkeys(out))


def _nx_graph(g, directed):
    import networkx as nx
    h = nx.DiGraph() if directed else nx.Graph()
    h.add_nodes_from(range(g.n))
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.row_ptr.astype(np.int64)))
    h.add_edges_from(zip(src.tolist(), g.col_idx.tolist()))
    return h


def check_bfs(spec, g, k, gpu):
    import networkx as nx
    s = g.symmetrize()
    h = _nx_graph(s, False)
    bfs = G.GpuBfs(s) if gpu else None
    rows = []
    for src in sources_of(s, k):
        want = np.full(s.n, UNREACHED, dtype=np.uint64)
        for v, d in nx.single_source_shortest_path_length(h, src).items():
            want[v] = d
        runs = [("cpu_serial", s.cpu_bfs(src)[0])]
        if bfs is not None:
            runs.append(("gpu_do", bfs.run("gpu_do", src, trace=False).dist))
        for variant, dist in runs:
            bad = int(np.count_nonzero(dist.astype(np.uint64) != want))
            rows.append({"graph": spec, "algo": "bfs", "variant": variant, "source": src,
                         "reached": int(np.count_nonzero(want != UNREACHED)), "error": bad,
                         "status": "OK" if bad == 0 else "FAIL"})
    return rows


def _compare(ref, got):
    l1 = float(np.abs(got - ref).sum())
    top = np.argsort(-ref, kind="stable")[:1000]
    rel = float(np.max(np.abs(got[top] - ref[top]) / ref[top])) if len(ref) else 0.0
    return l1, rel, float(got.sum())


def check_pr(spec, g, gpu):
    import networkx as nx
    n = g.n
    if n == 0:
        return []
    pr = nx.pagerank(_nx_graph(g, True), alpha=0.85, tol=1e-10 / n, max_iter=10000)
    ref = np.array([pr[v] for v in range(n)])
    t = g.transpose() if g.directed else g
    od = g.degrees()
    rows = []
    ours, _, _ = t.cpu_pagerank(od, 0.85, iters=10000, tol=1e-10)
    l1, rel, s = _compare(ref, ours)
    ok = l1 < 1e-6 and rel < 1e-4
    rows.append({"graph": spec, "algo": "pr", "variant": "cpu_serial", "source": "", "reached": "",
                 "error": f"L1={l1:.3e} rel={rel:.3e}", "status": "OK" if ok else "FAIL"})
    if gpu:
        p = G.GpuPageRank(t, od)
        # ||r_k - r*||_1 <= d / (1 - d) * delta_k: stop when that bound is < 1e-4
        p.run("gpu_pull_adaptive", G.pr_options(tol=1e-4 * 0.15 / 0.85, iters=1000))
        l1, rel, s = _compare(ref, p.ranks().astype(np.float64))
        ok = l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4
        rows.append({"graph": spec, "algo": "pr", "variant": "gpu_pull_adaptive", "source": "", "reached": "",
                     "error": f"L1={l1:.3e} rel={rel:.3e} sum-1={s - 1:.1e}", "status": "OK" if ok else "FAIL"})
    return rows


def validate(specs, k=4, gpu=False):
    rows = []
    for spec in specs:
        g = load(spec)
        rows += check_bfs(spec, g, k, gpu)
        rows += check_pr(spec, g, gpu)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--graphs", nargs="*", default=["synth:" + s for s in SUITE])
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--gpu", action="store_true", help="also check the HIP gpu_do and gpu_pull_adaptive")
    ap.add_argument("--csv", default="")
    args = ap.parse_args(argv)
    rows = validate(args.graphs, args.sources, args.gpu)
    for r in rows:
        print(f"{r['graph']:>24} {r['algo']:>4} {r['variant']:>18} {str(r['source']):>8} {r['error']!s:>36} "
              f"{r['status']}")
    if args.csv:
        with open(args.csv, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=list(rows[0].keys()) if rows else ["graph"])
            w.writeheader()
            w.writerows(rows)
    failed = [r for r in rows if r["status"] != "OK"]
    print(f"{len(rows) - len(failed)} / {len(rows)} checks OK")
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
