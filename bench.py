#!/usr/bin/env python3
"""Benchmark of the BFS / PageRank hot path on one MI355X.

Workload (the scale-22 RMAT graph of the benchmark configurations): undirected
RMAT, scale 22, a/b/c/d = .57/.19/.19/.05, 8 generated edges per vertex,
deduplicated, self-loops removed, symmetrized, ids permuted from seed 2;
64 seeded sources in the largest component (pickSources, seed 2).

One step = one batch of the 64 sources through BFS on the GPU, inputs
already in HBM. `value` is the aggregate GTEPS of the batch (sum over sources
of the component's undirected edges / batch time). Beside it the line carries
the plain per-source gpu_do rate (and gpu_do under the host and the persistent
loops, heuristic and learned, with its roofline), the batch in the caller's
numbering, the level loops on the grid graph, PageRank time to L1 < 1e-4
against the FP64 oracle on the scale-23 graph and ms per iteration on the
small RMAT, scale-22 RMAT and grid graphs, the personalized-PageRank block,
the roofline of the dominant kernel (HIP events on the library's stream) and
the reference's own cpu_do / cpu_omp timed on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-pagerank]
                       [--only bfs|pagerank|loops|ppr|prw]
       python bench.py --copy-calibration   (a plain device copy of 1 GiB, three times: the
                                             reference kernel for the FETCH_SIZE / WRITE_SIZE scale)
       python bench.py --gather-calibration (random 128-byte and 8-byte gathers of known size)
The roofline's `traffic` (HBM bytes per launch) is read from
profiles/hbm_traffic.json, which holds the rocprofv3 FETCH_SIZE and WRITE_SIZE
readings of the dominant kernels, FETCH_SIZE scaled by the random-gather
calibration (the copy calibration, twice as large, is listed beside it).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
COPY_BYTES = 1 << 30


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter runs, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            rows = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    for r in rows:
        if r.get("kernel") == kernel and r.get("workload") == workload:
            return r.get("traffic_bytes")
    return None


def copy_calibration():
    """One plain device-to-device copy of COPY_BYTES, three times (torch's copy kernel)."""
    import torch
    x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "launches": 3}), flush=True)


GATHER_ROWS = 1 << 22


def gather_calibration():
    """Random gathers of known size from a 1 GiB table, three launches each
    (torch's index_select, seeded indices): GATHER_ROWS rows of 128 bytes
    (a whole L2 line each), then GATHER_ROWS single 8-byte elements. The
    counters' reading per launch against the bytes the rows are, and what an
    8-byte random gather costs as the counters see it."""
    import torch
    table = torch.ones((1 << 30) // 4, dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    rows128 = table.view(-1, 32)
    idx = torch.randint(0, rows128.shape[0], (GATHER_ROWS,), device="cuda", generator=gen)
    for _ in range(3):
        out = rows128.index_select(0, idx)
    torch.cuda.synchronize()
    elems8 = table.view(torch.int64)
    idx8 = torch.randint(0, elems8.shape[0], (GATHER_ROWS,), device="cuda", generator=gen)
    for _ in range(3):
        out8 = elems8.index_select(0, idx8)
    torch.cuda.synchronize()
    del out, out8
    print(json.dumps({"gather_rows": GATHER_ROWS, "row_bytes": [128, 8], "table_bytes": 1 << 30,
                      "launches_each": 3}), flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--graph", default="rmat22_c2")
    ap.add_argument("--no-pagerank", action="store_true")
    ap.add_argument("--pr-graph", default="rmat23_c5",
                    help="graph of the PageRank part (default: the scale-23 heavy-skew RMAT)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="auto", choices=["auto", "batched", "sequential"])
    ap.add_argument("--only", default="all", choices=["all", "bfs", "pagerank", "loops", "ppr", "prw"],
                    help="profiling runs: just one part (the JSON line then carries that part)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--gather-calibration", action="store_true",
                    help="random 128-byte and 8-byte gathers from a 1 GiB table (counter calibration runs)")
    ap.add_argument("--no-ref", action="store_true", help="skip the FP64 PageRank reference (profiling runs)")
    ap.add_argument("--no-loops", action="store_true", help="skip the level-loop comparison on the grid graph")
    ap.add_argument("--loop-modes", default="host,graph_k8,graph_k64,persistent,hybrid",
                    help="level loops the grid comparison runs (comma list)")
    ap.add_argument("--no-pr-workloads", action="store_true",
                    help="skip PageRank ms/iteration on the small RMAT, scale-22 RMAT and grid graphs")
    ap.add_argument("--no-ppr", action="store_true", help="skip the personalized-PageRank block")
    ap.add_argument("--ppr-seeds", type=int, default=64, help="columns of the personalized-PageRank block")
    ap.add_argument("--ppr-assign", default=None, metavar="ORDER:LARGE_MIN:LANES:TG",
                    help="fixed PPR work assignment instead of the autotuner's (profiling runs), e.g. degree:1024:64:1024")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gather_calibration:
        gather_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this build runs on a single "
              "MI355X; batches of sources and PPR seeds shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    import graph_analytics_applesilicon_amd as G
    from oracle import pyoracle as O

    ctx = G.default_context()
    name, cus, _ = ctx.info()
    bfs_res = pr_res = cpu = loops = ppr = prw = None
    if args.only in ("all", "bfs"):
        bfs_res = bfs_part(G, O, args, ctx)
    if args.only == "loops" or (args.only == "all" and not args.no_loops):
        loops = loops_part(G, O, ctx, modes=args.loop_modes.split(","))
    if args.only == "pagerank" or (args.only == "all" and not args.no_pagerank):
        pg = G.Graph.named(args.pr_graph)
        pr_res = pagerank_part(G, O, pg, ctx, args.pr_graph, with_ref=not args.no_ref)
        pr_res["graph"] = args.pr_graph
    if args.only == "prw" or (args.only == "all" and not args.no_pr_workloads):
        prw = pr_workloads_part(G, O, ctx)
    if args.only == "ppr" or (args.only == "all" and not args.no_ppr):
        pg = pg if pr_res is not None else G.Graph.named(args.pr_graph)
        ppr = ppr_part(G, pg, ctx, args.pr_graph, args.ppr_seeds, assign=args.ppr_assign)
    if bfs_res is not None and not args.no_cpu:
        cpu = cpu_baseline(O, bfs_res.pop("_graph"), bfs_res.pop("_sources"), bfs_res["comp_edges"])
    if bfs_res is not None:
        bfs_res.pop("_graph", None)
        bfs_res.pop("_sources", None)

    if bfs_res is not None:
        line = {
            "metric": "BFS GTEPS batched+learned vs plain gpu_do; PageRank ms to L1<1e-4 vs FP64",
            "value": bfs_res["value"],
            "unit": "GTEPS",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": bfs_res["ms_per_step"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "u32",
            "data": "synthetic",
            "config": {"workload": f"{args.graph}: undirected RMAT scale 22, ef 8, a/b/c .57/.19/.19, ids "
                                   f"permuted (seed 2); {bfs_res['sources']} seeded LCC sources per step",
                       "n": bfs_res["n"], "m_arcs": bfs_res["m"], "sources": bfs_res["sources"],
                       "path": bfs_res["path"], "device": name, "cus": cus},
            "roofline": bfs_res.pop("roofline"),
            "cpu_baseline": cpu,
            # beside the headline: the same batch traversed in the caller's numbering
            # (no renumbered copy, no un-permute), and the first call, which builds the
            # renumbered copy of the graph and is not a timed step
            "batch_caller_numbering": {"ms": bfs_res.get("given_order_batched_ms"),
                                       "gteps": (bfs_res["sources"] * bfs_res["comp_edges"]
                                                 / (bfs_res["given_order_batched_ms"] / 1e3) / 1e9
                                                 if bfs_res.get("given_order_batched_ms") else None)},
            "first_call_ms": bfs_res.get("first_call_ms"),
        }
        for k in ("value", "ms_per_step", "n", "m", "sources", "path"):
            bfs_res.pop(k, None)
        line["bfs"] = bfs_res
        if loops is not None:
            line["level_loops_grid_c3"] = loops
        if pr_res is not None:
            line["pagerank_roofline"] = pr_res.pop("roofline")
            line["pagerank"] = pr_res
        if ppr is not None:
            line["ppr"] = ppr
        if prw is not None:
            line["pagerank_workloads"] = prw
    elif ppr is not None and pr_res is None:
        line = {"metric": f"personalized PageRank block ms per sweep ({args.ppr_seeds} seeds)",
                "value": ppr["ms_per_sweep"], "unit": "ms", "n_gpus": 1, "steps": ppr["sweeps"], "warmup": 1,
                "ms_per_step": ppr["ms_per_sweep"], "higher_is_better": False, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": ppr["workload"], "device": name}, "roofline": ppr.pop("roofline"),
                "cpu_baseline": None, "ppr": ppr}
    elif pr_res is None and loops is None and prw is not None:
        c2 = prw["rmat22_c2"]["gpu_pull_adaptive"]
        line = {"metric": "PageRank ms per iteration (gpu_pull_adaptive, rmat22_c2)", "value": c2["ms_per_iteration"],
                "unit": "ms", "n_gpus": 1, "steps": 20, "warmup": 1, "ms_per_step": c2["ms_per_iteration"],
                "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic", "config": {"workload": "rmat16_c1, rmat22_c2, grid_c3", "device": name},
                "roofline": None, "cpu_baseline": None, "pagerank_workloads": prw}
    elif pr_res is None:
        line = {"metric": "BFS us per level on the grid graph (gpu_do)", "value": loops["persistent"]["us_per_level"],
                "unit": "us/level", "n_gpus": 1, "steps": 1, "warmup": 1,
                "ms_per_step": loops["persistent"]["ms_per_traversal"], "higher_is_better": False,
                "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
                "config": {"workload": loops["workload"], "device": name}, "roofline": None,
                "cpu_baseline": None, "level_loops_grid_c3": loops}
    else:
        jac = pr_res["jacobi_certified"]
        line = {
            "metric": "PageRank ms to L1<1e-4 vs FP64 (certified Jacobi)",
            "value": jac["ms"], "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": jac["ms"],
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{args.pr_graph}: RMAT scale 23, ef 16, a/b/c .60/.17/.17", "device": name},
            "roofline": pr_res.pop("roofline"), "cpu_baseline": None, "pagerank": pr_res,
        }
    print(json.dumps(line), flush=True)


def bfs_part(G, O, args, ctx):
    t0 = time.time()
    g = G.Graph.named(args.graph)
    n, m = g.n, g.m
    sources = g.pick_sources(args.sources, 2)
    lcc, _ = g.largest_component()
    comp_edges = lcc.m // 2  # every source lies in the largest component
    log(f"graph {args.graph}: n {n} m {m} (lcc undirected edges {comp_edges}), {len(sources)} sources, "
        f"generated in {time.time() - t0:.1f}s")
    del lcc
    bfs = G.GpuBfs(g, ctx)
    rp, ci = g.row_ptr, g.col_idx
    out = {"comp_edges": comp_edges, "n": n, "m": m, "sources": len(sources), "_graph": g, "_sources": sources}

    # ---------------------------------------------------------- plain gpu_do
    plain_opts = G.bfs_options(time_levels=0)
    for s in sources[:2]:
        bfs.run("gpu_do", int(s), plain_opts, trace=False, fetch=False)
    t = time.perf_counter()
    for s in sources:
        bfs.run("gpu_do", int(s), plain_opts, trace=False, fetch=False)
    plain_s = time.perf_counter() - t
    plain_gteps = len(sources) * comp_edges / plain_s / 1e9
    log(f"plain gpu_do: {len(sources)} sequential traversals {plain_s * 1e3:.2f} ms, {plain_gteps:.2f} GTEPS")
    bad = 0
    for s in sources[:4]:
        d = bfs.run("gpu_do", int(s), plain_opts, trace=False).dist
        bad += int(not np.array_equal(d, O.bfs_serial(rp, ci, int(s))))
    out["plain_gpu_do"] = {"gteps": plain_gteps, "ms_per_batch": plain_s * 1e3,
                           "bit_exact_sources_checked": 4, "mismatches": bad}
    out["single_source"] = single_source_part(G, bfs, sources, comp_edges, n, m, ctx)

    # ---------------------------------------------------------- timed steps
    batched = args.mode in ("auto", "batched")
    bopts = G.bfs_options(policy=G.POLICY_LEARNED)
    if batched:
        installed = G.install_policies(ctx)
        out["learned_policy_installed"] = bool(installed[G.POLICY_BATCHED])

    def step():
        if batched:
            _, res, _ = bfs.run_batch(sources, bopts, fetch=False)
            return res.gpu_ms
        for s in sources:
            bfs.run("gpu_do", int(s), plain_opts, trace=False, fetch=False)
        return 0.0

    # the first batched call also builds the engine's derived arrays (the degree-renumbered
    # copy of the graph, heavy items, LDS-slot col_idx): reported, not timed as a step
    t = time.perf_counter()
    step()
    ctx.sync()
    out["first_call_ms"] = (time.perf_counter() - t) * 1e3
    for _ in range(max(0, args.warmup - 1)):
        step()
    ctx.sync()
    t = time.perf_counter()
    kernel_ms = [step() for _ in range(args.steps)]
    ctx.sync()
    total_s = time.perf_counter() - t
    ms_per_step = total_s * 1e3 / args.steps
    value = len(sources) * comp_edges / (ms_per_step / 1e3) / 1e9
    log(f"timed: {args.steps} steps, {ms_per_step:.3f} ms/step, {value:.2f} GTEPS "
        f"({'batched' if batched else 'sequential gpu_do'})")
    out.update({"value": value, "ms_per_step": ms_per_step,
                "path": ("batched BFS, learned policy: persistent traversal of the degree-renumbered graph "
                         "+ un-permute of the depths" if batched else "64 sequential gpu_do"),
                "speedup_vs_plain_gpu_do": value / plain_gteps if plain_gteps else None})

    roof = None
    if batched:
        dist, res, _ = bfs.run_batch(sources, bopts, fetch=True)
        checked = list(range(0, len(sources), max(1, len(sources) // 8)))
        mism = sum(int(not np.array_equal(dist[i], O.bfs_serial(rp, ci, int(sources[i])))) for i in checked)
        out["parity"] = {"sources_checked": len(checked), "mismatches": mism, "levels": res.levels,
                         "level_actions": list(res.level_actions)}
        heur = G.bfs_options(policy=G.POLICY_HEURISTIC)
        bfs.run_batch(sources, heur, fetch=False)
        hms = float(np.median([bfs.run_batch(sources, heur, fetch=False)[1].gpu_ms for _ in range(3)]))
        out["heuristic_batched_ms"] = hms
        # the same batch traversed in the caller's numbering (no renumbering, no un-permute)
        given = G.bfs_options(policy=G.POLICY_LEARNED, batch_order=G.BATCH_ORDER_GIVEN)
        bfs.run_batch(sources, given, fetch=False)
        out["given_order_batched_ms"] = float(np.median([bfs.run_batch(sources, given, fetch=False)[1].gpu_ms
                                                         for _ in range(3)]))
        # S = 256 sources in one batch (the second batch size of the workload)
        s256 = g.pick_sources(256, 2)
        bfs.run_batch(s256, bopts, fetch=False)
        ms256 = float(np.median([bfs.run_batch(s256, bopts, fetch=False)[1].gpu_ms for _ in range(3)]))
        out["batch256"] = {"sources": len(s256), "gpu_ms": ms256,
                           "gteps": len(s256) * comp_edges / (ms256 / 1e3) / 1e9}
        # one batch = the traversal launch, the depth pass and the un-permute (events
        # around the three): 4(n+1) + 4m + 4*n*S algorithmic bytes (one CSR pass, S depth rows)
        bytes_batch = 4 * (n + 1) + 4 * m + 4 * n * len(sources)
        avg_ms = float(np.mean(kernel_ms))
        ach = bytes_batch / (avg_ms / 1e3) / 1e9
        roof = {"kernel": "k_msbfs<1> + k_ms_finish<1> + k_ms_unpermute<1> (one batch)", "bound": "hbm",
                "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": measured_traffic("k_msbfs<1>+k_ms_finish<1>+k_ms_unpermute<1>",
                                            f"{args.graph}/k{len(sources)}"),
                "bytes_per_launch": bytes_batch, "avg_launch_ms": avg_ms}
    out["roofline"] = roof
    return out


def single_source_part(G, bfs, sources, comp_edges, n, m, ctx, trials=3):
    """gpu_do from every source, host loop against the persistent loop, heuristic
    against the learned single-source policy: per source one warm-up and
    `trials` timed runs (wall clock around the call, depths left on the
    device), the median; GTEPS as the harmonic mean over sources
    (src/bench/main.cpp:308-370). The roofline counts one pass over the CSR
    and the depth array per traversal, 4(n+1) + 4m + 4n bytes."""
    G.install_policies(ctx)
    b_bfs = 4 * (n + 1) + 4 * m + 4 * n
    res = {"bytes_per_traversal": b_bfs, "sources": len(sources)}
    for lname, loop in (("host", G.LOOP_HOST), ("persistent", G.LOOP_PERSISTENT)):
        for pname, pol in (("heuristic", G.POLICY_HEURISTIC), ("learned", G.POLICY_LEARNED)):
            o = G.bfs_options(loop=loop, policy=pol, time_levels=0)
            def once(s):
                t = time.perf_counter()
                bfs.run("gpu_do", int(s), o, trace=False, fetch=False)
                return (time.perf_counter() - t) * 1e3

            med = []
            for s in sources:
                once(s)
                med.append(float(np.median([once(s) for _ in range(trials)])))
            mean_ms = float(np.mean(med))
            ach = b_bfs / (mean_ms / 1e3) / 1e9
            res[f"{lname}_{pname}"] = {"gteps": comp_edges / (mean_ms / 1e3) / 1e9, "mean_ms": mean_ms,
                                       "median_ms": float(np.median(med)), "achieved_gbs": ach,
                                       "frac": ach / HBM_PEAK_GBS}
            log(f"  gpu_do {lname} loop, {pname}: {mean_ms:.3f} ms per traversal, "
                f"{res[f'{lname}_{pname}']['gteps']:.1f} GTEPS, {ach / HBM_PEAK_GBS * 100:.1f}% of HBM")
    return res


def pr_workloads_part(G, O, ctx, iters=20, trials=3):
    """PageRank ms per iteration on the small RMAT, the scale-22 RMAT and the grid:
    gpu_pull_adaptive (in-degree bins) with the default options and with the
    autotuner's layout / order / bins, and gpu_pull_scalar (thread per vertex),
    `iters` fixed sweeps per run (HIP events, median of `trials`), as a fraction
    of HBM peak by 4(n+1) + 8m + 4n bytes per iteration; beside the reference's
    cpu_omp (pagerankOmp, OpenMP on this host's cores) per iteration."""
    rlib = O.ref_lib()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    out = {}
    for name in ("rmat16_c1", "rmat22_c2", "grid_c3"):
        g = G.Graph.named(name)
        n, m = g.n, g.m
        od = g.degrees()
        pr = G.GpuPageRank(g, od, ctx)  # undirected: the graph is its own transpose
        b_pr = 4 * (n + 1) + 8 * m + 4 * n
        ent = {"n": n, "m": m, "bytes_per_iteration": b_pr, "iterations_per_run": iters}
        tuned = pr.autotune()
        tuned.iters, tuned.tol = iters, 0.0
        for key, variant, o in (("gpu_pull_adaptive", "gpu_pull_adaptive", G.pr_options(iters=iters)),
                                ("gpu_pull_adaptive_tuned", "gpu_pull_adaptive", tuned),
                                ("gpu_pull_scalar", "gpu_pull_scalar", G.pr_options(iters=iters))):
            pr.run(variant, o)
            per = float(np.median([pr.run(variant, o).gpu_ms for _ in range(trials)])) / iters
            ach = b_pr / (per / 1e3) / 1e9
            ent[key] = {"ms_per_iteration": per, "achieved_gbs": ach, "frac": ach / HBM_PEAK_GBS}
        ent["tuned_options"] = {"layout": int(tuned.layout), "order": int(tuned.order),
                                "small_max": int(tuned.small_max), "large_min": int(tuned.large_min)}
        if rlib is not None:
            O.ref_pagerank(g.row_ptr, g.col_idx, od, 0.85, 1, 0.0, parallel=True)
            t1 = time.perf_counter()
            O.ref_pagerank(g.row_ptr, g.col_idx, od, 0.85, 5, 0.0, parallel=True)
            ent["cpu_omp"] = {"kind": "reference", "cores": cores,
                              "ms_per_iteration": (time.perf_counter() - t1) / 5 * 1e3}
        out[name] = ent
        log(f"  pagerank {name}: adaptive {ent['gpu_pull_adaptive']['ms_per_iteration'] * 1e3:.1f} us/it "
            f"({ent['gpu_pull_adaptive']['frac'] * 100:.1f}%), tuned "
            f"{ent['gpu_pull_adaptive_tuned']['ms_per_iteration'] * 1e3:.1f} us/it "
            f"({ent['gpu_pull_adaptive_tuned']['frac'] * 100:.1f}%, {ent['tuned_options']}), scalar "
            f"{ent['gpu_pull_scalar']['ms_per_iteration'] * 1e3:.1f} us/it"
            + (f", cpu_omp {ent['cpu_omp']['ms_per_iteration']:.2f} ms/it" if "cpu_omp" in ent else ""))
    return out


def loops_part(G, O, ctx, trials=3, modes=None):
    """gpu_do on the grid graph (>1500 levels) under each level loop: ms per traversal and us per level."""
    g = G.Graph.named("grid_c3")
    bfs = G.GpuBfs(g, ctx)
    src = [int(s) for s in g.pick_sources(4, 3)]
    out = {"workload": f"grid_c3: largest component of a 1024 x 2048 grid, edges kept with p = 0.7 (seed 3); "
                       f"n {g.n}, m {g.m}; gpu_do from {len(src)} pickSources sources (seed 3)"}
    all_modes = [("host", dict(loop=G.LOOP_HOST, time_levels=0), "gpu_do"),
                 ("graph_k8", dict(loop=G.LOOP_GRAPH, graph_levels=8), "gpu_do"),
                 ("graph_k64", dict(loop=G.LOOP_GRAPH, graph_levels=64), "gpu_do"),
                 ("persistent", dict(loop=G.LOOP_PERSISTENT), "gpu_do"),
                 ("hybrid", dict(loop=G.LOOP_HOST, time_levels=0), "hybrid")]
    if modes:
        unknown = set(modes) - {k for k, _, _ in all_modes}
        if unknown:
            raise SystemExit(f"unknown loop modes {sorted(unknown)}")
    ref = O.bfs_serial(g.row_ptr, g.col_idx, src[0])
    for key, kw, variant in all_modes:
        if modes and key not in modes:
            continue
        o = G.bfs_options(**kw)
        r = bfs.run(variant, src[0], o, trace=False)
        exact = bool(np.array_equal(r.dist, ref))
        times, levels = [], 0
        for s in src:
            for _ in range(trials):
                t = time.perf_counter()
                r = bfs.run(variant, s, o, trace=False, fetch=False)
                times.append((time.perf_counter() - t) * 1e3)
                levels = r.levels
        ms = float(np.median(times))
        out[key] = {"ms_per_traversal": ms, "levels": levels, "us_per_level": ms * 1e3 / max(levels, 1),
                    "bit_exact": exact}
        log(f"  grid_c3 {key}: {ms:.2f} ms per traversal, {levels} levels, {ms * 1e3 / max(levels, 1):.1f} us/level")
    return out


def pagerank_part(G, O, g, ctx, g_name="rmat23_c5", with_ref=True):
    """Time to L1 < 1e-4 against the FP64 oracle, and the per-sweep roofline."""
    n, m = g.n, g.m
    od = g.degrees()
    t = g if not g.directed else g.transpose()
    pr = G.GpuPageRank(t, od, ctx)  # undirected: the graph is its own transpose
    ref = None
    rlib = O.ref_lib() if with_ref else None
    t0 = time.time()
    if rlib is not None:
        ref, it, _ = O.ref_pagerank(t.row_ptr, t.col_idx, od, 0.85, 1000, 1e-10, parallel=True)
    log(f"FP64 converged oracle (the reference's pagerankOmp, tol 1e-10): {time.time() - t0:.1f}s")
    res = {}
    bytes_sweep = 4 * (n + 1) + 8 * m + 4 * n
    # per-sweep time of the fused Jacobi sweep per layout and vertex order (20 sweeps, no host sync)
    sweeps, combos = {}, {}
    for layout, lname, lmin in ((G.PR_LAYOUT_STREAM, "stream", 256), (G.PR_LAYOUT_HOT, "hot", 256),
                                (G.PR_LAYOUT_TILED, "tiled", 256), (G.PR_LAYOUT_HOT_TAIL, "hot_tail64", 64),
                                (G.PR_LAYOUT_HOT_TAIL, "hot_tail256", 256)):
        for order, oname in ((G.PR_ORDER_GIVEN, "given"), (G.PR_ORDER_DEGREE, "degree")):
            if layout == G.PR_LAYOUT_HOT_TAIL and order == G.PR_ORDER_GIVEN:
                continue  # built for the degree order (hot sources and hub rows first)
            name = f"{lname}/{oname}"
            fo = G.pr_options(iters=20, order=order, layout=layout, large_min=lmin)
            pr.run("gpu_pull_adaptive", fo)
            r = pr.run("gpu_pull_adaptive", fo)
            sweeps[name] = r.gpu_ms / 20
            combos[name] = (layout, order, lmin)
            log(f"  sweep ({name}): {sweeps[name] * 1e3:.1f} us")
    best_name = min(sweeps, key=sweeps.get)
    layout, order, lmin = combos[best_name]
    res["sweep_ms"] = sweeps
    # gpu_push (FP64 atomics along the out-arcs) on the degree-ordered copy, for comparison
    po = G.pr_options(iters=20, order=G.PR_ORDER_DEGREE)
    pr.run("gpu_push", po)
    res["push_sweep_ms"] = pr.run("gpu_push", po).gpu_ms / 20
    log(f"  sweep (gpu_push, degree): {res['push_sweep_ms'] * 1e3:.1f} us")
    # Jacobi, certified stop: ||r_k - r*||_1 <= d/(1-d) * delta_k  =>  delta_k < 1e-4 * (1-d)/d
    tol_cert = 1e-4 * 0.15 / 0.85
    for variant, opts, key in (("gpu_pull_adaptive", G.pr_options(tol=tol_cert, iters=1000, order=order,
                                                                  layout=layout, large_min=lmin), "jacobi_certified"),
                               ("gpu_pull_gs", G.pr_options(tol=1e-6, iters=1000, order=order), "gs_tol_1e-6")):
        pr.run(variant, opts)  # warm-up (builds the order)
        t1 = time.perf_counter()
        r = pr.run(variant, opts)
        ms = (time.perf_counter() - t1) * 1e3
        ent = {"ms": ms, "iters": r.iters, "gpu_ms": r.gpu_ms, "last_delta": r.last_delta, "order": best_name}
        if ref is not None:
            l1, rel, s = O.check_ranks(ref, pr.ranks())
            ent.update({"l1_vs_fp64": l1, "max_rel_top1000": rel, "sum": s, "l1_below_1e-4": bool(l1 < 1e-4)})
        res[key] = ent
    # The metric's own definition: k* = the first sweep count whose ranks are within
    # L1 1e-4 of the converged FP64 ranks (with the top-1000 and sum rules), found
    # against the oracle; the timed run is then k* fixed sweeps (no host sync between
    # sweeps). The certified runs above need no oracle and stop later.
    if ref is not None:
        for variant, key, mk in (("gpu_pull_adaptive", "jacobi_kstar",
                                  lambda k: G.pr_options(iters=k, order=order, layout=layout, large_min=lmin)),
                                 ("gpu_pull_gs", "gs_kstar", lambda k: G.pr_options(iters=k, order=order))):
            kstar = None
            for k in range(1, 61):
                pr.run(variant, mk(k))
                l1, rel, s = O.check_ranks(ref, pr.ranks())
                if l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4:
                    kstar = k
                    break
            if kstar is None:
                continue
            pr.run(variant, mk(kstar))
            t1 = time.perf_counter()
            r = pr.run(variant, mk(kstar))
            ms = (time.perf_counter() - t1) * 1e3
            l1, rel, s = O.check_ranks(ref, pr.ranks())
            res[key] = {"k": kstar, "ms": ms, "gpu_ms": r.gpu_ms, "l1_vs_fp64": l1, "max_rel_top1000": rel, "sum": s}
            log(f"  {key}: {kstar} sweeps, {ms:.2f} ms (L1 {l1:.2e})")
    # the reference's cpu_omp PageRank (pagerank_cpu.cpp:40, OpenMP on every host core) beside
    # the GPU: ms per iteration, and the same k* sweeps to L1 < 1e-4 (more if FP64 needs them)
    if rlib is not None:
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        t1 = time.perf_counter()
        O.ref_pagerank(t.row_ptr, t.col_idx, od, 0.85, 5, 0.0, parallel=True)
        per_it = (time.perf_counter() - t1) / 5 * 1e3
        k = res.get("jacobi_kstar", {}).get("k", 1)
        while k <= 60:
            t1 = time.perf_counter()
            rc, _, _ = O.ref_pagerank(t.row_ptr, t.col_idx, od, 0.85, k, 0.0, parallel=True)
            ms = (time.perf_counter() - t1) * 1e3
            l1, rel, s = O.check_ranks(ref, rc.astype(np.float32))
            if l1 < 1e-4 and rel < 1e-3 and abs(s - 1) < 1e-4:
                break
            k += 1
        res["cpu_omp_baseline"] = {"kind": "reference", "cores": cores, "ms_per_iteration": per_it,
                                   "iters_to_l1_1e-4": k, "ms_to_l1_1e-4": ms, "l1_vs_fp64": l1}
        log(f"  cpu_omp ({cores} cores): {per_it:.1f} ms/iteration, L1 < 1e-4 after {k} iterations in {ms:.1f} ms")
    per_sweep_ms = sweeps[best_name]
    ach = bytes_sweep / (per_sweep_ms / 1e3) / 1e9
    kname = {G.PR_LAYOUT_HOT: "k_pr_sweep_hot", G.PR_LAYOUT_TILED: "k_pr_sweep_tiled",
             G.PR_LAYOUT_HOT_TAIL: "k_pr_sweep_hot + k_pr_band + k_pr_tail + k_pr_tail_merge"}.get(layout, "k_pr_sweep")
    res["roofline"] = {"kernel": f"{kname} (gpu_pull_adaptive, Jacobi, {best_name} layout/order)",
                       "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(kname, g_name),
                       "bytes_per_launch": bytes_sweep,
                       "avg_launch_ms": per_sweep_ms}
    log(f"pagerank: certified Jacobi {res['jacobi_certified']['ms']:.2f} ms "
        f"({res['jacobi_certified']['iters']} it), GS {res['gs_tol_1e-6']['ms']:.2f} ms "
        f"({res['gs_tol_1e-6']['iters']} it), sweep {per_sweep_ms * 1e3:.1f} us = {ach:.0f} GB/s")
    return res


def ppr_part(G, g, ctx, g_name, B=64, sweeps=10, assign=None):
    """Personalized PageRank block of B seeds (pickSources, seed 5): work
    assignment and numbering from the autotuner, then the per-sweep time as the
    difference of a 2*sweeps and a sweeps run (the block's initialisation
    cancels). Every column must keep mass 1."""
    n, m = g.n, g.m
    od = g.degrees()
    t = g if not g.directed else g.transpose()
    pr = G.GpuPageRank(t, od, ctx)
    seeds = np.asarray(g.pick_sources(B, 5), dtype=np.uint32)
    t0 = time.perf_counter()
    if assign:
        order, large_min, lanes, tg = assign.split(":")
        o = G.pr_options(order=G.PR_ORDER_DEGREE if order == "degree" else G.PR_ORDER_GIVEN,
                         large_min=int(large_min), medium_lanes=int(lanes), tg=int(tg))
    else:
        o = pr.autotune_ppr(seeds, iters=3)
    tune_s = time.perf_counter() - t0
    o.tol = 0
    o.iters = sweeps
    r1 = pr.run_ppr(seeds, o)
    o.iters = 2 * sweeps
    r2 = pr.run_ppr(seeds, o)
    per = (r2.gpu_ms - r1.gpu_ms) / sweeps
    blk = pr.ppr_block()
    sums = blk.sum(axis=0, dtype=np.float64)
    mass_err = float(np.abs(sums - 1.0).max())
    seed_min = float(blk[seeds, np.arange(B)].min())
    del blk
    # rows of the in-CSR, the lists, out-degrees, one row gather of 4B bytes per
    # in-edge, rank and contribution rows written; beside it the per-iteration
    # count without the out-degree read and the contribution rows,
    # 4(n+1) + 4m + 4Bm + 4Bn
    bytes_sweep = 4 * (n + 1) + 4 * m + 4 * n + 4 * B * m + 8 * B * n
    bytes_min = 4 * (n + 1) + 4 * m + 4 * B * m + 4 * B * n
    ach = bytes_sweep / (per / 1e3) / 1e9
    res = {"workload": f"{g_name}: RMAT scale 23, ef 16, a/b/c .60/.17/.17; {B} pickSources seeds (seed 5)",
           "seeds": B, "sweeps": sweeps, "ms_per_sweep": per, "gpu_ms_2x_sweeps": r2.gpu_ms,
           "autotune_s": tune_s, "assignment": {"order": "degree" if o.order == G.PR_ORDER_DEGREE else "given",
                                               "large_min": o.large_min, "medium_lanes": o.medium_lanes, "tg": o.tg},
           "max_column_mass_error": mass_err, "min_seed_rank": seed_min,
           "bytes_4Bn_formula": bytes_min, "frac_4Bn_formula": bytes_min / (per / 1e3) / 1e9 / HBM_PEAK_GBS,
           "roofline": {"kernel": "k_ppr_chunks + k_ppr_sweep + k_ppr_reduce per sweep", "bound": "hbm",
                        "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                        "traffic": measured_traffic("k_ppr_chunks+k_ppr_sweep", g_name), "bytes_per_launch": bytes_sweep,
                        "avg_launch_ms": per}}
    log(f"ppr: {B} seeds, {per:.3f} ms/sweep = {ach:.0f} GB/s ({ach / HBM_PEAK_GBS * 100:.1f}% of HBM), "
        f"assignment {res['assignment']}, autotune {tune_s:.1f}s, column mass error {mass_err:.2e}")
    return res


def cpu_baseline(O, g, sources, comp_edges, trials=5, budget_s=15.0):
    """The reference's cpu_do (OpenMP, all host cores) timed with the reference's
    own protocol (per source one warm-up and `trials` timed runs, the median;
    GTEPS as the harmonic mean over sources, src/bench/main.cpp:308-370), on as
    many of the sources as fit a bounded budget of host time."""
    rlib = O.ref_lib()
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    rp, ci = g.row_ptr, g.col_idx
    kind = "reference" if rlib is not None else "port"

    def one(s):
        t = time.perf_counter()
        if rlib is not None:
            O.ref_bfs_cpu(rp, ci, int(s), "cpu_do", threads=threads, trace=False)
        else:
            O.bfs_do_levels(rp, ci, int(s))
        return time.perf_counter() - t

    medians = []
    t_start = time.time()
    for s in sources:
        one(s)
        medians.append(float(np.median([one(s) for _ in range(trials)])))
        if time.time() - t_start > budget_s:
            break
    per = float(np.mean(medians))  # the harmonic mean of per-source GTEPS (equal component edges)
    return {"value": comp_edges / per / 1e9, "unit": "GTEPS", "cores": threads if rlib is not None else 1,
            "kind": kind,
            "sample": f"cpu_do (alpha 15, beta 18), {len(medians)} of the {len(sources)} sources, 1 warm-up + "
                      f"{trials} timed runs each ({time.time() - t_start:.1f} s of host time); mean of the "
                      f"per-source medians {per * 1e3:.1f} ms per traversal"}


if __name__ == "__main__":
    main()
