#!/usr/bin/env python3
"""Trains the learned BFS direction policies on the GPU and writes their weights.

Training set: 40 seeded undirected RMAT graphs (scales 16-21, a in
[0.45, 0.65], b = c = (1 - a - 0.05) / 2, edge factor 4-16, ids permuted)
and 10 seeded grids (largest component of a 4-neighbour grid, each edge kept
with probability 0.6-0.9). For every graph the level recorders time each
level under all three actions (TD_EDGE, TD_BINNED, BU):
  * batched: the 64 pickSources sources (seed = graph seed) in one batch;
  * single source: gpu_do from 4 of those sources (host loop).
Targets are ln(ms). Grids contribute thousands of nearly identical levels,
so at most 256 evenly spaced levels per recording are kept. Each slot's MLP
is trained by the library's own GPU kernels (full-batch Adam) from a seeded
He initialisation; a held-out set of 6 more graphs reports regret = time of
the chosen actions / time of the best actions, and the end-to-end time of the
learned policy against the GAP heuristic: the 64-source batch, and gpu_do
from 4 sources under the host and the persistent loops.

Usage (needs the GPU):
    python -m graph_analytics_applesilicon_amd.policy.train_policy --out DIR [--quick]
writes DIR/batched.f32, DIR/single_source.f32 and DIR/train_report.json; the
package loads graph_analytics_applesilicon_amd/policy/{batched,single_source}.f32.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import graph_analytics_applesilicon_amd as G  # noqa: E402

MAX_LEVELS_PER_RECORD = 256


def log(msg):
    print(msg, flush=True)


def rmat_spec(seed):
    rng = np.random.default_rng(seed)
    a = float(rng.uniform(0.45, 0.65))
    bc = (1.0 - a - 0.05) / 2
    return dict(scale=int(16 + seed % 6), ef=int(rng.choice([4, 8, 12, 16])), a=a, b=bc, c=bc, seed=seed)


def grid_spec(seed):
    rng = np.random.default_rng(seed)
    return dict(rows=int(rng.choice([256, 384, 512, 768])), cols=int(rng.choice([512, 768, 1024])),
                keep=float(rng.uniform(0.6, 0.9)), seed=seed)


def make_graph(kind, spec):
    if kind == "rmat":
        g = G.Graph.rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"], directed=False,
                         seed=spec["seed"])
        return g.permute(G.random_perm(g.n, spec["seed"]))
    g = G.Graph.grid(spec["rows"], spec["cols"], spec["keep"], spec["seed"])
    lcc, _ = g.largest_component()
    return lcc


def thin(feat, times, limit=MAX_LEVELS_PER_RECORD):
    if len(feat) <= limit:
        return feat, times
    idx = np.unique(np.linspace(0, len(feat) - 1, limit).round().astype(int))
    return feat[idx], times[idx]


def record(kind, spec, singles=4):
    g = make_graph(kind, spec)
    bfs = G.GpuBfs(g)
    src = np.array(g.pick_sources(64, spec["seed"]), dtype=np.uint32)
    out = {"batched": [], "single": []}
    fb, tb, _ = bfs.record_batch_levels(src)
    out["batched"].append(thin(fb, tb))
    heur_b = bfs.run_batch(src, G.bfs_options(), fetch=False)[1].gpu_ms
    out["heuristic_batched_ms"] = heur_b
    for s in src[:singles]:
        fs, ts, _ = bfs.record_levels(int(s))
        out["single"].append(thin(fs, ts))
    out["n"], out["m"] = g.n, g.m
    return out


def stack(recs, key):
    fs = [f for r in recs for (f, _) in r[key]]
    ts = [t for r in recs for (_, t) in r[key]]
    return np.concatenate(fs).astype(np.float32), np.concatenate(ts).astype(np.float32)


def regret(params, feat, times):
    y = G.mlp_forward(params, feat)
    pick = np.argmin(y, axis=1)
    chosen = times[np.arange(len(times)), pick].sum()
    best = times.min(axis=1).sum()
    acc = float(np.mean(pick == np.argmin(times, axis=1)))
    return float(chosen / best), acc


def train_slot(feat, times, seed, epochs, lr):
    t = np.log(np.maximum(times, 1e-3)).astype(np.float32)
    best = None
    for restart in range(3):
        p0 = G.mlp_init(seed + restart)
        p, loss = G.mlp_train(p0, feat, t, epochs=epochs, lr=lr)
        if best is None or loss < best[1]:
            best = (p, loss)
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--quick", action="store_true", help="4 RMAT + 1 grid, for a smoke run")
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)

    n_rmat, n_grid = (4, 1) if args.quick else (40, 10)
    train = [("rmat", rmat_spec(100 + i)) for i in range(n_rmat)] + \
            [("grid", grid_spec(500 + i)) for i in range(n_grid)]
    held = [("rmat", rmat_spec(900 + i)) for i in range(4)] + [("grid", grid_spec(950 + i)) for i in range(2)]
    if args.quick:
        held = held[:1] + held[-1:]

    recs, hrecs = [], []
    t0 = time.time()
    for i, (kind, spec) in enumerate(train + held):
        r = record(kind, spec)
        (recs if i < len(train) else hrecs).append(r)
        log(f"[{time.time() - t0:6.1f}s] {'train' if i < len(train) else 'held'} {kind} {spec} n={r['n']} "
            f"m={r['m']} levels(batched)={len(r['batched'][0][0])}")

    report = {"train_graphs": len(train), "held_out_graphs": len(held)}
    for slot, key, name in ((G.POLICY_BATCHED, "batched", "batched.f32"),
                            (G.POLICY_SINGLE, "single", "single_source.f32")):
        feat, times = stack(recs, key)
        hf, ht = stack(hrecs, key)
        p, loss = train_slot(feat, times, seed=7 + slot, epochs=args.epochs, lr=args.lr)
        tr_reg, tr_acc = regret(p, feat, times)
        ho_reg, ho_acc = regret(p, hf, ht)
        p.astype(np.float32).tofile(os.path.join(args.out, name))
        report[key] = {"samples": int(len(feat)), "held_out_samples": int(len(hf)), "final_loss": loss,
                       "train_regret": tr_reg, "train_accuracy": tr_acc,
                       "held_out_regret": ho_reg, "held_out_accuracy": ho_acc}
        log(f"{key}: {len(feat)} samples, loss {loss:.4f}, regret train {tr_reg:.3f} held-out {ho_reg:.3f}, "
            f"accuracy train {tr_acc:.2f} held-out {ho_acc:.2f}")

    # end-to-end on the held-out graphs: batched BFS time, heuristic vs learned
    ctx = G.default_context()
    G.policy_set(np.fromfile(os.path.join(args.out, "batched.f32"), dtype=np.float32), ctx, G.POLICY_BATCHED)
    rows = []
    for kind, spec in held:
        g = make_graph(kind, spec)
        bfs = G.GpuBfs(g)
        src = np.array(g.pick_sources(64, spec["seed"]), dtype=np.uint32)
        ms = {}
        for pol in (G.POLICY_HEURISTIC, G.POLICY_LEARNED):
            o = G.bfs_options(policy=pol)
            bfs.run_batch(src, o, fetch=False)
            ms[pol] = float(np.median([bfs.run_batch(src, o, fetch=False)[1].gpu_ms for _ in range(5)]))
        rows.append({"graph": f"{kind} {spec}", "heuristic_ms": ms[G.POLICY_HEURISTIC],
                     "learned_ms": ms[G.POLICY_LEARNED]})
        log(f"held-out {kind} {spec}: heuristic {ms[G.POLICY_HEURISTIC]:.3f} ms, learned "
            f"{ms[G.POLICY_LEARNED]:.3f} ms")
    report["held_out_batched_ms"] = rows

    # end-to-end single-source gpu_do on the held-out graphs, host and persistent loops
    G.policy_set(np.fromfile(os.path.join(args.out, "single_source.f32"), dtype=np.float32), ctx, G.POLICY_SINGLE)
    rows = []
    for kind, spec in held:
        g = make_graph(kind, spec)
        bfs = G.GpuBfs(g)
        src = [int(s) for s in g.pick_sources(64, spec["seed"])[:4]]
        row = {"graph": f"{kind} {spec}"}
        for loop, lname in ((G.LOOP_HOST, "host"), (G.LOOP_PERSISTENT, "persistent")):
            for pol, pname in ((G.POLICY_HEURISTIC, "heuristic"), (G.POLICY_LEARNED, "learned")):
                o = G.bfs_options(policy=pol, loop=loop)
                tot = 0.0
                for s in src:
                    bfs.run("gpu_do", s, o, trace=False, fetch=False)
                    tot += float(np.median([bfs.run("gpu_do", s, o, trace=False, fetch=False).gpu_ms
                                            for _ in range(3)]))
                row[f"{lname}_{pname}_ms"] = tot
        rows.append(row)
        log(f"held-out single {kind} {spec}: " + ", ".join(f"{k} {v:.3f}" for k, v in row.items() if k != "graph"))
    report["held_out_single_ms"] = rows
    with open(os.path.join(args.out, "train_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    log(f"done in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main()
