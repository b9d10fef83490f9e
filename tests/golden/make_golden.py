#!/usr/bin/env python3
"""Writes tests/golden/ref_vectors.json from the reference's own CPU code.

Needs oracle/_ref/libggpu_ref.so, which oracle/Makefile compiles from the
reference checkout (graph generators, transforms, cpu_serial / cpu_do BFS,
pagerankSerial). Runs in the development container only; the fixture it
writes is plain data (inputs named by generator parameters, outputs as
numbers and SHA-1 digests of little-endian arrays), so the tests can pin the
oracle and the host library wherever the reference build is absent.

Usage: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402

SUITE = ["path1000", "star100k", "complete200", "two_components", "empty1000", "grid100x100", "rmat12",
         "rmat12_directed"]


def sha(a, dtype):
    return hashlib.sha1(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def sample_sources(n):
    s = [0, n // 2, n - 1]
    x = 12345
    for _ in range(3):
        x = (x * 6364136223846793005 + 1442695040888963407) & (2**64 - 1)
        s.append((x >> 33) % n)
    return s


def out_degrees(rp):
    return np.diff(rp).astype(np.uint32)


def acceptance_inputs():
    """Seeded (ref, got) pairs for checkRanks; tests/test_oracle_cpu.py rebuilds the same ones.
    Sizes around the top-1000 cut, and one vector with 60 equal reference values
    straddling rank 1000 whose errors differ."""
    out = []
    for n in (1, 7, 999, 1000, 1001, 4096):
        rng = np.random.default_rng(1000 + n)
        ref = rng.random(n) + 0.01
        ref /= ref.sum()
        got = (ref * (1 + rng.normal(0, 1e-4, n))).astype(np.float32)
        out.append((f"random{n}", ref, got))
    rng = np.random.default_rng(77)
    ref = np.concatenate([1.0 + rng.random(970), np.full(60, 0.5), 0.4 * rng.random(1970)])
    ref /= ref.sum()
    err = rng.normal(0, 1e-5, len(ref))
    err[970:1030] = np.linspace(1e-3, 2e-3, 60)[rng.permutation(60)]  # the tied values' errors differ
    got = (ref * (1 + err)).astype(np.float32)
    out.append(("ties_at_1000", ref, got))
    return out


def acceptance_points():
    """(l1, rel, sum) triples on and beside each threshold of PrCheck::ok and
    gsNoWorseThanJacobi, and Jacobi references for the latter."""
    vals = [0.0, 5e-5, np.nextafter(1e-4, 0), 1e-4, np.nextafter(1e-4, 1), 1.5e-4, 2e-4]
    rels = [0.0, np.nextafter(1e-3, 0), 1e-3, np.nextafter(1e-3, 1), 1.5e-3 * (1 + 1e-4), 2e-3]
    pts = [(l1, rel, 1.0 + ds) for l1 in vals for rel in rels for ds in (0.0, -9e-5, 1e-4, -1e-4, 1.5e-4, 2e-4)]
    jacs = [(5e-5, 5e-4, 1.0), (1.5e-4, 1.5e-3, 1.0 - 1e-5), (1e-4, 1e-3, 1.0)]
    return pts, jacs


def main():
    if O.ref_lib() is None:
        sys.exit("oracle/_ref/libggpu_ref.so is not built (make -C oracle)")
    fx = {"suite": {}, "rmat": [], "random_perm": [], "orderings": []}
    for name in SUITE:
        rp, ci, fl = O.ref_suite(name)
        n = len(rp) - 1
        srp, sci, sfl = O.ref_symmetrize(rp, ci, fl)
        entry = {"n": n, "m": int(len(ci)), "flags": fl, "row_ptr_sha1": sha(rp, "<u4"),
                 "col_idx_sha1": sha(ci, "<u4"), "sym_m": int(len(sci)), "sym_col_idx_sha1": sha(sci, "<u4"),
                 "bfs": []}
        for s in sample_sources(n):
            d = O.ref_bfs_serial(srp, sci, s)
            d2, levels = O.ref_bfs_cpu(srp, sci, s, "cpu_do", threads=1)
            assert np.array_equal(d, d2)
            entry["bfs"].append({"source": s, "dist_sha1": sha(d, "<u4"),
                                 "reached": int((d != O.INF).sum()),
                                 "max_depth": int(d[d != O.INF].max()),
                                 "do_levels": [[dr, fr, ed] for dr, fr, ed in levels]})
        trp, tci, _ = O.ref_transpose(rp, ci, fl)
        od = out_degrees(rp)
        r, it, delta = O.ref_pagerank(trp, tci, od, 0.85, 20, 0.0)
        idx = sorted(set(np.linspace(0, n - 1, 16).round().astype(int).tolist())) if n else []
        entry["pagerank20"] = {"iters": it, "last_delta": delta, "sum": float(r.sum()),
                               "at": {str(i): float(r[i]) for i in idx}, "max": float(r.max()) if n else 0.0}
        fx["suite"][name] = entry
        print(f"{name}: n={n} m={len(ci)}", flush=True)
    for scale, ef, directed, seed in [(10, 16, True, 1), (10, 16, False, 1), (12, 16, True, 43), (11, 8, False, 7)]:
        rp, ci, fl = O.ref_rmat(scale, ef, directed, seed)
        fx["rmat"].append({"scale": scale, "edgefactor": ef, "directed": directed, "seed": seed, "m": int(len(ci)),
                           "row_ptr_sha1": sha(rp, "<u4"), "col_idx_sha1": sha(ci, "<u4")})
    for n, seed in [(10, 1), (1000, 2), (65536, 5)]:
        p = O.ref_random_perm(n, seed)
        fx["random_perm"].append({"n": n, "seed": seed, "sha1": sha(p, "<u4"), "head": p[:8].tolist()})
    for name in ("two_components", "grid100x100", "rmat12", "star100k"):
        rp, ci, fl = O.ref_symmetrize(*O.ref_suite(name))
        for ordering, seed in (("random", 3), ("degree", 1), ("rcm", 1)):
            p = O.ref_order_perm(rp, ci, fl, ordering, seed)
            qrp, qci, qfl = O.ref_permute(rp, ci, fl, p)
            gap, same = O.ref_locality(qrp, qci, qfl)
            fx["orderings"].append({"graph": name, "ordering": ordering, "seed": seed, "sha1": sha(p, "<u4"),
                                    "head": p[:8].tolist(), "mean_gap": gap, "same_block64": same})
    fx["check_ranks"] = []
    for name, ref, got in acceptance_inputs():
        l1, mx, sm = O.ref_check_ranks(ref, got)
        fx["check_ranks"].append({"name": name, "l1": float.hex(l1), "max_rel": float.hex(mx), "sum": float.hex(sm)})
    pts, jacs = acceptance_points()
    fx["ranks_ok"] = [int(O.ref_ranks_ok(*p)) for p in pts]
    fx["gs_no_worse"] = [[int(O.ref_gs_no_worse(p, j)) for p in pts] for j in jacs]
    path = os.path.join(HERE, "ref_vectors.json")
    with open(path, "w") as f:
        json.dump(fx, f, indent=0, sort_keys=True)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
