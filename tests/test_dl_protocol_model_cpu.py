"""Host model of the persistent single-source BFS loop's control protocol.

tests/model/dl_protocol_model.cpp runs 256 workgroups in 8 cache groups (one
per XCD) as coroutines under a seeded adversarial scheduler, with the
product's own decision code (fold_level / decide_action, sb_ok, the command
word of csrc/hip/dl_protocol.hpp). Two protocols:

* old: every workgroup decides from the per-block counters in part slot
  (level + 1) & 1; workgroup 0 runs small top-down levels alone and hands the
  state back with plain stores, a release fence and a flag;
* new: workgroup 0 alone decides and publishes commands; every cross-workgroup
  datum is written through and read coherently.

What the model shows:
* with writebacks that land after the flag is seen, the old protocol ends in
  the timeout gbench reported on rmat12 (group arrivals 96 128 128 128 128 128
  128 128, top 31, generation 3, episodes 2, level 5): every workgroup of
  group 0 saw the traversal end, the other groups ran one more grid level;
* with textbook fences the old protocol still diverges when an odd number of
  single-workgroup levels is followed by a grid level (its part slot is
  rewritten while slower workgroups still sum it);
* the new protocol finishes with the host loop's decision sequence in every
  case, in both memory modes, with and without stalls.
"""
import json
import os
import subprocess

import pytest

import graph_analytics_applesilicon_amd as G
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "model", "dl_protocol_model.cpp")
HIPCC = "/opt/rocm/bin/hipcc"

# An odd-length run of single-workgroup levels (level 1) between grid levels 0
# and 2: level 2's counters land in the part slot level 1's decision reads.
SLOT_REUSE = "4000001 100000000\n1 20000\n3000 15000\n2000000 50000000\n1000000 40000000\n997000 9960000\n0 0\n"


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("model") / "dl_protocol_model")
    inc = ["-I" + os.path.join(ROOT, p) for p in ("include", "graph_analytics_applesilicon_amd/csrc/host",
                                                  "graph_analytics_applesilicon_amd/csrc/hip")]
    subprocess.run([HIPCC, "-std=c++20", "-O2", "--cuda-host-only", "-x", "hip", *inc, SRC, "-o", exe], check=True,
                   capture_output=True)
    return exe


def profile(g, src, script=None):
    """n m, then (frontier size, frontier edges) per level, ending with 0 0."""
    deg = g.degrees()
    d = O.bfs_serial(g.row_ptr, g.col_idx, int(src))
    last = int(d[d != O.INF].max())
    rows = [f"{int((d == l).sum())} {int(deg[d == l].sum())}" for l in range(last + 1)] + ["0 0"]
    text = f"{g.n} {g.m}\n" + "\n".join(rows) + "\n"
    if script:
        text += "script " + " ".join(str(a) for a in script) + "\n"
    return text


def run(model, text, *args):
    r = subprocess.run([model, *map(str, args)], input=text, capture_output=True, text=True, check=True)
    return [json.loads(x) for x in r.stdout.splitlines()]


@pytest.fixture(scope="module")
def profiles():
    rmat = G.Graph.named("rmat12").symmetrize()
    grid = G.Graph.named("grid100x100").symmetrize()
    star = G.Graph.named("star100k").symmetrize()
    out = {f"rmat12:{s}": profile(rmat, s) for s in rmat.pick_sources(3, 1)}
    out["grid100x100:0"] = profile(grid, 0)
    out["grid100x100:alternate"] = profile(grid, 0, script=[2, 0])  # bottom-up / top-down: a hand-off per level
    out["star100k:leaf"] = profile(star, 1)
    out["slot-reuse"] = SLOT_REUSE
    return out


def test_new_protocol_finishes_with_the_host_loop_decisions(model, profiles):
    for name, text in profiles.items():
        for mode in ("textbook", "late"):
            for extra in ((), ("--stall", 0.02, "--stall-steps", 200000)):
                res = run(model, text, "--protocol", "new", "--mode", mode, "--seed", 1, "--runs", 6, "--spin", 20000,
                          *extra)
                bad = [r for r in res if r["outcome"] != "ok"]
                assert not bad, (name, mode, extra, bad[:1])
                # every workgroup ran workgroup 0's grid levels; nothing arrived beyond them
                r = res[0]
                assert r["top"] == 8 * r["grid_levels"], (name, r)
                assert r["group_arrivals"] == [32 * r["grid_levels"]] * 8, (name, r)


def test_old_protocol_with_late_writeback_reproduces_the_gbench_timeout(model, profiles):
    (r,) = run(model, profiles["rmat12:2320"], "--protocol", "old", "--mode", "late", "--drain", 0.002,
               "--seed", 1075, "--spin", 20000)
    assert r["outcome"] == "timeout"
    assert r["group_arrivals"] == [96] + [128] * 7
    assert (r["top"], r["generation"], r["episodes"], r["level"]) == (31, 3, 2, 5)
    assert r["exited_groups"] == [0]
    # the class, over seeds: some runs end that way, none finishes wrong without a timeout
    res = run(model, profiles["rmat12:2320"], "--protocol", "old", "--mode", "late", "--drain", 0.05, "--seed", 100,
              "--runs", 40, "--spin", 20000)
    assert any(x["outcome"] == "timeout" and x["episodes"] == 2 and len(x["exited_groups"]) in (1, 7) for x in res)
    assert all(x["outcome"] in ("ok", "timeout") for x in res)


def test_old_protocol_is_sound_on_rmat12_only_with_prompt_writeback(model, profiles):
    for name in ("rmat12:2320", "rmat12:1820", "rmat12:1819"):
        res = run(model, profiles[name], "--protocol", "old", "--mode", "textbook", "--seed", 1, "--runs", 6,
                  "--spin", 20000)
        assert all(r["outcome"] == "ok" for r in res), (name, res[:1])


def test_old_protocol_rewrites_a_part_slot_while_it_is_read(model, profiles):
    # (on grid100x100 with alternating actions the rewritten counters are too
    # small to change a branch; the slot-reuse profile's are not)
    res = run(model, profiles["slot-reuse"], "--protocol", "old", "--mode", "textbook", "--seed", 1, "--runs", 6,
              "--spin", 20000)
    assert sum(r["outcome"] != "ok" for r in res) >= 3, [r["outcome"] for r in res]
