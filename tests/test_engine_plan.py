"""The hip engine's fuser and code generator on the CPU: run programs with the
engine in compile-only mode (BH_HIP_DRY_RUN=1, every kernel is planned and
compiled by hiprtc for gfx950, nothing is launched) and check the plans."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
import json, sys
import numpy as np
sys.path.insert(0, %(root)r)
import bohrium_amd as bh
from bohrium_amd import workloads
plans = {}
def plan(name, fn):
    fn()
    bh.flush()
    plans[name] = [l.split()[0] for l in bh.message("plan log").splitlines() if l.strip()]
plan("mc_pi", lambda: workloads.mc_pi(10 ** 6))
S, K, T = workloads.black_scholes_inputs(1000)
bh.flush(); bh.message("plan log")
plan("bs_step", lambda: workloads.black_scholes_step(S, K, T, 3))
G = workloads.jacobi_grid(64)
bh.flush(); bh.message("plan log")
plan("jacobi_step", lambda: workloads.jacobi_step(G))
G2 = workloads.jacobi_grid(64)
bh.flush(); bh.message("plan log")
plan("jacobi_repeat", lambda: workloads.jacobi_repeat(G2, 3))
S2, K2, T2 = workloads.black_scholes_inputs(1000)
bh.flush(); bh.message("plan log")
plan("bs_repeat", lambda: workloads.black_scholes_repeat(S2, K2, T2, 4))
plan("scan", lambda: (bh.arange(1000) > 5).astype(np.int64).cumsum())
plan("row_sum", lambda: bh.ones((100, 50)).sum(axis=1))
plan("col_sum", lambda: bh.ones((100, 50)).sum(axis=0))
plan("cumsum_axis0", lambda: bh.ones((100, 50)).cumsum(axis=0))
plan("gather", lambda: bh.take(bh.arange(50).astype(np.float32), bh.array(np.arange(10) * 3)))
# a random program whose column and row sums of one shape the fusers once put in one loop nest
from bohrium_amd import _bh_api
_bh_api.set_flush_threshold(10 ** 7)
for fl in ("greedy, ", "serial, "):
    bh.message("fuser_list: " + fl + "push_reductions_inwards, collapse_redundant_axes")
    outs = workloads.random_dag(1001, 337, sides=(128, 256))
    bh.asnumpy_many([o[:1] if o.ndim == 1 else o[:1, :1] for o in outs])
    del outs
    plans["dag_" + fl.split(",")[0]] = [l.split()[0] for l in bh.message("plan log").splitlines() if l.strip()]
print("PLANS=" + json.dumps(plans))
"""


@pytest.fixture(scope="module")
def plans():
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    r = subprocess.run([sys.executable, "-c", DRIVER % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("PLANS=")][0]
    return json.loads(line[6:])


def test_mc_pi_is_one_fused_reduction(plans):
    """philox x2, casts, x*x+y*y, compare and the int64 sum run as one launch."""
    assert plans["mc_pi"] == ["reduce"]


def test_black_scholes_step_is_one_fused_reduction(plans):
    assert plans["bs_step"] == ["reduce"]


def test_jacobi_step_plan(plans):
    """The flattened max over |tmp - c| is restated on the stencil's 2-d shape, so the
    stencil, the delta and its max are one reduction; the copy back into the grid's
    interior overlaps the stencil's reads and is a second launch."""
    assert plans["jacobi_step"] == ["reduce", "map"]


def test_repeat_bodies_fold_scalar_tails(plans):
    """The scalar stores after a reduction (delta[...] = ..., sums[i] = ...) run in the
    reduction's finalize kernel, so each repeat is one reduce (+ the Jacobi copy)."""
    assert plans["jacobi_repeat"] == ["reduce", "map"] * 3
    # (the first launch builds the time offsets before the loop)
    assert plans["bs_repeat"] == ["map"] + ["reduce"] * 4


def test_scan_plan(plans):
    assert plans["scan"] == ["scan"]


def test_axis_reductions_run_on_parallel_skeletons(plans):
    # a row sum is one wave per row (lanes stride the row); a column sum a map over columns
    assert plans["row_sum"] == ["row"]
    assert "serial" not in plans["col_sum"]
    assert "serial" not in plans["cumsum_axis0"]


def test_gather_plan(plans):
    assert "serial" not in plans["gather"]


def test_fusers_keep_random_programs_parallel(plans):
    """Sweeps along two axes of one shape are never fused into one loop nest (it would run on one thread)."""
    for name in ("dag_greedy", "dag_serial"):
        assert plans[name], name
        assert "serial" not in plans[name], (name, plans[name])


GRID_DRIVER = r"""
import sys
sys.path.insert(0, %(root)r)
import bohrium_amd as bh
from bohrium_amd import workloads
def reduce_grids(fn):
    fn()
    bh.flush()
    return [int(l.split()[1]) for l in bh.message("plan log").splitlines() if l.startswith("reduce")]
G = workloads.jacobi_grid(8192)
bh.flush(); bh.message("plan log")
print("JACOBI", reduce_grids(lambda: workloads.jacobi_step(G)))
S, K, T = workloads.black_scholes_inputs(20_000_000)
bh.flush(); bh.message("plan log")
print("BS", reduce_grids(lambda: workloads.black_scholes_step(S, K, T, 0)))
"""


def test_light_reductions_take_half_the_workgroup_cap():
    """Full reductions with a light loop body (the Jacobi stencil + max|delta|) use half of
    reduce_max_grid (4096) workgroups, heavy ones (a Black-Scholes step) the whole cap; measured on
    MI355X, each is the faster choice for its kind."""
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    r = subprocess.run([sys.executable, "-c", GRID_DRIVER % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:]
    out = dict(l.split(" ", 1) for l in r.stdout.splitlines() if l.startswith(("JACOBI", "BS")))
    assert json.loads(out["JACOBI"]) == [2048]
    assert json.loads(out["BS"]) == [4096]


def test_full_reduce_main_loop_is_the_first_sweeping_loop():
    """With small sweeping tail blocks after it, a FULL_REDUCE launch is sized by its main loop
    (5000 elements: 5 workgroups of 256 threads x 4 elements), not by the last tail loop."""
    prog = r"""
import sys
sys.path.insert(0, %(root)r)
import numpy as np
import bohrium_amd as bh
from bohrium_amd import workloads, _bh_api as api
api.set_flush_threshold(10 ** 7)
bh.message("fuser_list: reshapable_first, push_reductions_inwards, collapse_redundant_axes")
sizes = np.array([5000, 200, 3000, 100])
xs, idxs = workloads.batched_inputs(sizes)
bh.flush()
bh.message("plan log")
res = workloads.batched_step(xs, idxs)
bh.flush()
print(bh.message("plan log"))
""" % {"root": ROOT}
    r = subprocess.run([sys.executable, "-c", prog], env=dict(os.environ, BH_HIP_DRY_RUN="1"),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    grids = [int(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("reduce ")]
    # the 5000- and 3000-element main loops size their launches (5 and 3 workgroups); sized by a
    # tail loop they would get 1
    assert grids and 5 in grids and 3 in grids, r.stdout


BATCH_DRIVER = r"""
import sys
import numpy as np
sys.path.insert(0, %(root)r)
import bohrium_amd as bh
from bohrium_amd import workloads, _bh_api
_bh_api.set_flush_threshold(10 ** 7)
bh.message("batch_persistent: on")
bh.message("batch_xcds: 8")
sizes = workloads.batched_sizes(40, 256, 40000, seed=7)
xs, idxs = workloads.batched_inputs(sizes, seed=7)
bh.flush()
sums, cums, gathered = workloads.batched_step(xs, idxs)
bh.flush()
print("SIZES=" + ",".join(str(int(n)) for n in sizes))
"""


def test_batched_scan_tiles_are_placed_per_xcd(tmp_path):
    """A batched scan launch (40 uneven arrays, 1-20 tiles of 2048 elements each) gives every
    workgroup an item or none: each item gets exactly as many workgroups as it has tiles, all with the
    same workgroup number mod 8 (the XCD the workgroup goes to, whose L2 then holds the item's gather
    source), the XCDs' tile counts differ by at most one item's, and the workgroups without an item
    only pad the XCDs to the same count."""
    env = dict(os.environ, BH_HIP_DRY_RUN="1", BH_HIP_DUMP_DIR=str(tmp_path))
    r = subprocess.run([sys.executable, "-c", BATCH_DRIVER % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    sizes = [int(v) for v in [l for l in r.stdout.splitlines() if l.startswith("SIZES=")][0][6:].split(",")]
    scans = []
    for f in sorted(tmp_path.glob("batch_*.txt")):
        fields = {}
        for line in f.read_text().splitlines():
            if line == "code":
                break
            k, _, v = line.partition(" ")
            fields[k] = v
        if fields["strategy"] == "scan":
            scans.append(fields)
    assert len(scans) == 1
    s = scans[0]
    items, blocks = int(s["items"]), int(s["blocks"])
    assert int(s["items_arg"]) == blocks and blocks % 8 == 0
    table = [int(v) for v in s["table"].split()]
    assert items == len(sizes) and len(table) == blocks + items
    counts = table[blocks:]
    assert sorted(counts) == sorted((n + 2047) // 2048 for n in sizes)
    seen = [0] * items
    xcd = {}
    load = [0] * 8
    for b in range(blocks):
        it = table[b]
        if it == 0xFFFFFFFF:
            continue
        seen[it] += 1
        assert xcd.setdefault(it, b % 8) == b % 8
        load[b % 8] += 1
    assert seen == counts
    assert max(load) - min(load) <= max(counts)
    assert blocks == 8 * max(load)


def test_persistent_plan_covers_every_launch_in_dependency_order(tmp_path):
    """The batched schedule as one persistent launch (40 uneven arrays: map+sum, then scan+gather of
    the map's output): every workgroup of every plan is one task (grid-stride plans with fewer, longer
    workgroups than their own grid, scans with one per tile), a full reduction's finalize comes in
    a later level than all its main workgroups, the scans (which gather the map's output) in a later
    level than every map workgroup, and each scan item's tiles share one XCD (position in the level
    mod 8)."""
    env = dict(os.environ, BH_HIP_DRY_RUN="1", BH_HIP_DUMP_DIR=str(tmp_path))
    r = subprocess.run([sys.executable, "-c", BATCH_DRIVER % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    plans = []
    for f in sorted(tmp_path.glob("persist_*.txt")):
        fields = {}
        for line in f.read_text().splitlines():
            if line == "code":
                break
            k, _, v = line.partition(" ")
            fields[k] = v
        if "reduce" in fields.get("code_strategies", ""):
            plans.append(fields)
    assert len(plans) == 1
    p = plans[0]
    strategies = p["code_strategies"].split()
    grids = [int(v) for v in p["grids"].split()]
    nblk = [int(v) for v in p["nblk"].split()]
    arg_off = [int(v) for v in p["arg_off"].split()]
    levels = [int(v) for v in p["levels"].split()]
    words = [int(v) for v in p["tasks"].split()]
    tasks = [tuple(words[i:i + 4]) for i in range(0, len(words), 4)]
    assert levels[0] == 0 and levels[-1] == len(tasks) and levels == sorted(levels)
    plan_of = {off // 16: i for i, off in enumerate(arg_off)}
    level_of = lambda t: max(L for L in range(len(levels) - 1) if levels[L] <= t)
    mains, fins = {}, {}
    for t, (case, b, nb, off16) in enumerate(tasks):
        if case == 0xFFFFFFFF:
            continue
        i = plan_of[off16]
        assert nb == (1 if case % 2 else nblk[i]) and 1 <= nblk[i] <= grids[i]
        (fins if case % 2 else mains).setdefault(i, []).append((t, b, strategies[case // 2]))
    assert sorted(mains) == list(range(len(grids)))
    for i, ts in mains.items():
        kind = ts[0][2]
        assert len(ts) == nblk[i]
        if kind != "scan":
            assert sorted(b for _, b, _ in ts) == list(range(nblk[i]))
        else:
            assert nblk[i] == grids[i]
            assert len({(t - levels[level_of(t)]) % 8 for t, _, _ in ts}) == 1
    reduce_plans = [i for i, ts in mains.items() if ts[0][2] == "reduce"]
    scan_plans = [i for i, ts in mains.items() if ts[0][2] == "scan"]
    assert len(reduce_plans) == len(scan_plans) == 40 and sorted(fins) == sorted(reduce_plans)
    last_map_level = max(level_of(t) for i in reduce_plans for t, _, _ in mains[i])
    for i in reduce_plans:
        assert level_of(fins[i][0][0]) > last_map_level
    # the gathers y[idx] read the maps' results: they run after them (in maps of their own: they share
    # no array with the cumsums, which read only the inputs and may run at any level)
    gather_plans = [i for i, ts in mains.items() if ts[0][2] == "map"]
    assert len(gather_plans) == 40
    for i in gather_plans:
        assert min(level_of(t) for t, _, _ in mains[i]) > last_map_level
