"""In-place updates through a temporary (`c[...] = f(views of G)`, c a view of G) run as a buffer
exchange with a twin of G: the new values are computed straight into the twin, which then becomes
G's buffer (rewrite_in_place_updates).

CPU (compile-only engine): the launches planned per sweep (the copy-back map disappears, the full
copy G -> twin is planned only when the twin is not known to equal G outside the update).
GPU: results against the oracle on the same bytecode for every way the twin can fall out of step:
the first update, an update after another write of G in the same flush, reads and writes of G after
the update in the same flush, host writes between flushes, syncs, and repeat bodies replayed as
period-2 graphs with an odd and an even number of iterations."""
import os
import subprocess
import sys

import numpy as np
import pytest

import bohrium_amd as bh
from bohrium_amd import workloads
from tests.helpers import run_both, assert_parity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 512  # float64 512 x 512 = 2 MiB, above the engine's twin threshold of 1 MiB


def stat(name):
    for line in bh.message("statistic").splitlines():
        if line.startswith(name + ":"):
            return int(float(line.split(":")[1]))
    raise KeyError(name)


PLANS = r"""
import sys
sys.path.insert(0, %(root)r)
import bohrium_amd as bh
from bohrium_amd import workloads
G = workloads.jacobi_grid(1024)
bh.flush()
bh.message("plan log")
for it in range(3):
    d = workloads.jacobi_step(G)
    bh.flush()
    print("sweep", it, " ".join(l.split()[0] for l in bh.message("plan log").splitlines() if l.strip()))
G[5, 5] = 1.0   # another write of G: the twin falls out of step
bh.flush()
bh.message("plan log")
d = workloads.jacobi_step(G)
bh.flush()
print("after-write", " ".join(l.split()[0] for l in bh.message("plan log").splitlines() if l.strip()))
for l in bh.message("statistic").splitlines():
    if l.startswith("twin_"):
        print(l.replace(":", ""))
"""


def test_planned_launches_per_sweep():
    r = subprocess.run([sys.executable, "-c", PLANS % {"root": ROOT}], env=dict(os.environ, BH_HIP_DRY_RUN="1"),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    out = {l.split(" ", 1)[0] + (l.split()[1] if l.startswith("sweep") else ""): l for l in r.stdout.splitlines()}
    # first sweep: full copy G -> twin (a map) and the stencil + max|delta| (a full reduction)
    assert out["sweep0"].split()[2:] == ["map", "reduce"], r.stdout
    # then the stencil alone: no copy-back map, no copy
    assert out["sweep1"].split()[2:] == ["reduce"], r.stdout
    assert out["sweep2"].split()[2:] == ["reduce"], r.stdout
    assert out["after-write"].split()[1:] == ["map", "reduce"], r.stdout
    assert "twin_exchanges 4" in r.stdout and "twin_full_copies 2" in r.stdout, r.stdout


def _grid(seed=0):
    rng = np.random.RandomState(seed)
    return rng.uniform(-1, 1, (N, N))


@pytest.mark.gpu
def test_sweeps_in_separate_flushes():
    g0 = _grid(1)

    def prog():
        G = bh.array(g0)
        ds = []
        for _ in range(5):
            ds.append(workloads.jacobi_step(G))
            bh.flush()
        return [G] + ds

    before = stat("twin_exchanges")
    got, want = run_both(prog)
    assert stat("twin_exchanges") >= before + 5
    for g, w in zip(got, want):
        assert_parity(g, w, np.float64, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_writes_and_reads_around_the_update_in_one_flush():
    g0 = _grid(2)

    def prog():
        G = bh.array(g0)
        bh.flush()
        out = []
        for it in range(4):
            G[0, :] = float(it)            # a write of G before the update: full copy first
            c = G[1:-1, 1:-1]
            tmp = (c + G[:-2, 1:-1] + G[2:, 1:-1] + G[1:-1, :-2] + G[1:-1, 2:]) * 0.2
            c[...] = tmp
            del tmp
            out.append(G[2:5, 2:5] * 2.0)  # a read of G after the update, in the same flush
            if it == 2:
                G[-1, 3] = 7.0             # a write after the update: the twin is out of step next time
            bh.flush()
        return [G] + out

    got, want = run_both(prog)
    for g, w in zip(got, want):
        assert_parity(g, w, np.float64, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_host_writes_and_syncs_between_updates():
    g0 = _grid(3)
    G = bh.array(g0)
    ref = g0.copy()

    def step_np(a):
        a = a.copy()
        a[1:-1, 1:-1] = (a[1:-1, 1:-1] + a[:-2, 1:-1] + a[2:, 1:-1] + a[1:-1, :-2] + a[1:-1, 2:]) * 0.2
        return a

    for it in range(6):
        workloads.jacobi_step(G)
        bh.flush()
        ref = step_np(ref)
        if it % 2 == 0:
            host = G.asnumpy()                   # sync: the host copy of the current buffer
            np.testing.assert_allclose(host, ref, rtol=1e-12, atol=1e-12)
        if it == 3:
            new = ref * 0.5                      # a host write: the twin is out of step
            G[...] = bh.array(new)
            ref = new
    np.testing.assert_allclose(G.asnumpy(), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("iters", [9, 10])
def test_repeat_body_replayed_as_period_two_graph(iters):
    g0 = _grid(4).astype(np.float32)

    def prog():
        G = bh.array(g0)
        d = workloads.jacobi_repeat(G, iters)
        return [G, d]

    r0 = int(bh.message("graph replays"))
    got, want = run_both(prog)
    assert int(bh.message("graph replays")) > r0, "the repeat body was not replayed as a graph"
    for g, w in zip(got, want):
        assert_parity(g, w, np.float32, rtol=1e-5, atol=1e-5)
