"""The autotuner (MI355X): kernels of at least 2^20 elements are launched under each distinct
candidate launch shape until every candidate has its samples, then under the fastest.

Every candidate must compute the same results: maps, integer reductions and scans bit for bit;
float sums within n * eps * sum|x| of the configured shape's result (a different grid is a
different summation order).  The last iteration is also checked against the oracle.  A second
process with the same table file takes the choice from the table without trials."""
import os
import subprocess
import sys

import numpy as np
import pytest

import bohrium_amd as bh
from tests.helpers import run_both, assert_parity

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 1 << 22


def stat(name):
    for line in bh.message("statistic").splitlines():
        if line.startswith(name + ":"):
            return float(line.split(":")[1])
    raise KeyError(name)


@pytest.fixture
def tuning():
    bh.message("autotune: forget")
    bh.message("autotune_samples: 1")
    bh.message("autotune: on")
    yield
    bh.message("autotune_samples: 3")
    bh.message("autotune: forget")


def _x(dtype, seed=5):
    rng = np.random.RandomState(seed)
    return rng.uniform(-2.0, 2.0, N).astype(dtype)


def _programs(x):
    """Three flushes, so that each skeleton is a kernel of its own."""
    a = bh.array(x)
    y = a * 1.5 + bh.ufunc("sqrt")(bh.ufunc("absolute")(a))          # MAP
    bh.flush()
    s = (a * a).sum()                                                  # FULL_REDUCE (float)
    c = (a > 0.0).astype(np.int64).sum()                               # FULL_REDUCE (int64)
    bh.flush()
    k = bh.ufunc("add").accumulate((a > 0.5).astype(np.int64))         # SCAN (int64)
    return [y, s, c, k]


def test_every_candidate_computes_the_same_results(tuning):
    x = _x(np.float64)
    bh.message("autotune: off")
    ref = [r.asnumpy() for r in _programs(x)]
    bh.message("autotune: on")
    d0 = stat("autotune_decisions")
    t0 = stat("autotune_trials")
    bound = N * np.finfo(np.float64).eps * float(np.sum(x * x))
    for it in range(40):
        before = stat("autotune_trials")
        got = [r.asnumpy() for r in _programs(x)]
        np.testing.assert_array_equal(got[0], ref[0])
        assert abs(float(got[1]) - float(ref[1])) <= bound, (it, float(got[1]), float(ref[1]))
        np.testing.assert_array_equal(got[2], ref[2])
        np.testing.assert_array_equal(got[3], ref[3])
        if stat("autotune_trials") == before and it > 0:
            break  # every kernel of the program has settled
        bh.message("autotune")  # settles trials whose samples are in
    assert stat("autotune_trials") > t0, "no kernel was tuned"
    assert stat("autotune_decisions") >= d0 + 3, bh.message("autotune")
    table = bh.message("autotune")
    assert "map:" in table and "reduce:" in table and "scan:" in table, table
    # after the decisions: the chosen shapes, checked against the oracle
    got, want = run_both(lambda: _programs(x))
    np.testing.assert_allclose(got[0], want[0], rtol=1e-12)
    assert abs(float(got[1]) - float(want[1])) <= bound
    assert_parity(got[2], want[2], np.int64)
    assert_parity(got[3], want[3], np.int64)


def test_table_is_reused_by_a_new_process(tmp_path):
    prog = r"""
import sys
sys.path.insert(0, %(root)r)
import numpy as np
import bohrium_amd as bh
bh.message("autotune_samples: 1")
x = np.random.RandomState(5).uniform(-2, 2, %(n)d)
def stat(name):
    for line in bh.message("statistic").splitlines():
        if line.startswith(name + ":"):
            return float(line.split(":")[1])
for it in range(30):
    y = (bh.array(x) * 1.5 + 2.0).asnumpy()
    bh.message("autotune")
    if stat("autotune_decisions") + stat("autotune_table_hits") > 0 and it > 2:
        break
print("trials", stat("autotune_trials"), "decisions", stat("autotune_decisions"),
      "hits", stat("autotune_table_hits"))
"""
    env = dict(os.environ, BH_HIP_CACHE_DIR=str(tmp_path))
    src = prog % {"root": ROOT, "n": N}
    outs = []
    for _ in range(2):
        r = subprocess.run([sys.executable, "-c", src], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-3000:]
        outs.append(r.stdout.split())
    first, second = outs[0], outs[1]
    assert float(first[first.index("decisions") + 1]) >= 1, first
    assert float(second[second.index("trials") + 1]) == 0, second
    assert float(second[second.index("hits") + 1]) >= 1, second
    files = [f for f in os.listdir(tmp_path) if f.startswith("autotune_")]
    assert files, os.listdir(tmp_path)
