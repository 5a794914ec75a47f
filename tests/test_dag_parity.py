"""Random array programs (the cost-model workload) on the hip engine against the oracle, under
each fuser: every partition the fusers produce, and every skeleton the planner picks for it, must
compute what the bytecode says.

The programs (workloads.random_dag) mix element-wise float32 operations, row and column sums
broadcast back, transposed, reversed and half-size views; each runs as one flush.  Tolerance:
rtol 1e-4 plus an absolute 1e-4 of the largest magnitude among all the program's outputs, since
a chain of up to a few hundred float32 operations with sums of 64-128 terms reassociates and
rounds (and contracts into FMAs) differently on the device and the host, and an output such as
x - mean(x) of a row that is constant holds only rounding noise (up to 3e-5 on the host where the
device computes 1e-9, for outputs of scale 0.65); a wrong element, index or loop order is off by
O(1) of the program's scale.  (The generator takes sqrt(|x| + 1), not sqrt(|x|), so that no
output is the square root of such noise.)"""
import os

import numpy as np
import pytest

import bohrium_amd as bh
from bohrium_amd import _bh_api as api
from bohrium_amd import workloads
from tests.helpers import run_both

pytestmark = pytest.mark.gpu

DEFAULT_FUSERS = "greedy, push_reductions_inwards, collapse_redundant_axes"
FUSERS = {
    "greedy": DEFAULT_FUSERS,
    "serial": "serial, push_reductions_inwards, collapse_redundant_axes",
    "none": "push_reductions_inwards, collapse_redundant_axes",
    # the MLP-scored fuser with the model trained on MI355X (bohrium_amd/data)
    "learned": "learned, push_reductions_inwards, collapse_redundant_axes",
}
SHIPPED_MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bohrium_amd", "data",
                             "costmodel_mi355x.npz")
# programs whose outputs are all finite (exp chains overflow in some others)
CASES = [(1001, 337), (1002, 374), (1006, 172), (1007, 209)]
_ORACLE = {}


@pytest.fixture(autouse=True)
def one_flush_per_program():
    api.set_flush_threshold(10 ** 7)
    yield
    api.set_flush_threshold(1000)
    bh.message("fuser_list: " + DEFAULT_FUSERS)


@pytest.fixture(scope="module")
def shipped_model():
    from bohrium_amd import costmodel
    m = costmodel.CostModel.load(SHIPPED_MODEL)
    m.install()
    yield m
    costmodel.uninstall()


def _program(seed, n):
    return lambda: workloads.random_dag(seed, n, sides=(64, 128))


@pytest.mark.parametrize("fuser", sorted(FUSERS))
@pytest.mark.parametrize("seed,n", CASES)
def test_random_program(seed, n, fuser, shipped_model):
    bh.message("fuser_list: " + FUSERS[fuser])
    if (seed, n) not in _ORACLE:
        got, want = run_both(_program(seed, n))
        _ORACLE[(seed, n)] = want
    else:
        got = [a.asnumpy() for a in _program(seed, n)()]
        want = _ORACLE[(seed, n)]
    assert len(got) == len(want)
    scale = max([1.0] + [float(np.abs(w).max()) for w in want if w.size])
    worst = max(float(np.abs(g.astype(np.float64) - w).max()) for g, w in zip(got, want) if w.size)
    print("seed %d fuser %s: %d outputs, scale %.3g, largest difference %.3g" % (seed, fuser, len(want), scale, worst))
    for g, w in zip(got, want):
        assert g.shape == w.shape and g.dtype == w.dtype
        np.testing.assert_allclose(g, w, rtol=1e-4, atol=1e-4 * scale)


def test_learned_plans_at_2048_squared(shipped_model):
    """The learned fuser's partition of one random program on 2048 x 2048 float32 arrays against
    the oracle; same tolerance as above (row and column sums now have 2048 terms, whose float32
    rounding stays far below 1e-4 of the scale)."""
    bh.message("fuser_list: " + FUSERS["learned"])
    got, want = run_both(lambda: workloads.random_dag(1006, 172, sides=(2048,)))
    assert len(got) == len(want) and len(got) > 0
    scale = max([1.0] + [float(np.abs(w).max()) for w in want if w.size])
    for g, w in zip(got, want):
        assert g.shape == w.shape and np.isfinite(w).all()
        np.testing.assert_allclose(g, w, rtol=1e-4, atol=1e-4 * scale)


def test_learned_plans_at_4096_squared_match_greedy(shipped_model):
    """At 4096 x 4096 (the largest arrays of the cost-model workload) the oracle needs minutes on
    the host, so the learned partition is checked against the greedy one on the device (greedy's
    partitions are checked against the oracle above): the same bytecode, so the same results
    within the tolerance above."""
    outs = {}
    for fuser in ("greedy", "learned"):
        bh.message("fuser_list: " + FUSERS[fuser])
        outs[fuser] = [a.asnumpy() for a in workloads.random_dag(1006, 172, sides=(4096,))]
    g, l = outs["greedy"], outs["learned"]
    assert len(g) == len(l) and len(g) > 0
    scale = max([1.0] + [float(np.abs(w).max()) for w in g if w.size])
    for a, b in zip(l, g):
        assert a.shape == b.shape == (4096, 4096) or a.shape == b.shape
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-4 * scale)
