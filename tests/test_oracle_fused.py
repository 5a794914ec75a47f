"""The oracle's fused OpenMP form (the CPU baseline, oracle.omp_ve.FusedProgram) against the oracle's
instruction-by-instruction replay on the same traces (CPU only).

The fused form writes runs of instructions as one loop nest with the reference OpenMP engine's
pragmas (parallel for on the outermost loop with reduction clauses, atomic or critical updates for
sweeps that cannot have a clause, simd on the innermost loop, scalar temporaries); it must compute
what the replay computes: integers and booleans exactly, floats within the difference of summation
orders (parallel reduction clauses, per-thread partial sums): 1e-12 relative for float64 results,
1e-5 for float32 (plus 1e-4 of the output scale for the random float32 programs, as in their GPU
parity test)."""
import numpy as np
import pytest

import bohrium_amd as bh
from bohrium_amd import workloads
from oracle import omp_ve
from tests.helpers import record


def _compare(program, rtol64=1e-12, rtol32=1e-5, scale_atol=0.0):
    res, trace = record(program)
    want = omp_ve.replay(trace)
    prog = omp_ve.FusedProgram(trace)
    _, got = prog.run_all()
    assert set(want) <= set(got)
    scale = max([1.0] + [float(np.abs(w).max()) for w in want.values() if w.size and w.dtype.kind == "f"])
    for b, w in want.items():
        g = got[b]
        if w.dtype.kind in "biu":
            np.testing.assert_array_equal(g, w)
        else:
            rtol = rtol64 if w.dtype == np.float64 else rtol32
            np.testing.assert_allclose(g, w, rtol=rtol, atol=rtol + scale_atol * scale)
    return prog


def test_black_scholes_one_parallel_loop_per_step():
    prog = _compare(lambda: workloads.black_scholes(10001, steps=3))
    assert "reduction(+:" in prog.source and "parallel for simd" in prog.source
    assert max(n for n, _ in prog.runs) >= 40  # the whole step is one loop


def test_mc_pi_count_is_exact():
    prog = _compare(lambda: workloads.mc_pi(100003))
    assert "oracle_random123" in prog.source


def test_jacobi_two_dimensional_blocks():
    def program():
        G = workloads.jacobi_grid(64)
        return [workloads.jacobi_step(G) for _ in range(3)] + [G]
    prog = _compare(program, rtol32=1e-6)
    assert any(len(S) == 2 and n >= 4 for n, S in prog.runs), prog.runs


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int64, np.int32])
def test_axis_reductions_clause_atomic_critical_simd(dtype):
    rng = np.random.RandomState(3)
    x = rng.uniform(-3, 3, (37, 53))
    x = x.astype(dtype) if np.dtype(dtype).kind == "f" else (x * 10).astype(dtype)

    def program():
        a = bh.array(x)
        y = a * 2 + 1
        return [y.sum(axis=0), y.sum(axis=1), bh.ufunc("maximum").reduce(y, 0), bh.ufunc("minimum").reduce(y, 1),
                bh.ufunc("multiply").reduce(a * 0 + 1, 0), (y > 0).sum(), y.max()]
    prog = _compare(program)
    assert "#pragma omp atomic" in prog.source      # column sums (axis 0 of a 2-D block)
    assert "#pragma omp critical" in prog.source    # column maxima

    def rows_only():
        y = bh.array(x) * 2 + 1
        return [y.sum(axis=1), bh.ufunc("minimum").reduce(y, 1)]
    prog = _compare(rows_only)
    # a block without atomic or critical updates: its row reductions are simd accumulators
    assert "#pragma omp simd reduction(+:" in prog.source and "reduction(min:" in prog.source


def test_logical_and_bitwise_reductions():
    rng = np.random.RandomState(4)
    b = rng.randint(0, 2, (30, 40)).astype(bool)
    i = rng.randint(0, 1 << 20, (30, 40)).astype(np.int64)

    def program():
        B, I = bh.array(b), bh.array(i)
        return [bh.ufunc("logical_and").reduce(B, 0), bh.ufunc("logical_or").reduce(B, 1),
                bh.ufunc("bitwise_xor").reduce(I, 0), bh.ufunc("bitwise_and").reduce(I + 1, 1),
                bh.ufunc("bitwise_or").reduce(I.reshape(1200) * 3)]
    _compare(program)


@pytest.mark.parametrize("seed,n", [(1001, 337), (1006, 172)])
def test_random_programs(seed, n):
    _compare(lambda: workloads.random_dag(seed, n, sides=(64, 128)), rtol32=1e-4, scale_atol=1e-4)
