"""Parity of the hip engine with the oracle on the same bridge trace (MI355X).

Each test runs a program on the engine with the trace recorded, replays the
identical trace on the oracle and compares: bit-exact for integer/bool/index/
philox results; fp64 within 1e-12 and fp32 within 1e-5 relative (plus the
stated absolute slack where a reduction order differs)."""
import math

import numpy as np
import pytest

import bohrium_amd as bh
from bohrium_amd import workloads
from tests.helpers import run_both, assert_parity

pytestmark = pytest.mark.gpu

INT_TYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64]
FLOAT_TYPES = [np.float32, np.float64]


@pytest.fixture(scope="module", autouse=True)
def engine_is_hip():
    info = bh.message("info")
    assert "HIP vector engine" in info, "the hip engine is not the active component: %r" % info
    yield


def _inputs(dtype, n=1000, seed=0, positive=False, nonzero=False):
    rng = np.random.RandomState(seed)
    dtype = np.dtype(dtype)
    if dtype.kind == "b":
        return rng.randint(0, 2, n).astype(bool)
    if dtype.kind in "iu":
        lo = 1 if (positive or nonzero or dtype.kind == "u") else -50
        a = rng.randint(lo, 50, n).astype(dtype)
        if nonzero:
            a[a == 0] = 1
        return a
    if dtype.kind == "c":
        return (rng.uniform(-2, 2, n) + 1j * rng.uniform(-2, 2, n)).astype(dtype)
    return rng.uniform(0.1 if positive else -2.0, 2.0, n).astype(dtype)


@pytest.mark.parametrize("name", ["add", "subtract", "multiply", "maximum", "minimum", "greater", "less_equal",
                                  "equal", "not_equal"])
@pytest.mark.parametrize("dtype", INT_TYPES + FLOAT_TYPES)
def test_binary(name, dtype):
    x, y = _inputs(dtype, seed=1), _inputs(dtype, seed=2)
    got, want = run_both(lambda: bh.ufunc(name)(bh.array(x), bh.array(y)))
    assert_parity(got[0], want[0], got[0].dtype)


@pytest.mark.parametrize("dtype", [np.int8, np.int32, np.int64, np.uint16, np.uint64])
def test_integer_division_remainder_bitwise(dtype):
    x, y = _inputs(dtype, seed=3), _inputs(dtype, seed=4, nonzero=True)
    prog = lambda: [bh.array(x) // bh.array(y), bh.ufunc("remainder")(bh.array(x), bh.array(y)),
                    bh.ufunc("fmod")(bh.array(x), bh.array(y)), bh.array(x) & bh.array(y),
                    bh.array(x) ^ bh.array(y), bh.ufunc("left_shift")(bh.array(x), bh.array(y) % 3),
                    bh.ufunc("invert")(bh.array(x)), bh.ufunc("absolute")(bh.array(x)), bh.array(x) ** 2]
    got, want = run_both(prog)
    for g, w in zip(got, want):
        assert_parity(g, w, g.dtype)


UNARY = ["sqrt", "exp", "log", "sin", "cos", "tan", "tanh", "sinh", "cosh", "arctan", "arcsin", "arccos",
         "exp2", "expm1", "log2", "log10", "log1p", "floor", "ceil", "trunc", "rint", "absolute", "sign",
         "arcsinh", "isnan", "isinf", "isfinite"]


@pytest.mark.parametrize("dtype", FLOAT_TYPES)
def test_unary_float(dtype):
    x = _inputs(dtype, seed=5, positive=True) * np.asarray(0.45, dtype)
    got, want = run_both(lambda: [bh.ufunc(n)(bh.array(x)) for n in UNARY])
    for n, g, w in zip(UNARY, got, want):
        assert_parity(g, w, g.dtype, rtol=2e-6 if dtype == np.float32 else 1e-12, atol=1e-300)


@pytest.mark.parametrize("src", [np.float64, np.float32, np.int64, np.int32, np.uint8, np.bool_])
def test_identity_casts(src):
    x = _inputs(src, seed=6, positive=True)
    dsts = [np.float64, np.float32, np.int64, np.int16, np.uint32, np.bool_]
    got, want = run_both(lambda: [bh.array(x).astype(d) for d in dsts])
    for g, w in zip(got, want):
        assert_parity(g, w, g.dtype)


def test_complex():
    x, y = _inputs(np.complex128, seed=7), _inputs(np.complex128, seed=8)
    prog = lambda: [bh.array(x) + bh.array(y), bh.array(x) * bh.array(y), bh.array(x) / bh.array(y),
                    bh.ufunc("exp")(bh.array(x)), bh.ufunc("real")(bh.array(x)), bh.ufunc("sqrt")(bh.array(x))]
    got, want = run_both(prog)
    for g, w in zip(got, want):
        assert_parity(g, w, g.dtype, rtol=1e-12)


def _random_view(shape, seed):
    rng = np.random.RandomState(seed)
    key = []
    for n in shape:
        start = rng.randint(0, max(1, n // 3))
        step = int(rng.choice([1, 2, -1]))
        key.append(slice(start, n, step) if step > 0 else slice(n - 1 - start, None, step))
    return tuple(key)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("ndim", [1, 2, 3, 4])
def test_reduce_views_every_axis(seed, ndim):
    shape = (9, 7, 5, 6)[:ndim]
    x = np.random.RandomState(seed).uniform(-1, 1, shape).astype(np.float32)
    xi = (x * 100).astype(np.int64)
    key = _random_view(shape, seed)

    def prog():
        out = []
        for axis in range(ndim):
            out.append(bh.array(x)[key].sum(axis=axis))
            out.append(bh.ufunc("maximum").reduce(bh.array(xi)[key], axis))
            out.append(bh.ufunc("minimum").reduce(bh.array(x)[key], axis))
        return out
    got, want = run_both(prog)
    for g, w in zip(got, want):
        assert_parity(g, w, g.dtype, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("name", ["add", "multiply", "minimum", "maximum", "bitwise_or", "bitwise_xor",
                                  "bitwise_and", "logical_or", "logical_and"])
@pytest.mark.parametrize("n", [1, 10, 1000, 1 << 20])
def test_full_reductions(name, n):
    dts = [np.int64, np.uint32, np.float64] if not name.startswith("logical") else [np.bool_]
    if name.startswith("bitwise"):
        dts = [np.int64, np.uint32, np.int8]
    for dt in dts:
        x = _inputs(dt, n=n, seed=9, positive=True)
        if name == "multiply" and np.dtype(dt).kind == "f":
            x = (1.0 + (x - 1.0) * 1e-6).astype(dt)
        got, want = run_both(lambda: bh.ufunc(name).reduce(bh.array(x)))
        if np.dtype(dt).kind != "f" or name in ("minimum", "maximum"):
            assert_parity(got[0], want[0], got[0].dtype)  # exact: integers, and min/max in any order
            continue
        # float64 sums and products in another order: each result lies within (n - 1) eps of the
        # exact value (times sum|x| for sums; relative for products of positive factors), so the two
        # differ by at most twice that: 2.2e-13 relative at n = 1000, 2.3e-10 at n = 2^20 (the
        # worst case; rounding errors of a tree and a sequential sum typically grow with sqrt(n)).
        eps = np.finfo(np.float64).eps
        g, w = float(got[0]), float(want[0])
        if name == "add":
            exact = math.fsum(x.astype(np.float64))
            tol = (n - 1) * eps * float(np.abs(x).sum())
            assert abs(g - exact) <= tol and abs(w - exact) <= tol, (g, w, exact, tol)
        else:
            exact = math.exp(math.fsum(np.log(x.astype(np.float64))))
            tol = 2 * (n - 1) * eps * abs(w) + 1e-12 * abs(w) * (n > 1)
            assert abs(g - w) <= max(tol, 1e-300), (g, w, tol)
            assert abs(g - exact) <= 1e-12 * abs(exact) + 4 * n * eps * abs(exact), (g, exact)


def test_reduce_identity_quirks():
    neg = -np.arange(1.0, 1001.0)
    big = np.array([0x7FFFFFFF00000FFF, 0x0FFFFFFFFFFFFF0F] * 500, dtype=np.int64)
    got, want = run_both(lambda: [bh.array(neg).max(), bh.ufunc("bitwise_and").reduce(bh.array(big))])
    assert got[0] == want[0] == np.finfo(np.float64).tiny
    assert got[1] == want[1]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_accumulate_views(axis):
    x = np.random.RandomState(10).uniform(-1, 1, (30, 20, 10))
    xi = (x * 10).astype(np.int64)
    key = _random_view(x.shape, 4)
    got, want = run_both(lambda: [bh.array(x)[key].cumsum(axis=axis), bh.array(xi)[key].cumprod(axis=axis)])
    assert_parity(got[0], want[0], np.float64, rtol=1e-12, atol=1e-12)
    assert_parity(got[1], want[1], np.int64)


@pytest.mark.parametrize("n", [1, 100, 2048, 2049, 100000, 3000001])
def test_scan_1d_int64_exact(n):
    x = np.random.RandomState(11).randint(-3, 5, n).astype(np.int32)
    got, want = run_both(lambda: (bh.array(x) > 0).astype(np.int64).cumsum())
    assert_parity(got[0], want[0], np.int64)
    np.testing.assert_array_equal(got[0], np.cumsum(x > 0))


def test_scan_float_and_product():
    """float64 cumsum and cumprod of 50,000 positive values near 1: the device scans in tiles and
    the oracle in sequence, so element k of each lies within k eps (relative, all terms positive)
    of its exact value and the two differ by at most 2 k eps: 1e-12 relative up to k = 4500,
    1.1e-11 at the last element.  The test holds every element to that bound and the first 4,500
    to 1e-12."""
    x = np.random.RandomState(12).uniform(0.999, 1.001, 50000)
    got, want = run_both(lambda: [bh.array(x).cumsum(), bh.array(x).cumprod()])
    eps = np.finfo(np.float64).eps
    k = np.arange(1, x.size + 1, dtype=np.float64)
    for g, w in zip(got, want):
        rel = np.abs(g - w) / np.abs(w)
        assert (rel <= 2 * k * eps).all(), float((rel / (2 * k * eps)).max())
        assert rel[:4500].max() <= 1e-12, float(rel[:4500].max())


def test_gather_scatter():
    rng = np.random.RandomState(13)
    data = rng.uniform(size=100000).astype(np.float32)
    idx = rng.randint(0, 100000, 60000).astype(np.int64)
    uidx = rng.permutation(100000)[:30000].astype(np.int64)
    vals = rng.uniform(size=30000).astype(np.float32)
    mask = rng.randint(0, 2, 30000).astype(bool)

    def prog():
        g = bh.take(bh.array(data), bh.array(idx))
        a = bh.array(data)
        bh.put(a, bh.array(uidx), bh.array(vals))
        b = bh.array(data)
        bh.put(b, bh.array(uidx), bh.array(vals), mask=bh.array(mask))
        return [g, a, b]
    got, want = run_both(prog)
    for g, w in zip(got, want):
        assert_parity(g, w, g.dtype)


def test_gather_out_of_bounds_raises():
    with pytest.raises(Exception):
        bh.take(bh.arange(10), bh.array(np.array([3, 11], np.int64))).asnumpy()
    # The engine stays usable afterwards.
    np.testing.assert_array_equal(bh.arange(5).asnumpy(), np.arange(5))


@pytest.mark.parametrize("n", [1, 1000, 1000003])
def test_random_is_bit_exact(n):
    got, want = run_both(lambda: [bh.random.RandomState(42).random123(n),
                                  bh.random.RandomState(7).random_sample((n, 3))[:, 1]])
    assert_parity(got[0], want[0], np.uint64)
    # float64(u64) / 2**64 is exact on both sides, so the uniforms are bit-exact too.
    np.testing.assert_array_equal(got[1], want[1])


def test_mc_pi_1e6_bit_exact():
    got, want = run_both(lambda: workloads.mc_pi(10 ** 6))
    assert int(got[0]) == int(want[0])


def test_mc_pi_full_config_bit_exact():
    """The benchmark size: N = 1e7 each, seed 42; the count must be bit-exact."""
    got, want = run_both(lambda: workloads.mc_pi(10 ** 7))
    assert int(got[0]) == int(want[0])
    assert abs(4.0 * int(got[0]) / 1e7 - np.pi) < 2e-3


def test_black_scholes_small():
    got, want = run_both(lambda: workloads.black_scholes(100003, steps=3))
    for g, w in zip(got, want):
        assert_parity(g, w, np.float64, rtol=1e-12)


def test_jacobi_small():
    def prog():
        G, deltas = workloads.jacobi(130, 20)
        return [G] + deltas
    got, want = run_both(prog)
    for g, w in zip(got, want):
        assert_parity(g, w, np.float32, rtol=1e-5, atol=1e-3)


def test_do_while_known_answers():
    def prog():
        a = bh.zeros(4)

        def body(a):
            a += 1
        bh.do_while(body, 5, a)
        b = bh.zeros(4)

        def body2(b):
            b += 1
            return bh.sum(b) < 10
        bh.do_while(body2, None, b)
        return [a, b]
    got, want = run_both(prog)
    np.testing.assert_array_equal(got[0], [5, 5, 5, 5])
    np.testing.assert_array_equal(got[1], [3, 3, 3, 3])
    np.testing.assert_array_equal(want[1], [3, 3, 3, 3])


def test_empty_and_ragged():
    def prog():
        e = bh.zeros(0)
        s = e.sum()
        r = bh.array(np.arange(7.0))[::3] * 2
        one = bh.array(np.array([5.0])).sum()
        return [s, r, one]
    got, want = run_both(prog)
    for g, w in zip(got, want):
        assert_parity(g, w, np.float64)


def test_sliding_views_known_answers():
    def prog():
        a = bh.array(np.arange(1.0, 11.0))

        def body(a):
            i = bh.get_iterator(1)
            a[i] += a[i - 1]
        bh.do_while(body, 9, a)
        b = bh.arange(1, 6)

        def kernel(b):
            i = bh.get_iterator(1)
            b[i] *= b[i - 1]
        bh.do_while(kernel, 4, b)
        return [a, b]
    got, want = run_both(prog)
    np.testing.assert_array_equal(got[0], [1, 3, 6, 10, 15, 21, 28, 36, 45, 55])
    np.testing.assert_array_equal(got[1], [1, 2, 6, 24, 120])
    np.testing.assert_array_equal(want[1], [1, 2, 6, 24, 120])


def test_black_scholes_repeat_full_size_sums():
    """The benchmark size (N = 2e7, 10 steps): engine vs oracle on the same bytecode."""
    def prog():
        S, K, T = workloads.black_scholes_inputs(20_000_000)
        return workloads.black_scholes_repeat(S, K, T, steps=10)
    got, want = run_both(prog)
    assert_parity(got[0], want[0], np.float64, rtol=1e-12)


def _jacobi_repeat_prog(n, iters):
    def prog():
        G = workloads.jacobi_grid(n)
        d = workloads.jacobi_repeat(G, iters)
        return [G, d]
    return prog


def test_jacobi_repeat_graph_replay_matches_eager_and_oracle():
    """The repeat body is graph-captured after two identical iterations; the replayed
    result must equal the eager one bit for bit and the oracle within fp32 tolerance."""
    bh.message("graph: off")
    try:
        eager = [a.asnumpy() for a in _jacobi_repeat_prog(258, 12)()]
    finally:
        bh.message("graph: on")
    before = int(bh.message("graph replays"))
    got, want = run_both(_jacobi_repeat_prog(258, 12))
    assert int(bh.message("graph replays")) > before, \
        "the repeat body was not replayed as a graph:\n" + bh.message("statistic")
    for g, e in zip(got, eager):
        np.testing.assert_array_equal(g, e)
    for g, w in zip(got, want):
        assert_parity(g, w, np.float32, rtol=1e-5, atol=1e-3)


def test_jacobi_full_grid_20_sweeps():
    """8192 x 8192 fp32 grid (the benchmark's size), 20 replayed sweeps vs the oracle."""
    got, want = run_both(_jacobi_repeat_prog(8192, 20))
    assert_parity(got[1], want[1], np.float32, rtol=1e-5)
    assert_parity(got[0], want[0], np.float32, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("unroll", [1, 2, 3, 4, 8])
def test_unrolled_grid_stride_parity(unroll):
    """Unrolled map and reduce skeletons (sizes that leave a remainder) against the oracle; one-rank walks
    take groups of `unroll` consecutive elements (merged wide loads), 2-D ones grid-strided groups."""
    old = bh.message("codegen:")
    bh.message("codegen: map_unroll=%d reduce_unroll=%d max_grid=64" % (unroll, unroll))
    try:
        def prog():
            S, K, T = workloads.black_scholes_inputs(64 * 256 * unroll * 3 + 77)
            x = bh.random.RandomState(3).random_sample((301, 257))
            y = x[1:, :-1] * 2.0 + x[:-1, 1:]
            return [workloads.black_scholes_step(S, K, T, 1), y, y.sum(), (y > 1.0).sum(), workloads.mc_pi(100003)]
        got, want = run_both(prog)
        assert_parity(got[0], want[0], np.float64, rtol=1e-12)
        assert_parity(got[1], want[1], np.float64)
        assert_parity(got[2], want[2], np.float64, rtol=1e-12)
        assert int(got[3]) == int(want[3])
        assert int(got[4]) == int(want[4])
    finally:
        bh.message("codegen: " + old)


def _batched_prog(count, seed=7):
    sizes = workloads.batched_sizes(count, lo=256, hi=20000, seed=seed)

    def prog():
        xs, idxs = workloads.batched_inputs(sizes, seed=seed)
        sums, cums, gathered = workloads.batched_step(xs, idxs)
        return sums + cums + gathered + xs
    return prog, count


BATCH_MODES = {"persistent": ("batch_persistent: on", "persist_launches:"),
               "levels": ("batch_persistent: off", "batch_launches:")}


def _stat(stats, name):
    return int([l for l in stats.splitlines() if l.startswith(name)][0].split()[1])


@pytest.mark.parametrize("mode", sorted(BATCH_MODES))
@pytest.mark.parametrize("count", [16, 200])
def test_batched_uneven_blocks_parity(count, mode):
    """Many uneven arrays in one flush: the map+sum, scan+gather launches of every array run as
    one persistent launch (or as batched launches level by level); results match the oracle (sums
    fp32, cumsums and gathers exact) and the unbatched engine bit for bit (the persistent launch's
    sums within the summation-order bound: it splits a sum into fewer, longer workgroups)."""
    from bohrium_amd import _bh_api as api
    prog, n = _batched_prog(count)
    msg, counter = BATCH_MODES[mode]
    api.set_flush_threshold(10 ** 7)
    try:
        bh.message("batch_min: 0")
        plain = [a.asnumpy() for a in prog()]
        bh.message("batch_min: 8")
        bh.message(msg)
        before = _stat(bh.message("statistic"), counter)
        got, want = run_both(prog)
        stats1 = bh.message("statistic")
        after = _stat(stats1, counter)
    finally:
        api.set_flush_threshold(1000)
        bh.message("batch_min: 8")
        bh.message("batch_persistent: off")
    assert after > before, "no %s launch:\n" % mode + stats1
    for i, (g, p) in enumerate(zip(got, plain)):
        if mode == "levels" or i >= n:
            np.testing.assert_array_equal(g, p)
    for i in range(n):
        # fp32 sums in different orders: both within the bound n * 2^-24 * sum|y| of the exact sum
        y = np.tanh(want[3 * n + i].astype(np.float64) * 0.5 + 0.1)
        bound = y.size * 2.0 ** -24 * np.abs(y).sum()
        exact = math.fsum(np.tanh(want[3 * n + i] * np.float32(0.5) + np.float32(0.1)).astype(np.float64))
        assert abs(float(got[i]) - exact) <= bound + 1e-5 * abs(exact)
        assert abs(float(want[i]) - exact) <= bound + 1e-5 * abs(exact)
        np.testing.assert_array_equal(got[n + i], want[n + i])
        assert_parity(got[2 * n + i], want[2 * n + i], np.float32, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("shape", [(37, 1000), (3, 5, 300), (1, 64), (300, 33)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int64])
def test_row_reductions_one_wave_per_row(shape, dtype):
    """Reductions over the innermost axis of an expression run as one wave per row; results
    match the oracle and the one-thread-per-row map skeleton."""
    rng = np.random.RandomState(len(shape) * 7 + shape[-1])
    x = rng.uniform(0.5, 1.5, shape).astype(dtype) if np.dtype(dtype).kind == "f" else \
        rng.randint(-50, 50, shape).astype(dtype)

    def prog():
        a = bh.array(x)
        v = a[..., ::-1]
        out = [(a * 2 + 1).sum(axis=-1), bh.ufunc("maximum").reduce(v - 3, -1),
               bh.ufunc("minimum").reduce(a * a, -1)]
        if np.dtype(dtype).kind == "i":
            out += [bh.ufunc("bitwise_xor").reduce(a, -1), bh.ufunc("bitwise_and").reduce(a + 100, -1),
                    bh.ufunc("logical_or").reduce(a > 40, -1)]
        else:
            out += [bh.ufunc("multiply").reduce(a * (1.0 / 1.0001), -1)]
        return out
    old = bh.message("codegen:")
    try:
        bh.message("codegen: row_reduce=0")
        plain = [o.asnumpy() for o in prog()]
        bh.message("codegen: row_reduce=1")
        got, want = run_both(prog)
    finally:
        bh.message("codegen: " + old)
    for g, p, w in zip(got, plain, want):
        # fp32 sums/products of up to 1000 terms in different orders: within n * 2^-24 relative
        rtol = 1e-4 if np.dtype(dtype) == np.float32 else 1e-12
        assert_parity(g, w, g.dtype, rtol=rtol, atol=0.0 if g.dtype.kind != "f" else 1e-6)
        assert_parity(g, p, g.dtype, rtol=rtol, atol=0.0 if g.dtype.kind != "f" else 1e-6)


@pytest.mark.parametrize("mode", sorted(BATCH_MODES))
def test_batched_full_configuration(mode):
    """The batched-arrays configuration at its full size: 4096 float32 arrays of
    floor(exp(U[ln 256, ln 262144])) elements (seed 42, 153,874,193 elements), x ~ N(0, 1) and
    int64 indices from philox, y = tanh(0.5 x + 0.1), sum(y), int64 cumsum(x > 0), y[idx], run as one
    persistent launch (or as batched launches level by level).  The oracle would need minutes on the host at this size, so each result is
    checked against what the fetched inputs fix: the cumsums equal numpy's exactly, every gathered
    element is y at its index (tanh within 1e-5 relative), and every sum lies within
    n * 2^-24 * sum|y| (summation order) + 1e-5 * sum|y| (tanh) of the float64 sum of y."""
    from bohrium_amd import _bh_api as api
    sizes = workloads.batched_sizes(4096)
    assert int(sizes.sum()) == 153874193 and sizes.min() >= 256 and sizes.max() <= 262144
    msg, counter = BATCH_MODES[mode]
    api.set_flush_threshold(10 ** 7)
    try:
        bh.message("batch_min: 8")
        bh.message(msg)
        xs, idxs = workloads.batched_inputs(sizes)
        bh.flush()
        stats0 = bh.message("statistic")
        sums, cums, gathered = workloads.batched_step(xs, idxs)
        got_s = bh.asnumpy_many(sums)
        stats1 = bh.message("statistic")
        got_c = bh.asnumpy_many(cums)
        got_g = bh.asnumpy_many(gathered)
        host_x = bh.asnumpy_many(xs)
        host_i = bh.asnumpy_many(idxs)
    finally:
        api.set_flush_threshold(1000)
        bh.message("batch_persistent: off")
    assert _stat(stats1, counter) > _stat(stats0, counter), "no %s launch" % mode
    for k in range(len(sizes)):
        x, idx = host_x[k], host_i[k]
        n = int(sizes[k])
        assert x.shape == (n,) and idx.shape == (n,) and idx.min() >= 0 and idx.max() < n
        np.testing.assert_array_equal(got_c[k], np.cumsum(x > 0, dtype=np.int64))
        y = np.tanh(x * np.float32(0.5) + np.float32(0.1))
        np.testing.assert_allclose(got_g[k], y[idx], rtol=1e-5, atol=1e-6)
        ay = float(np.abs(y.astype(np.float64)).sum())
        exact = float(y.astype(np.float64).sum())
        assert abs(float(got_s[k]) - exact) <= n * 2.0 ** -24 * ay + 1e-5 * ay, (k, n, float(got_s[k]), exact)


def test_full_reduce_with_sweeping_tail_blocks():
    """A FULL_REDUCE launch whose unit carries small sweeping loops after its main loop (a 200- and a
    100-element sum and scan, which the finalize kernel runs in one thread): the main loop is the
    first sweeping loop, not the last.  The reshapable-first fuser (the greedy fuser's fallback for
    flushes with more than 10,000 dependency edges, e.g. the 4096-array batched configuration)
    builds such kernels from a few arrays."""
    from bohrium_amd import _bh_api as api
    sizes = np.array([5000, 200, 3000, 100])

    def prog():
        xs, idxs = workloads.batched_inputs(sizes)
        sums, cums, gathered = workloads.batched_step(xs, idxs)
        return list(sums) + list(cums) + list(gathered) + list(xs)

    api.set_flush_threshold(10 ** 7)
    try:
        bh.message("fuser_list: reshapable_first, push_reductions_inwards, collapse_redundant_axes")
        got, want = run_both(prog)
    finally:
        bh.message("fuser_list: greedy, push_reductions_inwards, collapse_redundant_axes")
        api.set_flush_threshold(1000)
    n = len(sizes)
    for i in range(n):
        y = np.tanh(want[3 * n + i] * np.float32(0.5) + np.float32(0.1)).astype(np.float64)
        bound = y.size * 2.0 ** -24 * np.abs(y).sum() + 1e-5 * np.abs(y).sum()
        assert abs(float(got[i]) - float(want[i])) <= 2 * bound, (i, float(got[i]), float(want[i]))
        np.testing.assert_array_equal(got[n + i], want[n + i])
        assert_parity(got[2 * n + i], want[2 * n + i], np.float32, rtol=1e-5, atol=1e-6)


def test_persistent_grid_barrier_timeout_is_reported():
    """A grid barrier of the persistent launch that gives up waiting raises an engine error instead of
    hanging, and persistent launches are then switched off.  Forced from the host: with a limit of
    zero polls every workgroup that reaches a barrier before the last one gives up at its first poll
    (the kernel still runs to its end, with wrong results).  The next flush runs as level launches
    and matches the oracle."""
    from bohrium_amd import _bh_api as api
    prog, n = _batched_prog(16)
    api.set_flush_threshold(10 ** 7)
    bh.message("batch_min: 8")
    bh.message("batch_persistent: on")
    bh.message("persist_spin: 0")
    try:
        with pytest.raises(Exception, match="grid barrier"):
            bh.asnumpy_many(prog())
        assert bh.message("batch_persistent") == "off"
        bh.message("persist_spin: 4194304")
        got, want = run_both(prog)
    finally:
        bh.message("persist_spin: 4194304")
        bh.message("batch_persistent: off")
        api.set_flush_threshold(1000)
    for i in range(n):
        np.testing.assert_array_equal(got[n + i], want[n + i])  # cumsums
        assert_parity(got[2 * n + i], want[2 * n + i], np.float32, rtol=1e-5, atol=1e-6)  # gathers


@pytest.mark.parametrize("shape", [(1000003,), (257, 259), (4099, 1001), (3, 70000)])
def test_stencils_walked_in_xcd_parts(shape):
    """Kernels that read one base through several shifted views (stencils) walk their iteration space
    in one contiguous part per XCD; on odd sizes, 2-D shapes and grids smaller than 8 workgroups the
    results equal the oracle's (fp32 sums of 3 or 5 terms: 1e-6 relative) and, element for element,
    those of the plain grid-stride walk (the same arithmetic per element: bit-exact). The max-abs
    reduction over the stencil equals the plain walk's exactly and the oracle's within 1e-6 relative
    (the fused kernel contracts `sum * 0.2 - c` into one FMA, the oracle rounds the product first)."""
    rng = np.random.RandomState(sum(shape))
    x = rng.uniform(-1.0, 1.0, shape).astype(np.float32)

    def prog():
        a = bh.array(x)
        if len(shape) == 1:
            s = a[:-2] + a[1:-1] * 2.0 + a[2:]
            return [s, bh.ufunc("maximum").reduce(bh.ufunc("absolute")(s - a[1:-1]), 0)]
        c = a[1:-1, 1:-1]
        s = (c + a[:-2, 1:-1] + a[2:, 1:-1] + a[1:-1, :-2] + a[1:-1, 2:]) * 0.2
        d = bh.ufunc("absolute")(s - c)
        return [s, bh.ufunc("maximum").reduce(bh.ufunc("maximum").reduce(d, 1), 0)]

    old = bh.message("codegen:")
    try:
        bh.message("codegen: xcd_parts=0")
        plain = [o.asnumpy() for o in prog()]
        bh.message("codegen: xcd_parts=1")
        got, want = run_both(prog)
    finally:
        bh.message("codegen: " + old)
    np.testing.assert_array_equal(got[0], plain[0])
    np.testing.assert_allclose(got[0], want[0], rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(got[1], plain[1])
    np.testing.assert_allclose(got[1], want[1], rtol=1e-6, atol=0)
