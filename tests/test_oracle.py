"""Pin the oracle (CPU restatement of the reference OpenMP engine) against
golden vectors from the reference's own random123(), NumPy, closed forms and
the reference's documented known answers.  CPU only."""
import ctypes
import json
import os

import numpy as np
import pytest

import bohrium_amd as bh
from bohrium_amd import _bh_api as api
from oracle import omp_ve
from tests.helpers import run_oracle, record

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "philox2x32_10.json")
REF_LIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libr123_ref.so")


def numpy_philox(start, key, index):
    """Independent vectorised philox2x32-10 (Salmon et al. SC'11) in NumPy."""
    ctr = (np.uint64(start) + np.asarray(index, np.uint64)).astype(np.uint64)
    c0 = (ctr & np.uint64(0xFFFFFFFF)).astype(np.uint64)
    c1 = (ctr >> np.uint64(32)).astype(np.uint64)
    k = np.uint64(key & 0xFFFFFFFF)
    for _ in range(10):
        prod = c0 * np.uint64(0xD256D193)
        hi, lo = prod >> np.uint64(32), prod & np.uint64(0xFFFFFFFF)
        c0, c1 = hi ^ k ^ c1, lo
        k = (k + np.uint64(0x9E3779B9)) & np.uint64(0xFFFFFFFF)
    return c0 | (c1 << np.uint64(32))


def golden_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["name"])
def test_oracle_philox_matches_reference_golden(case):
    """The oracle's RANDOM instruction reproduces the reference random123() vectors."""
    idx = np.array(case["index"], dtype=np.int64)
    want = np.array([int(o) for o in case["out"]], dtype=np.uint64)

    def prog():
        n = int(idx.max()) + 1
        if n > 2_000_000:
            # Sparse far indices: use a 1-element view per index (memory offset = index).
            outs = []
            for i in idx:
                a = bh.empty(1, np.uint64)
                api.random123(a.base.handle, 0, (1,), (1,), case["start"] + int(i), case["key"])
                outs.append(a)
            return outs
        a = bh.empty(n, np.uint64)
        api.random123(a.base.handle, 0, (n,), (1,), case["start"], case["key"])
        return [a]

    res = run_oracle(prog)
    if len(res) == 1:
        got = res[0][idx]
    else:
        got = np.array([r[0] for r in res], dtype=np.uint64)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(numpy_philox(case["start"], case["key"], idx), want)


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built (make -C oracle ref)")
def test_oracle_philox_matches_reference_build_random_inputs():
    lib = ctypes.CDLL(REF_LIB)
    lib.ref_random123.restype = ctypes.c_uint64
    lib.ref_random123.argtypes = [ctypes.c_uint64] * 3
    rng = np.random.RandomState(7)
    for _ in range(200):
        start = int(rng.randint(0, 2 ** 63, dtype=np.int64))
        key = int(rng.randint(0, 2 ** 63, dtype=np.int64))
        i = int(rng.randint(0, 2 ** 40, dtype=np.int64))
        assert int(numpy_philox(start, key, [i])[0]) == int(lib.ref_random123(start, key, i))


def test_random_sample_is_exact_power_of_two_scaling():
    """random_sample = float64(u64) / float64(2**64-1) and the divisor rounds to 2**64."""
    assert float(np.float64(np.iinfo(np.uint64).max)) == 2.0 ** 64

    def prog():
        return bh.random.RandomState(42).random_sample(1000)

    got = run_oracle(prog)[0]
    want = numpy_philox(0, 42, np.arange(1000)).astype(np.float64) / 2.0 ** 64
    np.testing.assert_array_equal(got, want)


# ------------------------------------------------------------------ elementwise vs NumPy

INT_TYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64]
FLOAT_TYPES = [np.float32, np.float64]


def _inputs(dtype, n=64, seed=0, positive=False, nonzero=False):
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
    lo = 0.1 if positive else -2.0
    return rng.uniform(lo, 2.0, n).astype(dtype)


BINARY = ["add", "subtract", "multiply", "maximum", "minimum", "greater", "greater_equal", "less",
          "less_equal", "equal", "not_equal"]


@pytest.mark.parametrize("name", BINARY)
@pytest.mark.parametrize("dtype", INT_TYPES + FLOAT_TYPES)
def test_binary_ops_match_numpy(name, dtype):
    x, y = _inputs(dtype, seed=1), _inputs(dtype, seed=2)
    got = run_oracle(lambda: bh.ufunc(name)(bh.array(x), bh.array(y)))[0]
    want = getattr(np, name)(x, y)
    if np.dtype(dtype).kind == "f":
        np.testing.assert_allclose(got, want, rtol=1e-6 if dtype == np.float32 else 1e-15)
    else:
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64])
def test_signed_floor_division_and_remainder(dtype):
    """Reference DIVIDE/REMAINDER on signed ints follow Python/NumPy (instruction.cpp:146-165, 411-437)."""
    x, y = _inputs(dtype, seed=3), _inputs(dtype, seed=4, nonzero=True)
    got_d = run_oracle(lambda: bh.array(x) // bh.array(y))[0]
    got_r = run_oracle(lambda: bh.ufunc("remainder")(bh.array(x), bh.array(y)))[0]
    np.testing.assert_array_equal(got_d, x // y)
    np.testing.assert_array_equal(got_r, np.remainder(x, y))
    # The defining identity of floor division.
    np.testing.assert_array_equal(got_d * y + got_r, x)


@pytest.mark.parametrize("dtype", FLOAT_TYPES)
def test_float_mod_is_fmod_and_remainder_is_python(dtype):
    x, y = _inputs(dtype, seed=5), _inputs(dtype, seed=6, positive=True)
    got_m = run_oracle(lambda: bh.ufunc("fmod")(bh.array(x), bh.array(y)))[0]
    got_r = run_oracle(lambda: bh.ufunc("remainder")(bh.array(x), bh.array(y)))[0]
    tol = 1e-5 if dtype == np.float32 else 1e-12
    np.testing.assert_allclose(got_m, np.fmod(x, y), rtol=tol, atol=tol)
    np.testing.assert_allclose(got_r, x - np.floor(x / y) * y, rtol=tol, atol=tol)


UNARY_FLOAT = ["sqrt", "exp", "log", "sin", "cos", "tan", "tanh", "sinh", "cosh", "arctan", "exp2", "expm1",
               "log2", "log10", "log1p", "floor", "ceil", "trunc", "rint", "absolute", "sign", "arcsinh"]


@pytest.mark.parametrize("name", UNARY_FLOAT)
@pytest.mark.parametrize("dtype", FLOAT_TYPES)
def test_unary_float_ops_match_numpy(name, dtype):
    x = _inputs(dtype, seed=7, positive=name in ("sqrt", "log", "log2", "log10", "log1p"))
    got = run_oracle(lambda: bh.ufunc(name)(bh.array(x)))[0]
    want = getattr(np, name)(x)
    np.testing.assert_allclose(got, want, rtol=2e-6 if dtype == np.float32 else 1e-14, atol=1e-7)


@pytest.mark.parametrize("dtype", INT_TYPES)
def test_integer_bitwise_shift_abs(dtype):
    x, y = _inputs(dtype, seed=8), _inputs(dtype, seed=9, positive=True) % 5
    for name in ["bitwise_and", "bitwise_or", "bitwise_xor"]:
        got = run_oracle(lambda: bh.ufunc(name)(bh.array(x), bh.array(y)))[0]
        np.testing.assert_array_equal(got, getattr(np, name)(x, y))
    got = run_oracle(lambda: bh.ufunc("left_shift")(bh.array(x), bh.array(y)))[0]
    np.testing.assert_array_equal(got, np.left_shift(x, y))
    got = run_oracle(lambda: bh.ufunc("invert")(bh.array(x)))[0]
    np.testing.assert_array_equal(got, np.invert(x))
    got = run_oracle(lambda: bh.ufunc("absolute")(bh.array(x)))[0]
    np.testing.assert_array_equal(got, np.absolute(x))


@pytest.mark.parametrize("src", [np.float64, np.float32, np.int64, np.int32, np.uint8, np.bool_])
@pytest.mark.parametrize("dst", [np.float64, np.float32, np.int64, np.int16, np.uint32, np.bool_])
def test_identity_casts_are_c_casts(src, dst):
    x = _inputs(src, seed=10, positive=np.dtype(dst).kind == "u")
    got = run_oracle(lambda: bh.array(x).astype(dst))[0]
    np.testing.assert_array_equal(got, x.astype(dst))


def test_logical_ops_on_bool():
    x, y = _inputs(np.bool_, seed=11), _inputs(np.bool_, seed=12)
    for name in ["logical_and", "logical_or", "logical_xor"]:
        got = run_oracle(lambda: bh.ufunc(name)(bh.array(x), bh.array(y)))[0]
        np.testing.assert_array_equal(got, getattr(np, name)(x, y))
    got = run_oracle(lambda: bh.ufunc("logical_not")(bh.array(x)))[0]
    np.testing.assert_array_equal(got, ~x)


def test_power_of_integers_goes_through_double():
    x = np.arange(-5, 6, dtype=np.int64)
    got = run_oracle(lambda: bh.array(x) ** 3)[0]
    np.testing.assert_array_equal(got, x ** 3)


def test_complex_arithmetic():
    x, y = _inputs(np.complex128, seed=13), _inputs(np.complex128, seed=14)
    for op in ["add", "subtract", "multiply", "true_divide"]:
        got = run_oracle(lambda: bh.ufunc(op)(bh.array(x), bh.array(y)))[0]
        np.testing.assert_allclose(got, getattr(np, op)(x, y), rtol=1e-13)
    got = run_oracle(lambda: bh.ufunc("exp")(bh.array(x)))[0]
    np.testing.assert_allclose(got, np.exp(x), rtol=1e-13)


# ------------------------------------------------------------------ views, reductions, scans

def _random_view(shape, seed):
    """A random strided view of an array of `shape` (reference test/python/util.py:99-139 style)."""
    rng = np.random.RandomState(seed)
    key = []
    for n in shape:
        start = rng.randint(0, max(1, n // 3))
        step = int(rng.choice([1, 2, -1]))
        if step > 0:
            key.append(slice(start, n, step))
        else:
            key.append(slice(n - 1 - start, None, step))
    return tuple(key)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_reduce_views_every_axis(seed, ndim):
    shape = (7, 5, 6)[:ndim]
    x = np.random.RandomState(seed).uniform(-1, 1, shape)
    key = _random_view(shape, seed)
    xv = x[key]
    for axis in range(ndim):
        got = run_oracle(lambda: bh.array(x)[key].sum(axis=axis))[0]
        np.testing.assert_allclose(got, xv.sum(axis=axis), rtol=1e-13)
    xi = (x * 100).astype(np.int64)
    for axis in range(ndim):
        got = run_oracle(lambda: bh.ufunc("maximum").reduce(bh.array(xi)[key], axis))[0]
        np.testing.assert_array_equal(got, xi[key].max(axis=axis))


@pytest.mark.parametrize("name", ["add", "multiply", "minimum", "maximum", "bitwise_or", "bitwise_xor"])
@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.uint64, np.float64])
def test_reduce_vector_primitives(name, dtype):
    if name.startswith("bitwise") and np.dtype(dtype).kind == "f":
        pytest.skip("bitwise on floats")
    x = _inputs(dtype, n=10, seed=15, positive=True)
    got = run_oracle(lambda: bh.ufunc(name).reduce(bh.array(x)))[0]
    want = getattr(np, name).reduce(x, dtype=x.dtype)  # the bytecode reduces in the input type
    if np.dtype(dtype).kind == "f":
        np.testing.assert_allclose(got, want, rtol=1e-14)
    else:
        np.testing.assert_array_equal(got, want)


def test_reduce_identity_quirks_of_the_reference():
    """sweep_identity (instruction.cpp:546-578) + bh_constant::get_min (bh_constant.cpp:52-55):
    max-reduce of floats starts at the smallest positive normal, and bitwise-and starts at
    the 32-bit all-ones value, so an int64 and-reduce keeps only the low 32 bits."""
    neg = -np.arange(1.0, 6.0)
    got = run_oracle(lambda: bh.array(neg).max())[0]
    assert got == np.finfo(np.float64).tiny
    big = np.array([0x7FFFFFFF00000FFF, 0x0FFFFFFFFFFFFF0F], dtype=np.int64)
    got = run_oracle(lambda: bh.ufunc("bitwise_and").reduce(bh.array(big)))[0]
    assert int(got) == (0x7FFFFFFF00000FFF & 0x0FFFFFFFFFFFFF0F) & 0xFFFFFFFF
    pos = np.arange(1.0, 6.0)
    got = run_oracle(lambda: bh.array(pos).max())[0]
    assert got == 5.0


def test_bool_sum_counts_true_values():
    x = _inputs(np.bool_, n=1000, seed=16)
    got = run_oracle(lambda: bh.array(x).sum())[0]
    assert int(got) == int(x.sum())


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_accumulate_views(axis):
    x = np.random.RandomState(17).uniform(-1, 1, (6, 5, 4))
    key = _random_view(x.shape, 3)
    got = run_oracle(lambda: bh.array(x)[key].cumsum(axis=axis))[0]
    np.testing.assert_allclose(got, np.cumsum(x[key], axis=axis), rtol=1e-13)
    xi = (x * 10).astype(np.int64)
    got = run_oracle(lambda: bh.array(xi)[key].cumprod(axis=axis))[0]
    np.testing.assert_array_equal(got, np.cumprod(xi[key], axis=axis))


def test_triangular_numbers_known_answer():
    """Reference docs acc_loop.rst:56-72: cumulative sum of 1..10."""
    got = run_oracle(lambda: (bh.arange(10) + 1).cumsum())[0]
    np.testing.assert_array_equal(got, [1, 3, 6, 10, 15, 21, 28, 36, 45, 55])


def test_gather_scatter_match_numpy():
    rng = np.random.RandomState(18)
    data = rng.uniform(size=100)
    idx = rng.randint(0, 100, 40).astype(np.int64)
    got = run_oracle(lambda: bh.take(bh.array(data), bh.array(idx)))[0]
    np.testing.assert_array_equal(got, np.take(data, idx))
    uidx = rng.permutation(100)[:30].astype(np.int64)
    vals = rng.uniform(size=30)

    def prog():
        a = bh.array(data)
        bh.put(a, bh.array(uidx), bh.array(vals))
        return a
    got = run_oracle(prog)[0]
    want = data.copy()
    np.put(want, uidx, vals)
    np.testing.assert_array_equal(got, want)
    mask = rng.randint(0, 2, 30).astype(bool)

    def prog2():
        a = bh.array(data)
        bh.put(a, bh.array(uidx), bh.array(vals), mask=bh.array(mask))
        return a
    got = run_oracle(prog2)[0]
    want = data.copy()
    np.put(want, uidx[mask], vals[mask])
    np.testing.assert_array_equal(got, want)


def test_gather_out_of_bounds_is_an_error():
    with pytest.raises(IndexError):
        run_oracle(lambda: bh.take(bh.arange(10), bh.array(np.array([3, 11], np.int64))))


def test_range_is_flat_offset():
    got = run_oracle(lambda: bh.arange(17))[0]
    np.testing.assert_array_equal(got, np.arange(17))


# ------------------------------------------------------------------ repeats

def test_do_while_fixed_count_known_answer():
    """Reference docs acc_loop.rst:29-36: a += 1 five times -> [5 5 5 5]."""
    def prog():
        a = bh.zeros(4)

        def body(a):
            a += 1
        bh.do_while(body, 5, a)
        return a
    np.testing.assert_array_equal(run_oracle(prog)[0], [5, 5, 5, 5])


def test_do_while_condition_known_answer():
    """Reference docs acc_loop.rst:38-46: loop while sum(a) < 10 -> [3 3 3 3]."""
    def prog():
        a = bh.zeros(4)

        def body(a):
            a += 1
            return bh.sum(a) < 10
        bh.do_while(body, None, a)
        return a
    np.testing.assert_array_equal(run_oracle(prog)[0], [3, 3, 3, 3])


# ------------------------------------------------------------------ workloads

def test_mc_pi_small_matches_numpy_restatement():
    from bohrium_amd import workloads
    n = 20000
    got = run_oracle(lambda: workloads.mc_pi(n))[0]
    x = numpy_philox(0, 42, np.arange(n)).astype(np.float64) / 2.0 ** 64
    y = numpy_philox(n, 42, np.arange(n)).astype(np.float64) / 2.0 ** 64
    assert int(got) == int(np.sum(x * x + y * y <= 1.0))


def test_black_scholes_small_matches_closed_form():
    from math import erf, exp, log, sqrt
    from bohrium_amd import workloads
    n = 257
    S, K, T = [run_oracle(lambda: workloads.black_scholes_inputs(n)[i])[0] for i in range(3)]
    got = run_oracle(lambda: workloads.black_scholes(n, steps=3))
    for step in range(3):
        total = 0.0
        for s, k, t in zip(S, K, T + step * 0.01):
            d1 = (log(s / k) + (0.02 + 0.045) * t) / (0.3 * sqrt(t))
            d2 = d1 - 0.3 * sqrt(t)
            cnd = lambda d: 0.5 * (1 + erf(d / sqrt(2)))
            total += s * cnd(d1) - k * exp(-0.02 * t) * cnd(d2)
        # The A&S polynomial approximates the normal CDF to ~7.5e-8 absolute.
        assert abs(float(got[step]) - total) <= 1e-6 * n * 60


def test_jacobi_small_matches_numpy():
    from bohrium_amd import workloads
    n, iters = 34, 5

    def prog():
        G, deltas = workloads.jacobi(n, iters)
        return [G] + deltas
    res = run_oracle(prog)
    G = np.zeros((n, n), np.float32)
    G[:, 0] = -273
    G[:, -1] = -273
    G[-1, :] = -273
    G[0, :] = 40
    for it in range(iters):
        c = G[1:-1, 1:-1]
        tmp = (c + G[:-2, 1:-1] + G[2:, 1:-1] + G[1:-1, :-2] + G[1:-1, 2:]) * np.float32(0.2)
        d = np.abs(tmp - c).max()
        np.testing.assert_allclose(res[1 + it], max(d, np.finfo(np.float32).tiny), rtol=1e-6)
        G[1:-1, 1:-1] = tmp
    np.testing.assert_allclose(res[0], G, rtol=1e-6, atol=1e-4)


def test_sliding_views_triangular_numbers_known_answer():
    """Reference docs acc_loop.rst:56-72: a[i+1] += a[i] with an iterator over 9 repeats."""
    def prog():
        a = bh.array(np.arange(1.0, 11.0))

        def body(a):
            i = bh.get_iterator(1)
            a[i] += a[i - 1]
        bh.do_while(body, 9, a)
        return a
    np.testing.assert_array_equal(run_oracle(prog)[0], [1, 3, 6, 10, 15, 21, 28, 36, 45, 55])


def test_sliding_views_factorial_known_answer():
    """Reference bohrium/loop.py:304-310 docstring: a[i] *= a[i-1] -> [1 2 6 24 120]."""
    def prog():
        a = bh.arange(1, 6)

        def kernel(a):
            i = bh.get_iterator(1)
            a[i] *= a[i - 1]
        bh.do_while(kernel, 4, a)
        return a
    np.testing.assert_array_equal(run_oracle(prog)[0], [1, 2, 6, 24, 120])


def test_black_scholes_repeat_equals_loop_form():
    from bohrium_amd import workloads
    n = 513
    loop = run_oracle(lambda: workloads.black_scholes(n, steps=5))

    def prog():
        S, K, T = workloads.black_scholes_inputs(n)
        return workloads.black_scholes_repeat(S, K, T, steps=5)
    rep = run_oracle(prog)[0]
    np.testing.assert_allclose(rep, np.array(loop).ravel(), rtol=1e-15)
