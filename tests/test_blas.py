"""The blas_* extension methods (libbh_blas_hip.so) against NumPy.

The reference computes them with CBLAS (extmethods/blas/templates/body_func.tpl:1-30); its own tests
compare with numpy.dot (test/python/tests/test_ext_blas.py style).  Tolerances: every element of C
within 1e-12 (float64, complex128) or 1e-5 (float32, complex64) of the float64 product, relative to
(|A| |B|)[i, j], which bounds the rounding of any summation order.

CPU tests run the engine compile-only (BH_HIP_DRY_RUN=1): shape and type checks of the methods."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

TOL = {np.float32: 1e-5, np.float64: 1e-12, np.complex64: 1e-5, np.complex128: 1e-12}


def _rand(shape, dtype, seed):
    r = np.random.RandomState(seed)
    x = r.uniform(-1, 1, shape)
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * r.uniform(-1, 1, shape)
    return x.astype(dtype)


def _check(got, a, b, dtype):
    a64 = a.astype(np.complex128 if np.issubdtype(dtype, np.complexfloating) else np.float64)
    b64 = b.astype(a64.dtype)
    ref = a64 @ b64
    bound = np.abs(a64) @ np.abs(b64) + 1e-300
    err = np.abs(got.astype(a64.dtype) - ref) / bound
    assert err.max() <= TOL[dtype], "max relative error %g" % err.max()


DRY = r"""
import sys
sys.path.insert(0, %(root)r)
import numpy as np
import bohrium_amd as bh
from bohrium_amd import _bh_api
def run(name, ops):
    try:
        _bh_api.extmethod(name, [o.view_desc() for o in ops]); bh.flush(); return "ok"
    except Exception as e:
        return "error: " + str(e).splitlines()[0]
f = lambda *s: bh.empty(s, np.float64)
print("G1", run("blas_gemm", [f(5, 7), f(5, 3), f(3, 7)]))
print("G2", run("blas_gemm", [f(5, 7), f(5, 3), f(4, 7)]))
print("G3", run("blas_gemm", [bh.empty((5, 7), np.int32), bh.empty((5, 3), np.int32), bh.empty((3, 7), np.int32)]))
print("G4", run("blas_gemm", [f(5, 7), f(5, 3), f(3, 14)[:, ::2]]))
print("G5", run("blas_gemmt", [f(3, 7), f(5, 3), f(5, 7)]))
print("G6", run("blas_syrk", [f(5, 5), f(5, 3)]))
print("G7", run("blas_syrk", [f(5, 4), f(5, 3)]))
print("G8", run("blas_gemm", [f(5, 7), f(5, 3), bh.empty((3, 7), np.float32)]))
z = lambda *s: bh.empty(s, np.complex128)
print("G9", run("blas_symm", [f(5, 7), f(5, 5), f(5, 7)]))
print("G10", run("blas_symm", [f(5, 7), f(5, 4), f(4, 7)]))
print("G11", run("blas_hemm", [f(5, 7), f(5, 5), f(5, 7)]))
print("G12", run("blas_hemm", [z(5, 7), z(5, 5), z(5, 7)]))
print("G13", run("blas_herk", [z(5, 5), z(5, 3)]))
print("G14", run("blas_herk", [f(5, 5), f(5, 3)]))
print("G15", run("blas_syr2k", [f(5, 5), f(5, 3), f(5, 3)]))
print("G16", run("blas_syr2k", [f(5, 5), f(5, 3), f(3, 5)]))
print("G17", run("blas_her2k", [z(5, 5), z(5, 3), z(5, 3)]))
print("G18", run("blas_trmm", [f(5, 7), f(5, 5), f(5, 7)]))
print("G19", run("blas_trsm", [f(5, 7), f(5, 5), f(4, 7)]))
print("G20", run("blas_trsm", [f(5, 7), f(5, 5), f(5, 7)]))
"""


@pytest.fixture(scope="module")
def dry():
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    r = subprocess.run([sys.executable, "-c", DRY % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    return dict(l.split(" ", 1) for l in r.stdout.splitlines() if l.startswith("G"))


def test_extension_methods_load_and_check_shapes(dry):
    assert dry["G1"] == "ok" and dry["G5"] == "ok" and dry["G6"] == "ok"
    assert "shapes do not chain" in dry["G2"]
    assert "not supported by BLAS" in dry["G3"]
    assert "C-contiguous 2-d" in dry["G4"]
    assert "C must be n x n" in dry["G7"]
    assert "differ in type" in dry["G8"]
    for g in ("G9", "G12", "G13", "G15", "G17", "G18", "G20"):
        assert dry[g] == "ok", (g, dry[g])
    assert "shapes do not chain" in dry["G10"] and "shapes do not chain" in dry["G16"]
    assert "shapes do not chain" in dry["G19"]
    # the Hermitian methods exist for complex types only, as in the reference's methods table
    assert "BH_FLOAT64 not supported by BLAS for 'blas_hemm'" in dry["G11"]
    assert "not supported by BLAS for 'blas_herk'" in dry["G14"]


GEMM_SHAPES = [(1, 1, 1), (16, 16, 16), (64, 64, 64), (65, 33, 17), (130, 257, 200), (7, 300, 5), (512, 384, 1),
               (3000, 2050, 129)]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm(dtype, m, n, k):
    import bohrium_amd as bh
    a, b = _rand((m, k), dtype, 1), _rand((k, n), dtype, 2)
    c = bh.blas.gemm(bh.array(a), bh.array(b))
    got = c.asnumpy()
    assert got.dtype == dtype and got.shape == (m, n)
    _check(got, a, b, dtype)


@pytest.mark.gpu
def test_gemm_alpha_and_views():
    import bohrium_amd as bh
    a, b = _rand((40, 70), np.float64, 3), _rand((70, 90), np.float64, 4)
    A = bh.array(a)
    # a transposed view is copied to a contiguous operand first; alpha scales A
    c = bh.blas.gemm(bh.array(np.ascontiguousarray(a.T)).T, bh.array(b), alpha=2.5)
    _check(c.asnumpy() / 2.5, a, b, np.float64)
    b32 = b.astype(np.float32)
    m = bh.matmul(A, bh.array(b32))  # promoted to float64
    _check(m.asnumpy(), a, b32.astype(np.float64), np.float64)
    # the product feeds further bytecode in the same flush sequence
    s = (bh.blas.gemm(A, bh.array(b)) * 2.0).sum().asnumpy()
    assert abs(s - 2.0 * (a @ b).sum()) <= 1e-9 * np.abs(a).sum() * np.abs(b).sum()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,k,n", [(50, 30, 20), (129, 257, 65), (1000, 64, 1000)])
def test_gemmt(dtype, m, k, n):
    import bohrium_amd as bh
    a, b = _rand((m, k), dtype, 5), _rand((m, n), dtype, 6)
    got = bh.blas.gemmt(bh.array(a), bh.array(b)).asnumpy()
    assert got.shape == (k, n)
    _check(got, np.ascontiguousarray(a.T), b, dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128])
@pytest.mark.parametrize("m,k", [(33, 17), (300, 129), (2100, 64)])
def test_syrk_upper_triangle(dtype, m, k):
    import bohrium_amd as bh
    a = _rand((m, k), dtype, 7)
    c0 = np.full((m, m), -7.0).astype(dtype)
    got = bh.blas.syrk(bh.array(a), c=bh.array(c0)).asnumpy()
    ref_mask = np.triu(np.ones((m, m), bool))
    a64 = a.astype(np.complex128 if np.issubdtype(dtype, np.complexfloating) else np.float64)
    ref = a64 @ a64.T
    bound = np.abs(a64) @ np.abs(a64).T + 1e-300
    assert (np.abs(got.astype(ref.dtype) - ref) / bound)[ref_mask].max() <= TOL[dtype]
    # the strictly lower triangle keeps the C that was passed in
    il = np.tril_indices(m, -1)
    assert np.all(got[il] == c0[il])


# ------------------------------------------- the rest of the reference's methods table


def _wide(x):
    return x.astype(np.complex128 if np.issubdtype(x.dtype, np.complexfloating) else np.float64)


def _close(got, ref, bound, dtype, mask=None):
    err = np.abs(_wide(got) - ref) / (bound + 1e-300)
    if mask is not None:
        err = err[mask]
    assert err.max() <= TOL[dtype], "max relative error %g" % err.max()


def _sym(a):
    """The symmetric matrix whose upper triangle is a's."""
    u = np.triu(a, 1)
    return u + u.T + np.diag(np.diag(a))


def _herm(a):
    u = np.triu(a, 1)
    return u + u.conj().T + np.diag(np.diag(a).real)


def _unit_upper(a):
    return np.triu(a, 1) + np.eye(a.shape[0], dtype=a.dtype)


ALL = [np.float64, np.float32, np.complex128, np.complex64]
COMPLEX = [np.complex128, np.complex64]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ALL)
def test_reference_suites(dtype):
    """test/python/tests/test_ext_blas.py's cases: arange matrices of every side up to 19 (gemm, gemmt,
    syr2k, her2k), 2..9 (symm, hemm, syrk, herk, trmm, trsm)."""
    import bohrium_amd as bh
    cplx = np.issubdtype(dtype, np.complexfloating)
    for r in range(1, 20):
        a = np.arange(r * r, dtype=dtype).reshape(r, r)
        b = np.arange(r * r, dtype=dtype).reshape(r, r)
        A, B = bh.array(a), bh.array(b)
        a64, b64 = _wide(a), _wide(b)
        _close(bh.blas.gemm(A, B).asnumpy(), a64 @ b64, np.abs(a64) @ np.abs(b64), dtype)
        if not cplx:
            _close(bh.blas.gemmt(A, B).asnumpy(), a64.T @ b64, np.abs(a64.T) @ np.abs(b64), dtype)
        ref = np.triu(a64 @ b64.T + b64 @ a64.T)
        bound = np.abs(a64) @ np.abs(b64).T + np.abs(b64) @ np.abs(a64).T
        _close(bh.blas.syr2k(A, B).asnumpy(), ref, bound, dtype)
        if cplx:  # real-valued data: B^H = B^T
            _close(bh.blas.her2k(A, B).asnumpy(), ref, bound, dtype)
    for r in range(2, 10):
        a = np.arange(r * r, dtype=dtype).reshape(r, r)
        a = (a + a.T) / 2
        b = np.arange(r * r, dtype=dtype).reshape(r, r)
        a64, b64 = _wide(a), _wide(b)
        ref, bound = a64 @ b64, np.abs(a64) @ np.abs(b64)
        _close(bh.blas.symm(bh.array(a), bh.array(b)).asnumpy(), ref, bound, dtype)
        if cplx:
            _close(bh.blas.hemm(bh.array(a), bh.array(b)).asnumpy(), ref, bound, dtype)
        a = np.arange(r * r, dtype=dtype).reshape(r, r)
        a64 = _wide(a)
        res = bh.zeros_like(bh.array(a))
        bh.blas.syrk(bh.array(a), c=res)
        _close(res.asnumpy(), np.triu(a64 @ a64.T), np.abs(a64) @ np.abs(a64).T, dtype)
        if cplx:
            res = bh.zeros_like(bh.array(a))
            bh.blas.herk(bh.array(a), c=res)
            _close(res.asnumpy(), np.triu(a64 @ a64.T), np.abs(a64) @ np.abs(a64).T, dtype)
        a = np.arange(r * r, dtype=dtype).reshape(r, r)
        np.fill_diagonal(a, 1)
        a = np.triu(a)
        a64 = _wide(a)
        got = bh.blas.trmm(bh.array(a), bh.array(b)).asnumpy()
        _close(got, a64 @ b64, np.abs(a64) @ np.abs(b64), dtype)
        x = bh.blas.trsm(bh.array(a), bh.array(b)).asnumpy()
        # the solve checked by its residual A X - B against |A| |X| + |B| (inv(a) @ b grows too fast
        # with r for an elementwise comparison in float32)
        _close(a64 @ _wide(x), b64, np.abs(a64) @ np.abs(_wide(x)) + np.abs(b64), dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ALL)
@pytest.mark.parametrize("m,n", [(33, 17), (300, 129), (1100, 700)])
def test_symm_hemm_read_the_upper_triangle(dtype, m, n):
    import bohrium_amd as bh
    a, b = _rand((m, m), dtype, 11), _rand((m, n), dtype, 12)
    s64, b64 = _sym(_wide(a)), _wide(b)
    _close(bh.blas.symm(bh.array(a), bh.array(b)).asnumpy(), s64 @ b64, np.abs(s64) @ np.abs(b64), dtype)
    if np.issubdtype(dtype, np.complexfloating):
        h64 = _herm(_wide(a))
        _close(bh.blas.hemm(bh.array(a), bh.array(b)).asnumpy(), h64 @ b64, np.abs(h64) @ np.abs(b64), dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ALL)
@pytest.mark.parametrize("n,k", [(33, 17), (300, 129), (1030, 64)])
def test_rank_k_updates(dtype, n, k):
    """syr2k (and her2k, herk for complex): the upper triangle, the lower one keeps C, and the
    Hermitian results have an exactly real diagonal.  The front end, like the reference's, takes
    square A and B only for the two-operand updates (it requires a.shape[1] == b.shape[0]); the
    n x k case is queued as the extension-method instruction directly."""
    import bohrium_amd as bh
    from bohrium_amd import _bh_api

    def direct(name, a, b, c):
        ops = [bh.array(c), bh.array(a), bh.array(b)]  # alive until the instruction is queued
        _bh_api.extmethod(name, [o.view_desc() for o in ops])
        return ops[0]

    a, b = _rand((n, k), dtype, 13), _rand((n, k), dtype, 14)
    a64, b64 = _wide(a), _wide(b)
    up = np.triu(np.ones((n, n), bool))
    il = np.tril_indices(n, -1)
    c0 = np.full((n, n), 3.0).astype(dtype)
    assert bh.blas.syr2k(bh.array(a), bh.array(b)) is None  # n x k with n != k: rejected
    got = direct("blas_syr2k", a, b, c0).asnumpy()
    _close(got, a64 @ b64.T + b64 @ a64.T, np.abs(a64) @ np.abs(b64).T + np.abs(b64) @ np.abs(a64).T, dtype, up)
    assert np.all(got[il] == c0[il])
    if np.issubdtype(dtype, np.complexfloating):
        got = direct("blas_her2k", a, b, c0).asnumpy()
        ref = a64 @ b64.conj().T + b64 @ a64.conj().T
        _close(got, ref, np.abs(a64) @ np.abs(b64).T + np.abs(b64) @ np.abs(a64).T, dtype, up)
        assert np.all(np.diag(got).imag == 0) and np.all(got[il] == c0[il])
        got = bh.blas.herk(bh.array(a), c=bh.array(c0)).asnumpy()
        _close(got, a64 @ a64.conj().T, np.abs(a64) @ np.abs(a64).T, dtype, up)
        assert np.all(np.diag(got).imag == 0) and np.all(got[il] == c0[il])
        # no C given: the lower triangle is zero
        got = bh.blas.herk(bh.array(a)).asnumpy()
        assert np.all(got[il] == 0)
    # the square case through the front end, with alpha scaling A (alpha^2 for syrk)
    sq, sq2 = _rand((k, k), dtype, 19), _rand((k, k), dtype, 20)
    s64, s264 = _wide(sq), _wide(sq2)
    got = bh.blas.syr2k(bh.array(sq), bh.array(sq2), alpha=2.0).asnumpy()
    ref = 2.0 * (s64 @ s264.T + s264 @ s64.T)
    _close(got, np.triu(ref), 2.0 * (np.abs(s64) @ np.abs(s264).T + np.abs(s264) @ np.abs(s64).T), dtype)
    got = bh.blas.syrk(bh.array(sq), alpha=2.0).asnumpy()
    _close(got, np.triu(4.0 * s64 @ s64.T), 4.0 * np.abs(s64) @ np.abs(s64).T, dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ALL)
@pytest.mark.parametrize("m,n", [(1, 5), (63, 10), (64, 64), (130, 257), (300, 129), (1000, 70)])
def test_trmm_trsm(dtype, m, n):
    """trmm and trsm with a unit upper triangular A whose diagonal and lower part hold noise (never
    read); the solve is checked by its residual, and trsm(trmm(B)) returns B."""
    import bohrium_amd as bh
    a = (_rand((m, m), dtype, 15) / np.sqrt(m)).astype(dtype)  # this scale keeps A well conditioned
    a[np.diag_indices(m)] = 9.0
    b = _rand((m, n), dtype, 16)
    u64, b64 = _unit_upper(_wide(a)), _wide(b)
    B = bh.array(b)
    out = bh.blas.trmm(bh.array(a), B)
    assert out is B
    prod = B.asnumpy()
    _close(prod, u64 @ b64, np.abs(u64) @ np.abs(b64), dtype)
    x = bh.blas.trsm(bh.array(a), bh.array(b)).asnumpy()
    _close(u64 @ _wide(x), b64, np.abs(u64) @ np.abs(_wide(x)) + np.abs(b64), dtype)
    back = bh.blas.trsm(bh.array(a), bh.array(prod)).asnumpy()
    scale = np.abs(np.linalg.inv(u64)) @ (np.abs(u64) @ np.abs(b64))
    _close(back, b64, scale, dtype) if m <= 300 else None


@pytest.mark.gpu
def test_trmm_alpha_and_strided_b():
    """alpha scales A before the call (its unit diagonal stays 1); a strided B is written back."""
    import bohrium_amd as bh
    a, b = _rand((40, 40), np.float64, 17), _rand((40, 60), np.float64, 18)
    big = bh.array(np.zeros((40, 120)))
    view = big[:, ::2]
    view[...] = bh.array(b)
    bh.blas.trmm(bh.array(a), view, alpha=2.0)
    u = np.triu(2.0 * a, 1) + np.eye(40)
    got = big.asnumpy()
    _close(got[:, ::2], u @ b, np.abs(u) @ np.abs(b), np.float64)
    assert np.all(got[:, 1::2] == 0)


@pytest.mark.gpu
def test_gemm_inside_a_repeated_body():
    """An extension method inside a do_while body runs in every iteration: X := A X, three times in
    one flush (the engine takes extension methods out of a flush's instruction list to run them,
    so a repeated body must get them back for the next iteration).  Bound: each product of 64
    terms rounds within 64 * 2^-53 of the exact one, relative to |A| |X|."""
    import bohrium_amd as bh
    a = _rand((64, 64), np.float64, 5) * 0.1
    x0 = _rand((64, 32), np.float64, 6)
    A, X = bh.array(a), bh.array(x0)

    def body():
        X[...] = bh.blas.gemm(A, X)

    bh.do_while(body, 3)
    got = X.asnumpy()
    want = a @ (a @ (a @ x0))
    mag = np.abs(a) @ (np.abs(a) @ (np.abs(a) @ np.abs(x0)))
    assert np.all(np.abs(got - want) <= 3 * 64 * 2.0 ** -53 * mag + 1e-300)
    assert not np.allclose(got, a @ x0)  # more than one iteration ran
