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
    assert "C must be m x m" in dry["G7"]
    assert "differ in type" in dry["G8"]


GEMM_SHAPES = [(1, 1, 1), (16, 16, 16), (64, 64, 64), (65, 33, 17), (130, 257, 200), (7, 300, 5), (512, 384, 1),
               (3000, 2050, 129)]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm(dtype, m, n, k):
    import bohrium_amd as bh
    if np.issubdtype(dtype, np.complexfloating) and m * n * k > 10 ** 8:
        pytest.skip("large complex case covered by the real kernels' tiling")
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
