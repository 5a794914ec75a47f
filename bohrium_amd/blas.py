"""Basic Linear Algebra Subprograms on the device (the blas_* extension methods).

Mirrors the reference's bohrium.blas (bridge/npbackend/bohrium/blas.py:12-117): operands are made
C-contiguous, alpha scales A and beta scales C in bytecode, and one extension-method instruction
(C, A[, B]) is queued; the hip engine runs it from libbh_blas_hip.so on the matrix cores.  As in the
reference, the extension method itself computes with alpha 1 and beta 0, so `beta` only affects the
C that is passed in and then overwritten (extmethods/blas/templates/methods.json "options"), and
alpha scales A itself: syrk/herk give alpha^2 A A^T (|alpha|^2 A A^H) and trmm's unit diagonal is
not scaled.  The k-update methods (syrk, herk, syr2k, her2k) write C's upper triangle only; when no
C is given the lower part is zero.
"""
import sys

import numpy as np

from . import _bh_api
from .bharray import BhArray, empty, zeros

_TYPES = (np.float32, np.float64, np.complex64, np.complex128)


def __blas(name, a, b, alpha=1.0, c=None, beta=0.0, shape_matters=True, c_shape=None, c_zero=False):
    if b is not None:
        if not (a.ndim == 2 and b.ndim == 2):
            sys.stderr.write("[ext] Matrices need to be two-dimensional.\n")
            return None
        if a.shape[1] != b.shape[0] and shape_matters:
            sys.stderr.write("[ext] Wrong shape of matrices: first argument has shape {} and second has shape {}.\n"
                             .format(a.shape, b.shape))
            return None
        if not b.iscontiguous():
            b = b.copy()
    if not a.iscontiguous():
        a = a.copy()
    if np.dtype(a.dtype).type not in _TYPES:
        raise TypeError("%s not supported by BLAS for '%s'" % (a.dtype, name))
    if c is None:
        c = (zeros if c_zero else empty)(c_shape or (a.shape[0], b.shape[1]), a.dtype)
    elif not c.iscontiguous():
        c = c.copy()
    if alpha != 1.0:
        a = a * alpha
    if beta != 0.0:
        c = c * beta
    ops = [c.view_desc(), a.view_desc()] + ([b.view_desc()] if b is not None else [])
    _bh_api.extmethod(name, ops)  # writes c
    return c


def gemm(a, b, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A * B"""
    return __blas("blas_gemm", a, b, alpha, c, beta)


def gemmt(a, b, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A^T * B  (A m x k, B m x n, C k x n)"""
    if a.ndim == 2 and b.ndim == 2 and a.shape[0] != b.shape[0]:
        sys.stderr.write("[ext] Wrong shape of matrices: first argument has shape {} and second has shape {}.\n"
                         .format(a.shape, b.shape))
        return None
    return __blas("blas_gemmt", a, b, alpha, c, beta, shape_matters=False,
                  c_shape=(a.shape[1], b.shape[1]) if a.ndim == 2 and b.ndim == 2 else None)


def symm(a, b, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A * B, A symmetric (its upper triangle is read)"""
    return __blas("blas_symm", a, b, alpha, c, beta)


def hemm(a, b, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A * B, A Hermitian (its upper triangle is read, the diagonal as real); complex only"""
    return __blas("blas_hemm", a, b, alpha, c, beta)


def _same_shape(a, b):
    if a.ndim == 2 and b.ndim == 2 and a.shape != b.shape:
        sys.stderr.write("[ext] Wrong shape of matrices: first argument has shape {} and second has shape {}.\n"
                         .format(a.shape, b.shape))
        return False
    return True


def syr2k(a, b, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A * B**T + alpha * B * A**T (upper triangle of C)"""
    if not _same_shape(a, b):
        return None
    return __blas("blas_syr2k", a, b, alpha, c, beta, c_zero=True)


def her2k(a, b, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A * B**H + conj(alpha) * B * A**H (upper triangle of C, real diagonal); complex only.
    As alpha scales A before the call, a complex alpha gives alpha A B^H + B (alpha A)^H."""
    if not _same_shape(a, b):
        return None
    return __blas("blas_her2k", a, b, alpha, c, beta, c_zero=True)


def syrk(a, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A * A**T (upper triangle of C; its strictly lower part keeps the values of the C given)"""
    if a.ndim != 2:
        sys.stderr.write("[ext] Matrices need to be two-dimensional.\n")
        return None
    return __blas("blas_syrk", a, None, alpha, c, beta, c_shape=(a.shape[0], a.shape[0]), c_zero=True)


def herk(a, alpha=1.0, c=None, beta=0.0):
    """C := alpha * A * A**H (upper triangle of C, real diagonal); complex only"""
    if a.ndim != 2:
        sys.stderr.write("[ext] Matrices need to be two-dimensional.\n")
        return None
    return __blas("blas_herk", a, None, alpha, c, beta, c_shape=(a.shape[0], a.shape[0]), c_zero=True)


def _in_place(name, a, b, alpha):
    """B := op(A) B for the unit upper triangular A, written into b itself (through a C-contiguous copy
    when b is not C-contiguous, which the reference would modify and drop)."""
    if not isinstance(b, BhArray) or b.iscontiguous():
        __blas(name, a, b, alpha)
        return b
    tmp = b.copy()
    if __blas(name, a, tmp, alpha) is not None:
        b[...] = tmp
    return b


def trmm(a, b, alpha=1.0):
    """B := alpha * A * B, A unit upper triangular (alpha scales A's strictly upper part; the unit
    diagonal and the lower part of A are not read)"""
    return _in_place("blas_trmm", a, b, alpha)


def trsm(a, b):
    """Solves A * X = B for X, A unit upper triangular; X is written into B"""
    return _in_place("blas_trsm", a, b, 1.0)


def matmul(a, b):
    """2-d matrix product on the matrix cores for float and complex operands."""
    if not (isinstance(a, BhArray) and isinstance(b, BhArray)) or a.ndim != 2 or b.ndim != 2:
        raise ValueError("matmul: two 2-d arrays expected")
    if a.shape[1] != b.shape[0]:
        raise ValueError("matmul: shapes %s and %s do not align" % (a.shape, b.shape))
    if a.dtype != b.dtype:
        t = np.promote_types(a.dtype, b.dtype)
        a, b = a.astype(t), b.astype(t)
    return gemm(a, b)


def _bench(n=8192, reps=3):
    """Time gemm on n x n float64 and float32 operands resident on the device (python -m bohrium_amd.blas)."""
    import time
    from . import random as _random
    for dt, peak in ((np.float64, 78.6), (np.float32, 157.3)):
        R = _random.RandomState(7)
        a = R.random_sample((n, n)).astype(dt)
        b = R.random_sample((n, n)).astype(dt)
        def product():
            # one element of C copied back (a view's asnumpy would copy all of C's base)
            return (gemm(a, b)[:1, :1] * 1.0).asnumpy()
        product()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            product()
            dt_s = time.perf_counter() - t0
            best = dt_s if best is None else min(best, dt_s)
        tf = 2.0 * n ** 3 / best / 1e12
        print("gemm %s n=%d: %.3f ms, %.1f TFLOP/s (%.0f%% of the %.1f TFLOP/s MFMA peak)"
              % (np.dtype(dt).name, n, best * 1e3, tf, 100 * tf / peak, peak), flush=True)
        # the triangular solve: n^3 flops (n x n unit upper triangular A, n right-hand sides)
        def solve():
            return (trsm(a, b.copy())[:1, :1] * 1.0).asnumpy()
        solve()
        t0 = time.perf_counter()
        solve()
        ts = time.perf_counter() - t0
        print("trsm %s n=%d: %.3f ms, %.1f TFLOP/s" % (np.dtype(dt).name, n, ts * 1e3, n ** 3 / ts / 1e12), flush=True)


if __name__ == "__main__":
    _bench()
