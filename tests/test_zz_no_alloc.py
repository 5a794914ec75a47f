"""bhblas_symm_noalloc, the allocation-free variant of bhblas_symm, through the C-ABI.

It stages the expanded symmetric matrix in a fixed per-device area instead of a stream-ordered
temporary; the GEMM that follows is the same, so its result must equal bhblas_symm's bit for bit.
Two calls in a row (the second smaller) check that the area's leftover contents do not leak into
the next result.  Device memory and the stream come from the engine's own bhhip C-ABI, so the
test process runs one HIP runtime."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "bohrium_amd", "lib")
DTYPES = {np.float32: 0, np.float64: 1, np.complex64: 2, np.complex128: 3}
SCRATCH_BYTES = 32 << 20


def _libs():
    ctypes.CDLL(os.path.join(LIB, "libbh.so"), mode=ctypes.RTLD_GLOBAL)
    hip = ctypes.CDLL(os.path.join(LIB, "libbh_ve_hip.so"), mode=ctypes.RTLD_GLOBAL)
    blas = ctypes.CDLL(os.path.join(LIB, "libbh_blas_hip.so"))
    i64, vp = ctypes.c_int64, ctypes.c_void_p
    for name in ("bhblas_symm", "bhblas_symm_noalloc"):
        f = getattr(blas, name)
        f.argtypes = [ctypes.c_int, i64, i64, vp, i64, vp, i64, vp, i64, vp]
        f.restype = ctypes.c_int
    blas.bhblas_last_error.restype = ctypes.c_char_p
    hip.bhhip_last_error.restype = ctypes.c_char_p
    hip.bhhip_stream.restype = vp
    hip.bhhip_malloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.bhhip_free.argtypes = [vp]
    hip.bhhip_memcpy_h2d.argtypes = [vp, vp, ctypes.c_size_t]
    hip.bhhip_memcpy_d2h.argtypes = [vp, vp, ctypes.c_size_t]
    return hip, blas


def test_oversized_call_is_refused_before_any_device_work():
    """m * m * element size above the 32 MiB staging area: a clear error, no device call."""
    _, blas = _libs()
    m = 2049  # float64: 2049^2 * 8 bytes > 32 MiB
    rc = blas.bhblas_symm_noalloc(1, m, 4, None, m, None, 4, None, 4, None)
    assert rc != 0
    msg = blas.bhblas_last_error().decode()
    assert "33554432" in msg and "m = 2049" in msg, msg


class _Device:
    def __init__(self, hip):
        self.hip = hip
        assert hip.bhhip_init(0) == 0, hip.bhhip_last_error()
        self.ptrs = []

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.hip.bhhip_malloc(ctypes.byref(p), max(a.nbytes, 1)) == 0, self.hip.bhhip_last_error()
        self.ptrs.append(p)
        assert self.hip.bhhip_memcpy_h2d(p, a.ctypes.data, a.nbytes) == 0, self.hip.bhhip_last_error()
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.hip.bhhip_memcpy_d2h(out.ctypes.data, p, out.nbytes) == 0, self.hip.bhhip_last_error()
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.bhhip_free(p)


def _rand(shape, dtype, rng):
    x = rng.uniform(-1, 1, shape)
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * rng.uniform(-1, 1, shape)
    return x.astype(dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
def test_symm_noalloc_equals_symm_twice_then_smaller(dtype):
    hip, blas = _libs()
    dev = _Device(hip)
    stream = hip.bhhip_stream()
    rng = np.random.RandomState(11)
    es = np.dtype(dtype).itemsize
    m_big = int(np.sqrt(SCRATCH_BYTES // es))  # the largest m the staging area holds
    try:
        for m, n in ((m_big, 96), (m_big, 96), (257, 130)):
            a = _rand((m, m), dtype, rng)
            b = _rand((m, n), dtype, rng)
            A, B = dev.upload(a), dev.upload(b)
            C1 = dev.upload(np.full((m, n), np.nan, dtype))
            C2 = dev.upload(np.full((m, n), np.nan, dtype))
            t = DTYPES[dtype]
            assert blas.bhblas_symm(t, m, n, A, m, B, n, C1, n, stream) == 0, blas.bhblas_last_error()
            assert blas.bhblas_symm_noalloc(t, m, n, A, m, B, n, C2, n, stream) == 0, blas.bhblas_last_error()
            assert hip.bhhip_synchronize() == 0, hip.bhhip_last_error()
            want = dev.download(C1, np.empty((m, n), dtype))
            got = dev.download(C2, np.empty((m, n), dtype))
            # same expanded matrix, same GEMM kernel: identical bits
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (m, n)
            # and the product of the symmetric matrix given by a's upper triangle
            up = np.triu(a.astype(np.complex128 if np.iscomplexobj(a) else np.float64))
            full = up + np.triu(up, 1).T
            ref = full @ b.astype(full.dtype)
            bound = np.abs(full) @ np.abs(b.astype(full.dtype))
            tol = 1e-12 if es in (8, 16) and dtype != np.complex64 else 1e-5
            assert (np.abs(got - ref) / bound).max() <= tol, (m, n)
    finally:
        dev.close()
