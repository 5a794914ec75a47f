"""Shared helpers: run a program on the hip engine with the bridge trace recorded,
replay the same trace on the oracle, and compare views."""
import numpy as np

from bohrium_amd import _bh_api as api
from oracle import omp_ve


def oracle_view(data, arr):
    """The oracle's value of BhArray `arr` (same base, offset, shape, strides)."""
    host = data[api._bid(arr.base.handle)]
    if arr.size == 0:
        return np.empty(arr.shape, host.dtype)
    it = host.itemsize
    v = np.lib.stride_tricks.as_strided(host[arr.offset:], shape=arr.shape,
                                        strides=tuple(s * it for s in arr.strides))
    return np.array(v, copy=True)


def record(program):
    """Run `program()` in record-only mode (no engine); returns (results, trace)."""
    with api.record_only() as r:
        res = program()
        api.flush()
    return res, r.trace


def run_oracle(program):
    """Record `program()` and replay it on the oracle; returns a list of numpy results."""
    res, trace = record(program)
    data = omp_ve.replay(trace)
    if not isinstance(res, (list, tuple)):
        res = [res]
    return [oracle_view(data, a) for a in res]


def run_both(program):
    """Run `program()` on the hip engine while tracing, then replay the trace on the
    oracle.  Returns (engine results, oracle results) as numpy arrays."""
    api.trace_start()
    try:
        res = program()
        if not isinstance(res, (list, tuple)):
            res = [res]
        got = [a.asnumpy() for a in res]
    finally:
        trace = api.trace_stop()
    data = omp_ve.replay(trace)
    want = [oracle_view(data, a) for a in res]
    return got, want


def assert_parity(got, want, dtype, rtol=None, atol=0.0):
    """Bit-exact for integer/bool results; within `rtol` (fp64 1e-12, fp32 1e-5 by default) otherwise."""
    dtype = np.dtype(dtype)
    if dtype.kind in "biu":
        np.testing.assert_array_equal(got, want)
        return
    if rtol is None:
        rtol = 1e-12 if dtype in (np.float64, np.complex128) else 1e-5
    np.testing.assert_allclose(got, want, rtol=rtol, atol=atol, equal_nan=True)
