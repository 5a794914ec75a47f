"""Host copies the engine hands out may be written by the caller (MI355X).

A caller that syncs an array and takes its host pointer (bhc_data_get with copy2host, the
reference's getMemoryPointer, bh_component.hpp:188-199) may write through it, as the reference's
NumPy bridge does when an array returns from NumPy (bridge/npbackend/src/memory.c:265-283); the
next kernel must read the written values.  The reference GPU engines drop the device copy on every
host sync for this reason (ve/cuda/main.cpp:91-108).  This engine either does the same
(host_write_tracking off) or keeps the device copy and hands out read-only host pages whose first
write is caught by bh_mem_signal (host_write_tracking on, the default), so an array that is only
read on the host is not uploaded again."""
import ctypes

import numpy as np
import pytest

import bohrium_amd as bh
from bohrium_amd import _bh_api as api

pytestmark = pytest.mark.gpu

N = 100_003


def stat(name):
    for line in bh.message("statistic").splitlines():
        if line.startswith(name + ":"):
            return int(float(line.split(":")[1]))
    raise KeyError(name)


@pytest.fixture(params=["on", "off"])
def tracking(request):
    bh.message("host_write_tracking: " + request.param)
    yield request.param
    bh.message("host_write_tracking: on")


def host_pointer(arr, nullify=False):
    """Sync `arr`'s base and return its host data as a NumPy array over the engine's memory."""
    api.sync(arr.base.handle)
    api.flush()
    p = api.data_get(arr.base.handle, copy2host=True, force_alloc=True, nullify=nullify)
    assert p
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(arr.base.nelem,))


def test_write_through_the_host_pointer_reaches_later_kernels(tracking):
    base = np.arange(N, dtype=np.float64)
    a = bh.array(base) * 2.0
    host = host_pointer(a)
    np.testing.assert_array_equal(host, base * 2.0)
    faults = stat("host_write_faults")
    host[::7] = -1.0
    want = base * 2.0
    want[::7] = -1.0
    np.testing.assert_array_equal((a + 1.0).asnumpy(), want + 1.0)
    np.testing.assert_array_equal(a.asnumpy(), want)
    if tracking == "on":
        assert stat("host_write_faults") == faults + 1
    # A kernel writes the array again: the device copy is the newest once more.
    a *= 3.0
    np.testing.assert_array_equal(a.asnumpy(), want * 3.0)
    host = host_pointer(a)
    host[1] = 5.0
    want = want * 3.0
    want[1] = 5.0
    np.testing.assert_array_equal((a - 0.5).asnumpy(), want - 0.5)


def test_host_reads_keep_the_device_copy(tracking):
    a = bh.array(np.ones(N)) * 3.0
    a.asnumpy()
    before = stat("h2d_bytes")
    np.testing.assert_array_equal((a + 1.0).asnumpy(), np.full(N, 4.0))
    uploaded = stat("h2d_bytes") - before
    if tracking == "on":
        assert uploaded == 0
    else:
        assert uploaded == N * 8


def test_host_pointer_taken_over_is_not_guarded():
    a = bh.array(np.arange(N, dtype=np.float64)) + 1.0
    faults = stat("host_write_faults")
    host = host_pointer(a, nullify=True)
    host[:] = 0.0  # the caller owns these pages now: no fault, no effect on the engine
    assert stat("host_write_faults") == faults
