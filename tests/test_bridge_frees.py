"""The bridge runtime deletes a freed base only after the flush that executes its BH_FREE.

Both cases run the bhc C-ABI in compile-only mode (BH_HIP_DRY_RUN=1: kernels are planned and
compiled, nothing touches a device) in a child process, with glibc's MALLOC_PERTURB_ filling freed
heap blocks so that a read of a deleted base shows up as a wrong dtype or size.

1. Auto-flush at the threshold.  With 1000 instructions queued, bhc_free's BH_FREE is the one that
   triggers the auto-flush.  The flush runs the 1000 earlier instructions; the base must survive it
   because its BH_FREE is only in the next batch (the reference queues before it flushes for this
   reason, bridge/cxx/src/Runtime.cpp:33-42).
2. A flush that throws.  The bases freed in it must still reach the engine's release before the
   bridge deletes them; otherwise a new base at the same address would inherit a stale device buffer.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = r"""
import ctypes, sys
sys.path.insert(0, %(root)r)
from bohrium_amd import _bh_api as api, opcodes as op
lib = api.lib()
F64 = 10

def view(base, n):
    v = api.BhcView()
    v.base = base
    v.start = 0
    v.ndim = 1
    v.shape[0] = n
    v.stride[0] = 1
    return v

def const_view():
    v = api.BhcView()
    v.base = None
    v.ndim = 0
    return v

def check(rc):
    if rc != 0:
        raise RuntimeError(lib.bhc_last_error().decode())

def stat(name):
    for line in api.message("statistic").splitlines():
        if line.startswith(name + ":"):
            return int(float(line.split(":")[1]))
    raise KeyError(name)

one = ctypes.c_double(1.0)
"""

AT_THRESHOLD = PRELUDE + r"""
check(lib.bhc_set_flush_threshold(1000))
keep = lib.bhc_new(F64, 4096)
victim = lib.bhc_new(F64, 64)
ops = (api.BhcView * 2)(view(keep, 4096), const_view())
for i in range(999):
    check(lib.bhc_op(op.IDENTITY, 2, ops, F64, ctypes.byref(one)))
vops = (api.BhcView * 2)(view(victim, 64), const_view())
check(lib.bhc_op(op.IDENTITY, 2, vops, F64, ctypes.byref(one)))
# 1000 instructions queued: this free's BH_FREE triggers the auto-flush of the first 1000.
check(lib.bhc_free(victim))
check(lib.bhc_flush())
check(lib.bhc_free(keep))
check(lib.bhc_flush())
print("OK releases", stat("releases"), "missed", stat("missed_frees"), flush=True)
"""

AFTER_FAILURE = PRELUDE + r"""
check(lib.bhc_set_flush_threshold(10 ** 7))
x = lib.bhc_new(F64, 4096)
y = lib.bhc_new(F64, 4096)
check(lib.bhc_flush())
r0 = stat("releases")
check(lib.bhc_op(op.IDENTITY, 2, (api.BhcView * 2)(view(x, 4096), const_view()), F64, ctypes.byref(one)))
# A shift of float64 operands has no C expression: the kernel fails to compile and the flush throws.
check(lib.bhc_op(op.LEFT_SHIFT, 3, (api.BhcView * 3)(view(y, 4096), view(x, 4096), view(x, 4096)), 0, None))
check(lib.bhc_free(x))
rc = lib.bhc_flush()
assert rc != 0, "the flush was expected to fail"
print("ERR", lib.bhc_last_error().decode().splitlines()[0][:120], flush=True)
r1 = stat("releases")
# The engine is still usable afterwards.
z = lib.bhc_new(F64, 4096)
check(lib.bhc_op(op.IDENTITY, 2, (api.BhcView * 2)(view(z, 4096), const_view()), F64, ctypes.byref(one)))
check(lib.bhc_free(z))
check(lib.bhc_free(y))
check(lib.bhc_flush())
print("OK released_after_failure", r1 - r0, flush=True)
"""


def _run(script):
    env = dict(os.environ, BH_HIP_DRY_RUN="1", MALLOC_PERTURB_="165")
    return subprocess.run([sys.executable, "-c", script % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=600)


def test_free_that_triggers_the_auto_flush_outlives_it():
    r = _run(AT_THRESHOLD)
    assert r.returncode == 0, r.stdout[-3000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("OK")]
    assert line, r.stdout[-3000:]
    assert line[0].endswith("missed 0"), line[0]


def test_bases_freed_in_a_failed_flush_are_released():
    r = _run(AFTER_FAILURE)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "ERR" in r.stdout, r.stdout[-3000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("OK")]
    assert line and line[0] == "OK released_after_failure 1", r.stdout[-3000:]
