"""User kernels (tag "hip"): `bohrium_amd.user_kernel.execute` compiles the user's `execute` with hiprtc
and launches it on the operands' device buffers (the reference OpenCL engine's userKernel,
ve/opencl/engine_opencl.cpp:611-669, with bh107's front end, bridge/bh107/bh107/user_kernel.py).

CPU tests run the engine compile-only (BH_HIP_DRY_RUN=1) in a subprocess: errors come back as the
engine's message, like the reference.  GPU tests run the kernels and compare with NumPy."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

ADD = r'''
#include <stdint.h>
extern "C" __global__ void execute(double *a, double *b, double *c) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 100) c[i] = a[i] + b[i] + i;
}'''

DRIVER = r"""
import sys
sys.path.insert(0, %(root)r)
import numpy as np
import bohrium_amd as bh
from bohrium_amd import user_kernel as uk
a = bh.ones(100, np.float64); b = bh.ones(100, np.float64); c = bh.empty(100, np.float64)
def run(src, **param):
    try:
        uk.execute(src, [a, b, c], param=param)
        return "ok"
    except RuntimeError as e:
        return "error: " + str(e).splitlines()[0]
print("R1", run(%(add)r, global_work_size=[100], local_work_size=[64]))
print("R2", run(%(add)r, global_work_size=[100]))
print("R3", run(%(add)r))
print("R4", run(%(add)r, global_work_size=[64, 2, 1, 1], local_work_size=[64, 1, 1, 1]))
print("R5", run(%(add)r, global_work_size=[64, 64], local_work_size=[64, 32]))
print("R6", run("not a kernel", global_work_size=[1], local_work_size=[1]))
try:
    uk.execute(%(add)r, [a, b, c[10:]], param=dict(global_work_size=[1], local_work_size=[1]))
    print("R7 ok")
except TypeError as e:
    print("R7 type error")
print("PROTO", repr(uk.gen_function_prototype([a, b, bh.empty(3, np.int64)])))
"""


@pytest.fixture(scope="module")
def dry_results():
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    r = subprocess.run([sys.executable, "-c", DRIVER % {"root": ROOT, "add": ADD}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    return dict(l.split(" ", 1) for l in r.stdout.splitlines() if l[:1] in "RP" and " " in l)


def test_compiles_and_accepts_a_launch_shape(dry_results):
    assert dry_results["R1"] == "ok"


def test_param_errors_are_messages(dry_results):
    assert "must be the same" in dry_results["R2"]
    assert "global_work_size and local_work_size" in dry_results["R3"]
    assert "one to three dimensions" in dry_results["R4"]
    assert "more than 1024 threads" in dry_results["R5"]


def test_compile_error_is_the_hiprtc_log(dry_results):
    assert dry_results["R6"].startswith("error: [HIP] hiprtc compile failed")


def test_non_behaving_operand_is_refused(dry_results):
    assert dry_results["R7"] == "type error"


def test_prototype(dry_results):
    assert "execute(double *a0, double *a1, int64_t *a2)" in dry_results["PROTO"]


@pytest.mark.gpu
def test_add_kernel_on_device():
    import bohrium_amd as bh
    x = np.arange(100, dtype=np.float64)
    a = bh.array(x) * 2.0  # lazily computed operand: flushed before the kernel runs
    b = bh.ones(100, np.float64)
    c = bh.empty(100, np.float64)
    bh.user_kernel.execute(ADD, [a, b, c], param={"global_work_size": [100], "local_work_size": [64]})
    np.testing.assert_array_equal(c.asnumpy(), 2.0 * x + 1.0 + np.arange(100))
    # the kernel's output feeds later bytecode
    np.testing.assert_array_equal((c + 1.0).sum().asnumpy(), (2.0 * x + 2.0 + np.arange(100)).sum())


@pytest.mark.gpu
def test_2d_launch_on_device():
    import bohrium_amd as bh
    src = r'''
extern "C" __global__ void execute(float *m, long *rows) {
    const unsigned x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x < 300 && y < 70) m[y * 300 + x] = (float)(y * 1000 + x);
    if (x == 0 && y < 70) rows[y] = y;
}'''
    m = bh.zeros((70, 300), np.float32)
    rows = bh.zeros(70, np.int64)
    bh.user_kernel.execute(src, [m, rows], param={"global_work_size": [300, 70], "local_work_size": [64, 4]})
    y, x = np.mgrid[0:70, 0:300]
    np.testing.assert_array_equal(m.asnumpy(), (y * 1000 + x).astype(np.float32))
    np.testing.assert_array_equal(rows.asnumpy(), np.arange(70))
