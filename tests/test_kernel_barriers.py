"""Workgroup barriers in the generated kernels on the CPU: compile the kernels of the benchmark
workloads compile-only (hiprtc for gfx950; the batched and persistent forms too) into a fresh
kernel cache, disassemble each code object with the ROCm llvm-objdump, and check that no loop whose
exit depends on a lane-varying value contains an s_barrier.

Why: when a loop that holds a workgroup barrier can be left by some lanes of a wave and not others,
the compiler keeps the other lanes looping under a narrowed exec mask, and the wave executes the
barrier a different number of times than its neighbours: the workgroup never leaves it.  The
persistent kernel's task loop and grid barrier, and every skeleton with a barrier, must therefore
branch only on wave-uniform values.  In the disassembly a lane-varying loop exit is a latch that
narrows exec (s_andn2_b64 exec, exec, sN) followed by the loop's back-edge (the first backward
branch after it other than an exec-zero exit to the join block); the loop spans from the back-edge's
target to the back-edge.  The check is validated on a kernel written with that mistake."""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"
HIPCC = "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(not (os.path.exists(OBJDUMP) and os.path.exists(HIPCC)),
                                reason="ROCm llvm-objdump / hipcc not installed")

_INS = re.compile(r'\s+(\S.*?)\s*//\s*([0-9A-F]+):.*?(?:<(\w+)\+0x([0-9a-f]+)>)?$')
_SYM = re.compile(r'([0-9a-f]+) <(\w+)>:')


def _instructions(text):
    """(symbol, offset, instruction text, branch target offset or None) per disassembled line."""
    out, base, sym = [], None, None
    for line in text.splitlines():
        m = _SYM.match(line)
        if m:
            base, sym = int(m.group(1), 16), m.group(2)
            continue
        m = _INS.match(line)
        if m and base is not None:
            tgt = int(m.group(4), 16) if m.group(3) == sym else None
            out.append((sym, int(m.group(2), 16) - base, m.group(1), tgt))
    return out


def divergent_loops_with_barriers(text):
    """Loops closed after an exec-narrowing latch that contain s_barrier: [(symbol, latch, start, end)]."""
    ins = _instructions(text)
    bad = []
    for i, (sym, a, t, _) in enumerate(ins):
        if not t.startswith("s_andn2_b64 exec, exec"):
            continue
        for sym2, b, u, tgt in ins[i + 1:]:
            if sym2 != sym:
                break
            if u.startswith("s_cbranch_execz"):
                continue  # lanes all done: exit to the join block, which restores exec
            if tgt is not None and tgt <= a and (u.startswith("s_branch") or u.startswith("s_cbranch")):
                n = sum(1 for s3, c, v, _ in ins if s3 == sym and tgt <= c <= b and v.startswith("s_barrier"))
                if n:
                    bad.append((sym, a, tgt, b))
                break
    return bad


def _disassemble(path):
    r = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", path], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


BAD_KERNEL = r"""
#include <hip/hip_runtime.h>
// A ticket loop whose exit tests a value read from LDS per lane (not known to be uniform), with a
// thread-0 region on either side of the barriers.
__global__ void __launch_bounds__(256) ticket_loop(const float *x, float *y, unsigned n, unsigned *ticket,
                                                   unsigned *done) {
  __shared__ unsigned t_s;
  for (;;) {
    if (threadIdx.x == 0) t_s = atomicAdd(ticket, 1u);
    __syncthreads();
    const unsigned t = t_s;
    if (t >= n) break;
    for (int j = 0; j < 16; ++j) {
      const size_t i = (size_t)t * 4096 + j * 256 + threadIdx.x;
      y[i] = x[i] * 2.0f + 1.0f;
    }
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(done + t, 1u);
  }
}
"""


def test_check_finds_a_barrier_in_a_divergent_loop(tmp_path):
    src = tmp_path / "bad.hip"
    src.write_text(BAD_KERNEL)
    obj = tmp_path / "bad.o"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "--offload-device-only", "--no-gpu-bundle-output",
                        "-c", "-o", str(obj), str(src)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout
    assert divergent_loops_with_barriers(_disassemble(str(obj)))


DRIVER = r"""
import sys
import numpy as np
sys.path.insert(0, %(root)r)
import bohrium_amd as bh
from bohrium_amd import workloads, _bh_api
_bh_api.set_flush_threshold(10 ** 7)
bh.message("batch_persistent: on")
sizes = workloads.batched_sizes(40, 256, 40000, seed=7)
xs, idxs = workloads.batched_inputs(sizes, seed=7)
bh.flush()
sums, cums, gathered = workloads.batched_step(xs, idxs)
bh.flush()
S, K, T = workloads.black_scholes_inputs(1000)
workloads.black_scholes_step(S, K, T, 3)
workloads.mc_pi(10 ** 5)
G = workloads.jacobi_grid(64)
workloads.jacobi_step(G)
bh.ones((100, 50)).sum(axis=1)
bh.flush()
outs = workloads.random_dag(1001, 200, sides=(128,))
bh.asnumpy_many([o[:1] if o.ndim == 1 else o[:1, :1] for o in outs])
print("DONE")
"""


def test_generated_kernels_have_no_barrier_in_a_divergent_loop(tmp_path):
    cache = tmp_path / "cache"
    env = dict(os.environ, BH_HIP_DRY_RUN="1", BH_HIP_DUMP_DIR=str(tmp_path), XDG_CACHE_HOME=str(cache))
    r = subprocess.run([sys.executable, "-c", DRIVER % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=900)
    assert r.returncode == 0 and "DONE" in r.stdout, r.stdout[-4000:]
    objs = sorted((cache / "bohrium_amd").glob("*.hsaco"))
    assert len(objs) >= 10
    kernels = set()
    for o in objs:
        text = _disassemble(str(o))
        kernels.update(sym for sym, _, _, _ in _instructions(text))
        assert divergent_loops_with_barriers(text) == [], o
    assert {"bh_kernel", "bh_kernel_fin", "bh_kernel_batch", "bh_kernel_persist"} <= kernels
    shutil.rmtree(str(cache), ignore_errors=True)
