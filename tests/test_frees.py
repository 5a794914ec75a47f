"""Every BH_FREE of a flush must reach the engine's release of its base.

The bridge deletes a freed base after the flush and the allocator may hand its address to a
new base; a device buffer still filed under that address would then be used for the new base,
whatever its size.  A transformer that drops a free (push_reductions_inwards once lost the frees
of the sub-blocks it swapped) shows up as `missed_frees` in the engine's statistics.  Runs the
random array programs in compile-only mode (BH_HIP_DRY_RUN=1, nothing launched)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
import sys
sys.path.insert(0, %(root)r)
import bohrium_amd as bh
from bohrium_amd import workloads, _bh_api as api
api.set_flush_threshold(10 ** 7)
for seed, n, f in %(cases)r:
    if True:
        bh.message("fuser_list: " + f + ", push_reductions_inwards, collapse_redundant_axes")
        outs = workloads.random_dag(seed, n, sides=(64, 128))
        bh.asnumpy_many([o[:1] if o.ndim == 1 else o[:1, :1] for o in outs])
        del outs
        miss = [l for l in bh.message("statistic").splitlines() if l.startswith("missed_frees:")]
        print("CASE", seed, n, repr(f), miss[0], flush=True)
"""


def test_random_programs_release_every_freed_base():
    # seed 104 (398 instructions) under the serial fuser lost 16 frees in push_reductions_inwards
    cases = [(104, 398, "serial"), (104, 398, "greedy"), (107, 159, "serial")]
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    r = subprocess.run([sys.executable, "-c", DRIVER % {"root": ROOT, "cases": cases}], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("CASE")]
    assert len(lines) == len(cases), r.stdout[-4000:]
    for l in lines:
        assert l.endswith("missed_frees: 0"), l
