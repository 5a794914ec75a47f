"""The bytecode trace that the oracle replays (bohrium_amd._bh_api trace_start / trace_stop).

Base handles are the runtime's object addresses, which the allocator hands out again once a base is
freed; the trace must still give every base its own id, or the replay would make two bases share one
array (an intermittent parity failure that depends on allocation history)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, %(root)r)
import numpy as np
import bohrium_amd as bh
from bohrium_amd import _bh_api as api
handles = []
new = api.new
def spy(*a, **k):
    h = new(*a, **k)
    handles.append(h)
    return h
api.new = spy
api.trace_start()
for i in range(300):
    a = bh.ones(10 + i %% 3)
    b = a + 1.0
    del a, b
    bh.flush()
t = api.trace_stop()
ids = [r[1] for r in t if r[0] == "new"]
print("handles", len(handles), "distinct", len(set(handles)))
print("ids", len(ids), "distinct", len(set(ids)))
"""


def test_trace_ids_stay_unique_when_handles_are_reused():
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    r = subprocess.run([sys.executable, "-c", SCRIPT % {"root": ROOT}], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    lines = dict(l.split(" ", 1) for l in r.stdout.splitlines() if l.startswith(("handles", "ids")))
    nh, dh = (int(v) for v in lines["handles"].split()[::2])
    ni, di = (int(v) for v in lines["ids"].split()[::2])
    assert dh < nh, "the allocator reused no handle; the check below would prove nothing"
    assert ni == nh and di == ni
