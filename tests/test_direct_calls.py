"""ufunc calls that go straight into the bridge extension (bharray.Ufunc._call_direct).

CPU: the same program run in compile-only mode with the direct path on and off plans the same
launches with the same arguments, constants included (the engine's launch dumps are compared file
by file).  GPU: results of direct calls equal NumPy's (tracing, which the parity tests use, turns
the direct path off, so these calls run untraced here)."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
import sys
import numpy as np
sys.path.insert(0, %(root)r)
import bohrium_amd as bh
from bohrium_amd import bharray
bharray._DIRECT[0] = %(direct)s
taken = [0]
_call_direct = bharray.Ufunc._call_direct
def counting(self, args):
    r = _call_direct(self, args)
    taken[0] += r is not None
    return r
bharray.Ufunc._call_direct = counting
x = bh.array(np.linspace(-2.0, 2.0, 1001, dtype=np.float32))
y = bh.array(np.linspace(1.0, 3.0, 1001, dtype=np.float32))
i = bh.array(np.arange(1001, dtype=np.int64))
d = bh.array(np.linspace(-1.0, 1.0, 1001))
bh.flush()
outs = [
    bh.ufunc("tanh")(x * 0.5 + 0.1),          # array * float, + float, unary
    x * y - y / x,                             # array op array
    x > 0,                                     # comparison with an int constant
    i * 3 + 7, i // 2, i %% 5,                 # int64 with int constants
    d * -0.0, d + 0.0, d * 1e300,              # signed zero and large float constants
    bh.ufunc("absolute")(-d), x * 2,           # float32 array with an int constant
    x.sum(), d.max(), (x > 0).astype(np.int64).cumsum(), i.astype(np.float32),
    bh.take(x, i // 2), bh.ones((30, 40)).sum(axis=1),
]
bh.flush()
print("DIRECT_CALLS=%%d" %% taken[0])
"""


def dumps(tmp_path, direct):
    out = tmp_path / ("direct" if direct else "general")
    out.mkdir()
    env = dict(os.environ, BH_HIP_DRY_RUN="1", BH_HIP_DUMP_DIR=str(out))
    r = subprocess.run([sys.executable, "-c", PROGRAM % {"root": ROOT, "direct": direct}], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:]
    calls = int(r.stdout.split("DIRECT_CALLS=")[1].split()[0])
    assert (calls >= 12) if direct else (calls == 0), calls
    return {f: (out / f).read_text() for f in sorted(os.listdir(out))}


def test_direct_calls_plan_the_same_launches(tmp_path):
    general = dumps(tmp_path, False)
    direct = dumps(tmp_path, True)
    assert general, "no launches were dumped"
    assert sorted(general) == sorted(direct)
    for name in general:
        assert general[name] == direct[name], name


@pytest.mark.gpu
def test_direct_calls_match_numpy():
    import bohrium_amd as bh
    from bohrium_amd import bharray
    assert bharray._DIRECT[0]
    xs = np.linspace(-2.0, 2.0, 100003, dtype=np.float32)
    ys = np.linspace(1.0, 3.0, 100003, dtype=np.float32)
    ii = np.arange(100003, dtype=np.int64)
    x, y, i = bh.array(xs), bh.array(ys), bh.array(ii)
    got = [(x * 0.5 + 0.1).asnumpy(), (x * y - y).asnumpy(), (x > 0).asnumpy(), (i * 3 + 7).asnumpy(),
           (x * -0.0).asnumpy(), bh.ufunc("tanh")(x).asnumpy()]
    want = [xs * np.float32(0.5) + np.float32(0.1), xs * ys - ys, xs > 0, ii * 3 + 7, xs * np.float32(-0.0),
            np.tanh(xs)]
    for g, w in zip(got, want):
        assert g.dtype == w.dtype
    # the kernels contract a*b + c into one fma (NumPy rounds twice): float32 within 1e-6 relative
    np.testing.assert_allclose(got[0], want[0], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got[1], want[1], rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(got[2], want[2])
    np.testing.assert_array_equal(got[3], want[3])
    np.testing.assert_array_equal(got[4], want[4])
    assert np.array_equal(np.signbit(got[4]), np.signbit(want[4]))
    np.testing.assert_allclose(got[5], want[5], rtol=2e-6, atol=1e-7)
