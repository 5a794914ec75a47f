"""Recorded flushes in the BhIR wire format (bhc_record_flushes / bhc_replay_file).

CPU: the committed fixtures (tests/golden/bhir, written by tests/golden/make_bhir_fixtures.py)
hold exactly the bytecode their programs emit today (compared instruction by instruction with the
bridge trace of the same program, up to base renaming), a recording replays through a
compile-only engine, and damaged files are refused.  GPU: each fixture replayed on the hip engine
gives, for every result base, what the oracle computes from the same file (bit-exact for integers,
1e-12 relative for float64 sums of up to a few thousand terms, 1e-5 for float32)."""
import importlib.util
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "bhir")
sys.path.insert(0, ROOT)

from oracle import bhir_wire  # noqa: E402

_spec = importlib.util.spec_from_file_location("make_bhir_fixtures",
                                               os.path.join(ROOT, "tests", "golden", "make_bhir_fixtures.py"))
fixtures = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(fixtures)
NAMES = sorted(fixtures.PROGRAMS)


def _canonical(trace):
    """Ops and flushes with bases renamed in order of first use (trace ids and addresses differ)."""
    ren = {}

    def r(b):
        return ren.setdefault(b, len(ren))

    out = []
    for rec in trace:
        if rec[0] == "op":
            ops = [None if o is None else (r(o[0]),) + tuple(o[1:]) for o in rec[2]]
            slides = rec[4] if len(rec) > 4 else None
            if slides is not None and not any(slides):
                slides = None
            out.append(("op", rec[1], ops, rec[3], slides))
        elif rec[0] == "free":
            out.append(("free", r(rec[1])))
        elif rec[0] == "flush" and (rec[1] != 1 or rec[2] is not None):
            # repeated flushes (do_while); plain flush boundaries depend on the auto-flush threshold
            out.append(("flush", rec[1], None if rec[2] is None else r(rec[2])))
    return out


@pytest.mark.parametrize("name", NAMES)
def test_fixture_holds_the_programs_bytecode(name):
    """The file's instructions equal the bridge trace of the same program (record-only mode)."""
    from bohrium_amd import _bh_api as api
    with api.record_only() as rec:
        res = fixtures.PROGRAMS[name]()
        api.flush()
        del res
    want = _canonical(rec.trace)
    trace, live = bhir_wire.to_trace(bhir_wire.read_file(os.path.join(GOLDEN, name + ".bhir")))
    got = _canonical(trace)
    frees = lambda t: [x for x in t if x[0] != "free"]
    # frees of arrays dropped after the last flush are not in the file
    assert frees(got) == frees(want)[:len(frees(got))]
    assert len(frees(got)) >= len(frees(want)) - 1
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    for k, v in meta["results"].items():
        assert v["address"] in live, k


DRY = r"""
import sys
sys.path.insert(0, %(root)r)
import numpy as np
import bohrium_amd as bh
from bohrium_amd import _bh_api as api, workloads
def flushes():
    for l in bh.message("statistic").splitlines():
        if l.startswith("flushes:"):
            return int(l.split(":")[1])
api.record_flushes(%(path)r)
s = workloads.black_scholes(1000, steps=2)
g = workloads.jacobi_grid(32)
d = workloads.jacobi_repeat(g, 2)
bh.flush()
api.record_flushes(None)
f0 = flushes()
api.replay_file(%(path)r)
print("replayed", flushes() - f0)
h = api.replay_base(api.base_address(d.base.handle))
print("base", api.lib().bhc_base_nelem(h))
try:
    api.replay_file(%(bad)r)
except api.BohriumError as e:
    print("refused", str(e).splitlines()[0])
"""


def test_replay_in_a_compile_only_engine(tmp_path):
    path = str(tmp_path / "rec.bhir")
    bad = str(tmp_path / "bad.bhir")
    env = dict(os.environ, BH_HIP_DRY_RUN="1")
    # record, then cut the file inside its second archive
    src = DRY % {"root": ROOT, "path": path, "bad": bad}
    with open(bad, "wb") as f:
        f.write(b"")
    r = subprocess.run([sys.executable, "-c", src.replace("api.replay_file(%r)" % bad, "pass")], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    n = len(bhir_wire.read_file(path))
    raw = open(path, "rb").read()
    with open(bad, "wb") as f:
        f.write(raw[:len(raw) // 2])
    r = subprocess.run([sys.executable, "-c", src], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    lines = dict(l.split(" ", 1) for l in r.stdout.splitlines() if " " in l)
    assert int(lines["replayed"]) >= n, r.stdout
    assert int(lines["base"]) == 1
    assert "truncated" in lines["refused"], r.stdout


def test_damaged_archives_are_refused(tmp_path):
    """Wrong magic, wrong version and trailing bytes raise instead of replaying garbage."""
    src = open(os.path.join(GOLDEN, "mc_pi_1e5.bhir"), "rb").read()
    cases = {"magic": b"XHIRFILE" + src[8:], "version": src[:8] + b"\x02\x00\x00\x00" + src[12:]}
    # an archive one byte longer than it should be
    n = int.from_bytes(src[12:20], "little")
    cases["trailing"] = src[:12] + (n + 1).to_bytes(8, "little") + src[20:20 + n] + b"\x00" + src[20 + n:]
    paths = {}
    for what, data in cases.items():
        p = tmp_path / (what + ".bhir")
        p.write_bytes(data)
        paths[what] = str(p)
    prog = ("import sys\nsys.path.insert(0, %r)\nfrom bohrium_amd import _bh_api as api\n"
            "for what, p in %r.items():\n"
            "    try:\n        api.replay_file(p)\n        print(what, 'accepted')\n"
            "    except api.BohriumError as e:\n        print(what, str(e).splitlines()[0])\n") % (ROOT, paths)
    r = subprocess.run([sys.executable, "-c", prog], env=dict(os.environ, BH_HIP_DRY_RUN="1"),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    out = dict(l.split(" ", 1) for l in r.stdout.splitlines() if " " in l)
    assert "not a recorded-flush file" in out["magic"], out
    assert "not a recorded-flush file" in out["version"], out
    assert "trailing bytes" in out["trailing"], out


def _tol(dtype):
    dtype = np.dtype(dtype)
    if dtype.kind in "biu":
        return None
    return 1e-12 if dtype == np.float64 else 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_fixture_replay_matches_the_oracle(name):
    import bohrium_amd as bh  # noqa: F401
    from bohrium_amd import _bh_api as api
    from oracle import omp_ve
    path = os.path.join(GOLDEN, name + ".bhir")
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))["results"]
    trace, live = bhir_wire.to_trace(bhir_wire.read_file(path))
    data = omp_ve.replay(trace)
    api.replay_file(path)
    handles = {k: api.replay_base(v["address"]) for k, v in meta.items()}
    for h in handles.values():
        api.sync(h)
    api.flush()
    for k, v in sorted(meta.items()):
        want = data[live[v["address"]]]
        got = api.data_copy_out(handles[k], np.dtype(v["dtype"]), want.size)
        tol = _tol(v["dtype"])
        if tol is None:
            np.testing.assert_array_equal(got, want, err_msg=k)
        else:
            np.testing.assert_allclose(got, want, rtol=tol, atol=tol, err_msg=k)
