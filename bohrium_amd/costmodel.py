"""The learned kernel cost model (include/bh_costmodel.h) from Python.

    CostModel          ctypes binding to the bhcm_* HIP implementation in libbh_ve_hip.so
    profiled_kernels   (features, device microseconds) of the kernels the engine just ran
    fit                standardise features, train on the GPU, return the model
    install            make a model the one the engine's learned fuser uses

The engine records the cost-model features of every jitk kernel with its profiled
launches ("launches" message: `... k=<kernel id> f=<features>`); a kernel's time is
the sum of its launches.  The model predicts log(microseconds).
"""
import ctypes
import os

import numpy as np

from . import _bh_api

FEATURES = ["log2_elements", "log2_bytes_in", "log2_bytes_out", "log2_bytes_temp", "log2_instrs", "transcendentals",
            "reductions", "accumulates", "gathers_scatters", "max_ndim", "strided_inputs", "inner_axis_reduce",
            "outer_axis_reduce", "full_reduce", "input_arrays", "fp64_fraction"]
NFEAT = len(FEATURES)
# The launch-shape model reads the block features followed by these (include/bh_costmodel.h).
LAUNCH_FEATURES = ["log2_workgroups", "log2_issued_together", "log2_elements_per_thread"]
NFEAT_LAUNCH = NFEAT + len(LAUNCH_FEATURES)

_LIB = None


def lib():
    """The engine library (the same loaded instance the component stack uses)."""
    global _LIB
    if _LIB is None:
        _bh_api.lib()  # make sure the stack (and with it the engine library) is loaded
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libbh_ve_hip.so")
        L = ctypes.CDLL(path)
        vp, i64, f32p = ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_float)
        L.bhcm_last_error.restype = ctypes.c_char_p
        L.bhcm_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]
        L.bhcm_destroy.argtypes = [vp]
        L.bhcm_num_params.argtypes = [vp]
        L.bhcm_num_params.restype = i64
        L.bhcm_get_params.argtypes = [vp, f32p]
        L.bhcm_set_params.argtypes = [vp, f32p]
        L.bhcm_set_normalization.argtypes = [vp, f32p, f32p]
        L.bhcm_predict.argtypes = [vp, f32p, i64, f32p]
        L.bhcm_gradient.argtypes = [vp, f32p, f32p, i64, f32p, f32p]
        L.bhcm_train.argtypes = [vp, f32p, f32p, i64, ctypes.c_int, ctypes.c_float, f32p]
        L.bhcm_install.argtypes = [vp]
        L.bhcm_install_launch.argtypes = [vp]
        _LIB = L
    return _LIB


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(rc):
    if rc != 0:
        raise RuntimeError("cost model: " + lib().bhcm_last_error().decode())


class CostModel:
    """MLP nfeat -> hidden -> hidden -> 1 (ReLU) on the GPU."""

    def __init__(self, nfeat=NFEAT, hidden=32, seed=1):
        self.nfeat, self.hidden = nfeat, hidden
        h = ctypes.c_void_p()
        _check(lib().bhcm_create(nfeat, hidden, seed, ctypes.byref(h)))
        self.handle = h
        self.mean = np.zeros(nfeat, np.float32)
        self.inv_std = np.ones(nfeat, np.float32)

    def __del__(self):
        if getattr(self, "handle", None) is not None and _LIB is not None:
            _LIB.bhcm_destroy(self.handle)
            self.handle = None

    def num_params(self):
        return int(lib().bhcm_num_params(self.handle))

    def params(self):
        out = np.empty(self.num_params(), np.float32)
        _check(lib().bhcm_get_params(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, p):
        p, ptr = _f32(p)
        assert p.size == self.num_params()
        _check(lib().bhcm_set_params(self.handle, ptr))

    def set_normalization(self, mean, inv_std):
        m, mp = _f32(mean)
        s, sp = _f32(inv_std)
        _check(lib().bhcm_set_normalization(self.handle, mp, sp))
        self.mean, self.inv_std = m.copy(), s.copy()

    def predict(self, x):
        x, xp = _f32(x)
        n = x.shape[0]
        y = np.empty(n, np.float32)
        _check(lib().bhcm_predict(self.handle, xp, n, y.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return y

    def gradient(self, x, t):
        x, xp = _f32(x)
        t, tp = _f32(t)
        g = np.empty(self.num_params(), np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().bhcm_gradient(self.handle, xp, tp, x.shape[0], g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   loss.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return float(loss[0]), g

    def train(self, x, t, steps, lr=1e-2):
        x, xp = _f32(x)
        t, tp = _f32(t)
        losses = np.zeros(steps, np.float32)
        _check(lib().bhcm_train(self.handle, xp, tp, x.shape[0], steps, lr,
                                losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return losses

    def install(self):
        """Use this model in the engine's learned fuser (cached plans are dropped)."""
        _check(lib().bhcm_install(self.handle))
        _bh_api.message("plans: drop")

    def install_launch(self):
        """Use this model (NFEAT_LAUNCH features) as the autotuner's launch-shape model."""
        assert self.nfeat == NFEAT_LAUNCH
        _check(lib().bhcm_install_launch(self.handle))
        _bh_api.message("plans: drop")

    def save(self, path):
        np.savez(path, params=self.params(), mean=self.mean, inv_std=self.inv_std,
                 shape=np.array([self.nfeat, self.hidden]))

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        m = cls(int(d["shape"][0]), int(d["shape"][1]))
        m.set_params(d["params"])
        m.set_normalization(d["mean"], d["inv_std"])
        return m


def uninstall():
    _check(lib().bhcm_install(None))
    _bh_api.message("plans: drop")


def uninstall_launch():
    _check(lib().bhcm_install_launch(None))
    _bh_api.message("plans: drop")


def tuning_samples(text):
    """(features n x NFEAT_LAUNCH, microseconds n) from the engine's "autotune samples" text."""
    X, y = [], []
    for line in text.splitlines():
        parts = line.split()
        if len(parts) != 2 or not parts[0].startswith("f=") or not parts[1].startswith("ms="):
            continue
        X.append([float(v) for v in parts[0][2:].split(",")])
        y.append(float(parts[1][3:]) * 1e3)
    if not X:
        return np.zeros((0, NFEAT_LAUNCH), np.float32), np.zeros(0, np.float32)
    return np.array(X, np.float32), np.array(y, np.float32)


def tuning_outcomes(text):
    """[(signature, chosen, measured ms per candidate, predicted ms per candidate or None)] from the
    engine's "autotune outcomes" text (candidate 0: the configured knobs)."""
    out = []
    for line in text.splitlines():
        f = dict(p.split("=", 1) for p in line.split() if "=" in p)
        if "ms" not in f:
            continue
        pred = None if f["predicted_ms"] == "none" else [float(v) for v in f["predicted_ms"].split(",")]
        out.append((f["key"], int(f["chosen"]), [float(v) for v in f["ms"].split(",")], pred))
    return out


def profiled_kernels(launch_text):
    """(features n x NFEAT, microseconds n) per jitk kernel from the engine's "launches" text."""
    per = {}
    for line in launch_text.splitlines():
        parts = line.split()
        if len(parts) < 7 or not parts[5].startswith("k=") or not parts[6].startswith("f="):
            continue
        kid, feats = parts[5][2:], parts[6][2:]
        if not feats:
            continue
        us = float(parts[3]) * 1e3
        if kid in per:
            per[kid][1] += us
        else:
            per[kid] = [np.array([float(v) for v in feats.split(",")], np.float32), us]
    if not per:
        return np.zeros((0, NFEAT), np.float32), np.zeros(0, np.float32)
    X = np.stack([v[0] for v in per.values()])
    y = np.array([v[1] for v in per.values()], np.float32)
    return X, y


def fit(X, us, hidden=32, steps=3000, lr=3e-3, seed=1):
    """Train a model on (features, microseconds); targets are log(microseconds)."""
    X = np.asarray(X, np.float32)
    t = np.log(np.maximum(np.asarray(us, np.float64), 1e-3)).astype(np.float32)
    m = CostModel(X.shape[1], hidden, seed)
    mean = X.mean(0)
    std = X.std(0)
    m.set_normalization(mean, 1.0 / np.where(std > 1e-6, std, 1.0))
    losses = m.train(X, t, steps, lr)
    return m, losses
