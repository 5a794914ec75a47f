"""The learned cost model: the float64 oracle network pinned by finite differences (CPU),
and the HIP training/inference kernels checked against it (MI355X)."""
import os

import numpy as np
import pytest

from oracle import mlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

F, H = 16, 8


def _data(n=40, seed=0):
    rng = np.random.RandomState(seed)
    x = rng.normal(size=(n, F))
    t = np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.2 * x[:, 3]
    p = rng.normal(scale=0.4, size=mlp.layout(F, H)["n"])
    mean = rng.normal(scale=0.1, size=F)
    inv_std = rng.uniform(0.5, 1.5, size=F)
    return x, t, p, mean, inv_std


def test_oracle_gradient_matches_finite_differences():
    x, t, p, mean, inv_std = _data()
    loss, g = mlp.loss_and_gradient(p, x, t, F, H, mean, inv_std)
    fd = mlp.finite_difference_gradient(p, x, t, F, H, mean, inv_std)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)
    assert loss > 0


def test_oracle_plants_caught():
    """A dropped ReLU mask or a transposed W2 changes the gradient (the check above has teeth)."""
    x, t, p, mean, inv_std = _data()
    _, g = mlp.loss_and_gradient(p, x, t, F, H, mean, inv_std)
    W1, b1, W2, b2, w3, b3 = mlp.unpack(p, F, H)
    q = p.copy()
    o = mlp.layout(F, H)
    q[o["W2"]:o["b2"]] = W2.T.ravel()
    _, gq = mlp.loss_and_gradient(q, x, t, F, H, mean, inv_std)
    fdq = mlp.finite_difference_gradient(p, x, t, F, H, mean, inv_std)
    assert not np.allclose(gq, fdq, rtol=1e-3, atol=1e-6)


def test_oracle_adam_reduces_loss():
    x, t, p, mean, inv_std = _data(200)
    _, losses = mlp.adam_steps(p, x, t, F, H, 300, 1e-2, mean, inv_std)
    assert losses[-1] < 0.5 * losses[0]


def test_feature_names_match_engine():
    import re
    from bohrium_amd import costmodel
    src = open(os.path.join(ROOT, "bohrium_amd", "csrc", "ve_hip", "features.cpp")).read()
    block = src[src.index("kFeatureNames[kNumFeatures] = {"):]
    block = block[:block.index("};")]
    assert re.findall(r'"(\w+)"', block) == costmodel.FEATURES


def test_random_dag_is_deterministic():
    from bohrium_amd import workloads, _bh_api as api

    def trace(seed):
        with api.record_only() as r:
            outs = workloads.random_dag(seed, 120)
            del outs
            api.flush()
        return [(e[0], e[1]) if e[0] == "op" else e[0] for e in r.trace]

    assert trace(3) == trace(3)
    assert trace(3) != trace(4)


# ---------------------------------------------------------------- on the GPU

@pytest.mark.gpu
def test_hip_forward_and_gradient_match_oracle():
    from bohrium_amd import costmodel
    x, t, _, mean, inv_std = _data(300, seed=2)
    m = costmodel.CostModel(F, H, seed=5)
    p = m.params().astype(np.float64)
    m.set_normalization(mean, inv_std)
    y, _ = mlp.forward(p, x, F, H, mean, inv_std)
    np.testing.assert_allclose(m.predict(x), y, rtol=1e-4, atol=1e-4)
    loss, g = m.gradient(x, t)
    lo, go = mlp.loss_and_gradient(p, x.astype(np.float32), t.astype(np.float32), F, H,
                                   mean.astype(np.float32), inv_std.astype(np.float32))
    assert abs(loss - lo) <= 1e-4 * max(1.0, lo)
    np.testing.assert_allclose(g, go, rtol=2e-3, atol=2e-5 * np.abs(go).max())


@pytest.mark.gpu
def test_hip_adam_training_tracks_oracle():
    from bohrium_amd import costmodel
    x, t, _, mean, inv_std = _data(256, seed=3)
    m = costmodel.CostModel(F, H, seed=7)
    m.set_normalization(mean, inv_std)
    p0 = m.params().astype(np.float64)
    losses = m.train(x, t, 20, lr=1e-2)
    _, ref = mlp.adam_steps(p0, x.astype(np.float32), t.astype(np.float32), F, H, 20, 1e-2,
                            mean.astype(np.float32), inv_std.astype(np.float32))
    np.testing.assert_allclose(losses, ref, rtol=2e-3, atol=1e-5)
    more = m.train(x, t, 400, lr=1e-2)
    assert more[-1] < 0.5 * losses[0]


@pytest.mark.gpu
def test_learned_fuser_with_the_shipped_model():
    """With the model trained on this device (bohrium_amd/data/costmodel_mi355x.npz, made by
    bohrium_amd.costmodel_train from programs 100-129), the learned fuser's partitions of held-out
    programs run within 10 % of the better of the analytic greedy and serial fusers, and well below
    the serial fuser, whose all-in-one blocks it scores as slower."""
    from bohrium_amd import _bh_api as api
    from bohrium_amd import costmodel
    from bohrium_amd import costmodel_train as ct
    model = costmodel.CostModel.load(ct.DEFAULT_MODEL)
    model.install()
    api.set_flush_threshold(10 ** 7)
    try:
        rows, tot, best = ct.evaluate([1003, 1005, 1006], pipelines=("greedy", "serial", "learned"),
                                      sides=(512, 1024), reps=2)
        print("device ms", tot, "per-program best", best)
        assert tot["learned"] <= 1.10 * min(tot["greedy"], tot["serial"])
        assert tot["learned"] < 0.8 * tot["serial"]
    finally:
        costmodel.uninstall()
        api.set_flush_threshold(1000)
        api.message("fuser_list: " + ct.PIPELINES["greedy"])


@pytest.mark.gpu
def test_launch_model_with_the_shipped_model():
    """The launch-shape model trained on this device (bohrium_amd/data/launchmodel_mi355x.npz, made by
    bohrium_amd.launchmodel_train from programs 100-107) against measurement on held-out programs:
    the autotuner measures every candidate shape of each kernel and reports the model's prediction
    of each, so the model's choices are judged on the same kernels.  Measuring only the configured
    knobs and the model's two best-predicted candidates (autotune_mode: prune) lands within 5 % of
    the measured-best candidates in total, and the model's own pick is at most 3 % slower than the
    configured knobs in total (the totals are sums of device times read from HIP events, which move
    by a few per cent between runs)."""
    from bohrium_amd import _bh_api as api
    from bohrium_amd import costmodel
    from bohrium_amd import launchmodel_train as lt
    model = costmodel.CostModel.load(lt.DEFAULT_MODEL)
    model.install_launch()
    api.set_flush_threshold(10 ** 7)
    try:
        picks, tot_model, tot_best, tot_default = lt.judge([1003, 1005], sides=(1024, 2048), extra=False)
        tot_prune = sum(min(ms[0], min(ms[j] for j in lt.top_k(pred, 2))) for _, _, ms, pred in picks)
        print("kernels %d model %.3f prune-2 %.3f best %.3f configured %.3f"
              % (len(picks), tot_model, tot_prune, tot_best, tot_default))
        assert len(picks) >= 10
        assert tot_prune <= 1.05 * tot_best
        assert tot_model <= 1.03 * tot_default
    finally:
        costmodel.uninstall_launch()
        api.set_flush_threshold(1000)
        api.message("autotune: forget")


@pytest.mark.gpu
def test_model_mode_picks_without_trials_and_keeps_results():
    """autotune_mode = model: kernels take the launch shape the model predicts fastest, no trial runs,
    and the results are those of the oracle on the same bytecode (tolerance of the random-program
    parity tests: rtol 1e-4 plus 1e-4 of the largest output magnitude, for float32 chains of a few
    hundred operations with 1024-term sums reassociated differently on the device)."""
    import bohrium_amd as bh
    from bohrium_amd import _bh_api as api
    from bohrium_amd import costmodel, workloads
    from bohrium_amd import launchmodel_train as lt
    from tests.helpers import run_both
    model = costmodel.CostModel.load(lt.DEFAULT_MODEL)
    model.install_launch()
    api.set_flush_threshold(10 ** 7)
    bh.message("autotune: forget")
    bh.message("autotune_mode: model")
    try:
        t0, m0 = lt._stat("autotune_trials"), lt._stat("autotune_model_picks")
        got, want = run_both(lambda: workloads.random_dag(1004, 120, sides=(1024,)))
        assert lt._stat("autotune_model_picks") > m0
        assert lt._stat("autotune_trials") == t0
        scale = max([1.0] + [float(np.abs(w).max()) for w in want if w.size])
        for g, w in zip(got, want):
            np.testing.assert_allclose(g, w, rtol=1e-4, atol=1e-4 * scale)
    finally:
        bh.message("autotune_mode: measure")
        costmodel.uninstall_launch()
        api.set_flush_threshold(1000)
        bh.message("autotune: forget")


@pytest.mark.gpu
def test_prune_mode_measures_the_best_predicted_shapes_only():
    """autotune_mode = prune: a tuned kernel's trial holds the configured knobs and the model's two
    best-predicted shapes only (the others are counted as pruned), and the results are those of the
    oracle on the same bytecode (tolerance of the random-program parity tests)."""
    import bohrium_amd as bh
    from bohrium_amd import _bh_api as api
    from bohrium_amd import costmodel
    from bohrium_amd import launchmodel_train as lt
    from tests.helpers import run_both
    model = costmodel.CostModel.load(lt.DEFAULT_MODEL)
    model.install_launch()
    api.set_flush_threshold(10 ** 7)
    bh.message("autotune: forget")
    bh.message("autotune_mode: prune")
    try:
        p0, t0 = lt._stat("autotune_pruned"), lt._stat("autotune_trials")
        got, want = run_both(lambda: workloads_dag(1006))
        for _ in range(12):  # let the trials finish
            bh.asnumpy_many([o[:1] if o.ndim == 1 else o[:1, :1] for o in workloads_dag(1006)])
            bh.message("autotune")
        assert lt._stat("autotune_pruned") > p0
        assert lt._stat("autotune_trials") > t0
        for line in bh.message("autotune outcomes").splitlines():
            f = dict(p.split("=", 1) for p in line.split() if "=" in p)
            if "ms" in f:
                assert len(f["ms"].split(",")) <= 3, line
        scale = max([1.0] + [float(np.abs(w).max()) for w in want if w.size])
        for g, w in zip(got, want):
            np.testing.assert_allclose(g, w, rtol=1e-4, atol=1e-4 * scale)
    finally:
        bh.message("autotune_mode: measure")
        costmodel.uninstall_launch()
        api.set_flush_threshold(1000)
        bh.message("autotune: forget")


def workloads_dag(seed):
    from bohrium_amd import workloads
    return workloads.random_dag(seed, 120, sides=(1024,))
