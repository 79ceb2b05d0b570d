"""End-to-end closure at small size: training pairs from the solver's own
fine runs, training with the library's kernels, and a held-out run where
coarse+closure beats the plain coarse run against the fine solver; the
fallback fires when the tolerance is made impossible to meet."""
import numpy as np
import pytest

from nemo2d import closure_pipeline as cp

pytestmark = pytest.mark.gpu

NC = 48
# gentle bathymetry for a 96 km basin (heavily overlapping mounts make the
# shallow-water run itself blow up): 6 mounts, radii 4-16 km, heights 10-40 m
KW = dict(n_mounts=6, radius=(4e3, 16e3), height=(10.0, 40.0))


def train_small():
    Xs, Ys = [], []
    for seed in range(3):
        X, Y = cp.generate_pairs(cp.Setup(NC, seed, **KW), n_snap=24)
        Xs.append(X)
        Ys.append(Y)
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    clos = cp.Closure(X, Y, seed=0)
    losses = clos.train(X, Y, iters=600, batch=8, lr=2e-3)
    return clos, losses


def test_closure_trains_and_helps_on_held_out_basin():
    clos, losses = train_small()
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.6 * losses[0], losses
    held = cp.Setup(NC, 100, **KW)
    ep, ec, fell = cp.evaluate(held, clos, n_windows=24, tol=2.0)
    print("plain ssh rms", ep.mean(), "closure ssh rms", ec.mean())
    assert not fell
    assert ec.mean() < ep.mean(), (ec.mean(), ep.mean())
    # impossible tolerance: the judge switches the closure off
    _, _, fell2 = cp.evaluate(held, clos, n_windows=10, tol=1e-9, judge_every=5)
    assert fell2


def test_device_correction_matches_host_correction_bitwise():
    """Closure.correct_model (features, network and increments on the
    device) leaves the same state as the host path Closure.correct."""
    setup = cp.Setup(24, 5, **KW)
    X, Y = cp.generate_pairs(setup, n_snap=3)
    clos = cp.Closure(X, Y, seed=1)
    clos.train(X, Y, iters=5, batch=2)
    m = setup.coarse_model()
    m.run(1, 30)
    host, _ = clos.correct(m.retrieve(), setup.coarse_grid)
    clos.correct_model(m)
    dev = m.retrieve()
    for k in ("sshn", "un", "vn", "sshn_u", "sshn_v"):
        assert np.array_equal(dev[k], host[k]), (k, np.abs(dev[k] - host[k]).max())
    # the corrected model steps on from the corrected state
    m.run(31, 40)
    assert np.all(np.isfinite(m.retrieve()["sshn"]))
    m.close()


def test_device_restriction_and_target_equal_host_path_bitwise():
    """The on-device restriction of a fine state and the training target
    equal the numpy restriction (restrict_state) and target bit for bit."""
    from nemo2d.closure import DevArray
    setup = cp.Setup(24, 7, **KW)
    fm, cm, tm = setup.fine_model(), setup.coarse_model(), setup.coarse_model()
    fm.run(1, 60)
    f = fm.retrieve()
    host = cp.restrict_state(f, setup.fine_grid, setup.coarse_grid)
    cm.restrict_from(fm)
    dev = cm.retrieve()
    for k in ("sshn", "un", "vn", "sshn_u", "sshn_v"):
        assert np.array_equal(dev[k], host[k]), (k, np.abs(dev[k] - host[k]).max())
    # the restricted model steps on like one given the same state by set_state
    ref = setup.coarse_model()
    ref.set_state(host)
    cm.run(31, 40)
    ref.run(31, 40)
    a, b = cm.retrieve(), ref.retrieve()
    for k in ("sshn", "un", "vn"):
        assert np.array_equal(a[k], b[k]), k
    # target: (truth - pred) / scale per cell, FP32
    fm.run(61, 80)
    tm.restrict_from(fm)
    truth = cp.restrict_state(fm.retrieve(), setup.fine_grid, setup.coarse_grid)
    scale = np.array([1e-3, 2e-4, 3e-4])
    y = DevArray((24 * 24, 8), np.float32)
    tm.closure_target(cm, scale, y)
    want = ((cp.t_fields(truth) - cp.t_fields(a)) / scale).astype(np.float32).reshape(-1, 3)
    got = y.download()
    assert np.array_equal(got[:, :3], want) and not got[:, 3:].any()
    for m in (fm, cm, tm, ref):
        m.close()


def test_online_training_is_deterministic():
    """Two trainings on the device from the same seeds give bitwise the same
    weights: the pairs are written on the models' streams into the batch the
    network reads on the default stream, and each rewrite waits for the
    network's previous pass (a rewrite racing the pass it feeds would make
    the result depend on timing)."""
    bathy = dict(n_mounts=8, radius=(20e3, 100e3), height=(10.0, 60.0))
    runs = []
    for _ in range(2):
        clos, info = cp.train_online(512, range(4), 8, group=2, bathy=bathy, lr=1e-3, rollout=2)
        assert info["pairs"] == 32
        runs.append(clos.net.weights())
    for a, b in zip(runs[0][0] + runs[0][1], runs[1][0] + runs[1][1]):
        assert np.array_equal(a, b)


def test_closure_at_1024_beats_plain_coarse_and_fallback_fires():
    """The closure at the benchmark's size: 1024^2 coarse basins, pairs from
    2048^2 fine runs made and used on the device (16 basins x 40 windows),
    judged on a held-out basin by the fine run: over 20 windows (200 coarse
    steps) the coarse+closure run's ssh is closer to the fine solution than
    the plain coarse run's; an impossible tolerance switches the closure off."""
    bathy = dict(n_mounts=32, radius=(20e3, 200e3), height=(10.0, 60.0))
    clos, info = cp.train_online(1024, range(16), 40, group=4, bathy=bathy, lr=1e-3, rollout=2)
    assert info["pairs"] == 640 and all(np.isfinite(l) for _, l in info["losses"])
    held = cp.Setup(1024, 1000, **bathy)
    ep, ec, fell = cp.evaluate_device(held, clos, 20, tol=1.0, judge_every=1)
    print("plain ssh rms", ep.mean(), "closure ssh rms", ec.mean(), "fell back", fell)
    assert ec.mean() < 0.9 * ep.mean(), (ec.mean(), ep.mean())
    _, _, fell2 = cp.evaluate_device(held, clos, 5, tol=1e-9, judge_every=1)
    assert fell2
