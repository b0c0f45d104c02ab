"""Learned sweep-axis selector: kernels against the numpy oracle
(oracle/mlp_ref.py), the judges (plain sweep: in-LDS; slab sweep: the slab
sweep's own per-scene counts) against the pipeline, numpy counts and brute
force, and the gain of the trained networks over PCA and the closed-form
automatic rule with identical pair sets."""
import os

import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import learn as L
from gsap_amd import workloads as W
from oracle import mlp_ref as M
from oracle import oracle as O
from tests.gpu_util import to_dev
from tests.test_axis_auto import slab_windows

pytestmark = pytest.mark.gpu


def _data(rng, B):
    x = rng.standard_normal((B, 64)).astype(np.float32) * 0.5
    t = np.abs(rng.standard_normal((B, 65))).astype(np.float32)
    return x, t


def test_mlp_forward_matches_oracle(dev):
    rng = np.random.default_rng(0)
    m = G.AxisModel(seed=3)
    p = m.params()
    x, _ = _data(rng, 256)
    y = m.forward(to_dev(x, dev)).cpu().numpy()
    yr, _ = M.forward(p, x)
    # bf16 activations: a sum landing on a rounding boundary may round the
    # other way (2^-8 relative), so compare with a bf16-scale tolerance
    assert np.allclose(y, yr, rtol=2e-2, atol=2e-2 * np.abs(yr).max())
    assert np.median(np.abs(y - yr)) < 1e-3 * np.abs(yr).max()


def test_mlp_gradient_matches_oracle(dev):
    rng = np.random.default_rng(1)
    m = G.AxisModel(seed=5)
    p = m.params()
    x, t = _data(rng, 512)
    loss, g = m.loss_grad(to_dev(x, dev), to_dev(t, dev), lambda_count=0.05)
    lr, gr, _ = M.loss_grad(p, x, t, G.axis_candidates(), lam=0.05)
    assert abs(loss - lr) <= 1e-3 * max(1.0, abs(lr))
    scale = np.abs(gr).max()
    assert np.allclose(g, gr, rtol=5e-2, atol=5e-3 * scale)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.999


def test_judge_matches_pipeline_and_brute_force(dev):
    boxes, off = W.scenes(6, seed=11, lo=300, hi=2000)
    bd = to_dev(boxes, dev)
    axes = G.axis_candidates()[[0, 1, 2, 5, 9, 40]]
    sap = G.Sap()
    tests, pairs = sap.axis_tests_batched(bd, off, axes)
    tests, pairs = tests.cpu().numpy(), pairs.cpu().numpy()
    for s in range(len(off) - 1):
        b = boxes[off[s]:off[s + 1]]
        assert pairs[s] == len(O.brute(b))
        for k, u in enumerate(axes):
            assert tests[s, k] == O.sweep_tests(b, u)
            fixed = G.Sap(axis_mode=G.AXIS_FIXED, fixed_axis=u)
            _, st = fixed.count(to_dev(b, dev))
            assert st.sweep_tests == tests[s, k]


def test_features_finite_deterministic_scale_free(dev):
    boxes, off = W.scenes(32, seed=12, lo=500, hi=5000)
    sap = G.Sap()
    f1 = sap.features_batched(to_dev(boxes, dev), off).cpu().numpy()
    f2 = sap.features_batched(to_dev(boxes, dev), off).cpu().numpy()
    assert np.all(np.isfinite(f1)) and np.array_equal(f1, f2)
    f3 = sap.features_batched(to_dev(boxes * 4.0, dev), off).cpu().numpy()
    assert np.allclose(f1, f3, atol=2e-3)  # positions and extents scaled together


def test_training_reduces_loss(dev):
    import torch
    boxes, off = W.scenes(3000, seed=21, lo=128, hi=1024)
    bd = to_dev(boxes, dev)
    sap = G.Sap()
    feats, tgt, tests, best = L.make_dataset(sap, bd, off, G.axis_candidates())
    assert feats.shape == (3000, 64) and tgt.shape == (3000, 65)
    m = G.AxisModel(seed=1)
    losses = m.train(feats, tgt, epochs=15, batch=500, lr=2e-3)
    assert losses[-1] < 0.6 * losses[0], losses
    # the snapped prediction's regret is far below that of a random candidate
    y = m.forward(feats).cpu().numpy()
    cand = G.axis_candidates()
    pick = np.argmax(np.abs(y[:, :3] @ cand.T), axis=1)
    t = tests.cpu().numpy().astype(np.float64)
    regret_all = t / t.min(1, keepdims=True) - 1
    regret = regret_all[np.arange(len(t)), pick]
    assert np.mean(regret) < 0.5 * np.mean(regret_all), (np.mean(regret), np.mean(regret_all))


def _model(dev, slabs=False):
    path = L.default_model(slabs)
    if os.path.exists(path):
        return G.AxisModel.load(path)
    if slabs:
        model, _ = L.train_selector(n_scenes=8192, lo=256, hi=65536, epochs=40, slabs=True,
                                    regret_scale=20.0, log=lambda *_: None)
    else:
        model, _ = L.train_selector(n_scenes=20000, epochs=30, regret_scale=4.0, log=lambda *_: None)
    return model


def _plane_pair(b, k):
    """The sweep and slab axes of the plane normal to coordinate axis k: the
    two other coordinate axes, the one of larger centre variance first."""
    c = (b[:, :3].astype(np.float64) + b[:, 3:].astype(np.float64)) * 0.5
    var = c.var(0)
    a0, a1 = [j for j in range(3) if j != k]
    u, v = (a0, a1) if var[a0] >= var[a1] else (a1, a0)
    return np.eye(3)[u], np.eye(3)[v]


def test_slab_judge_counts_the_slab_sweep(dev):
    """Slab-mode labels for scenes beyond the in-LDS judge's 2,048 objects:
    each coordinate-plane count matches a numpy count of that plane's slab
    sweep windows without their own entries (up to the ties of the 16-bit key
    quantisation); pairs match brute force."""
    boxes, off = W.scenes(5, seed=13, lo=2500, hi=12000)
    bd = to_dev(boxes, dev)
    axes = G.axis_candidates()[:3]
    sap = G.Sap(flags=G.FLAG_SLABS)
    tests, pairs = sap.axis_tests_batched(bd, off, axes)
    tests, pairs = tests.cpu().numpy(), pairs.cpu().numpy()
    for s in range(len(off) - 1):
        b = boxes[off[s]:off[s + 1]]
        assert pairs[s] == len(O.brute(b))
        for k in range(3):
            t, e = slab_windows(b, *_plane_pair(b, k))
            t -= e  # the sweep does not test an entry against itself
            assert abs(int(tests[s, k]) - t) <= 0.01 * t + 64, (s, k, int(tests[s, k]), t)
    # a second call on the same handle gives the same labels
    t2, p2 = sap.axis_tests_batched(bd, off, axes)
    assert np.array_equal(t2.cpu().numpy(), tests) and np.array_equal(p2.cpu().numpy(), pairs)


def test_learned_axis_cuts_sweep_tests_identical_pairs(dev):
    """The trained selector on 256 scenes of the 2,048-scene workload's
    generator (sizes log-uniform in [1024, 65536]): >= 20% fewer sweep tests
    than PCA, identical pair lists and per-scene offsets."""
    model = _model(dev)
    boxes, off = W.scenes(256, seed=5, lo=1024, hi=65536)
    ev = L.evaluate(model, to_dev(boxes, dev), off, conf_min=0.95)
    print(ev)
    assert ev["pairs_identical"]
    assert ev["reduction"] >= 0.20, ev


def test_learned_planes_beat_auto_identical_pairs(dev):
    """Slab sweep: the trained plane selector on 256 scenes of the 2,048-scene
    workload's generator runs fewer sweep tests than the closed-form automatic
    rule and less than half of PCA's, with pair lists and per-scene offsets
    identical to PCA's."""
    model = _model(dev, slabs=True)
    boxes, off = W.scenes(256, seed=5, lo=1024, hi=65536)
    ev = L.evaluate(model, to_dev(boxes, dev), off, conf_min=0.95, slabs=True)
    print(ev)
    assert ev["pairs_identical"]
    assert ev["tests_learned"] < ev["tests_auto"], ev
    assert ev["tests_learned"] < 0.5 * ev["tests_pca"], ev


def test_learned_planes_single_scene_same_pairs(dev):
    model = _model(dev, slabs=True)
    b = W.aniso(200_000, seed=3)
    sl = G.Sap(axis_mode=G.AXIS_LEARNED, flags=G.FLAG_SLABS)
    sl.set_learned_model(model)
    got, _, st = sl.collide(to_dev(b, dev))
    ref, _, _ = G.Sap(axis_mode=G.AXIS_PCA).collide(to_dev(b, dev))
    import torch
    assert torch.equal(got, ref)
    assert st.axis_source == 3
    assert sorted(np.abs(np.asarray(st.axis))) == [0.0, 0.0, 1.0]  # a coordinate sweep axis


def test_learned_mode_single_scene_same_pairs(dev):
    model = _model(dev)
    b = W.aniso(200_000, seed=3)
    sl = G.Sap(axis_mode=G.AXIS_LEARNED)
    sl.set_learned_model(model)
    got, _, st = sl.collide(to_dev(b, dev))
    ref, _, stp = G.Sap(axis_mode=G.AXIS_COORD_MAXVAR).collide(to_dev(b, dev))
    import torch
    assert torch.equal(got, ref)
    assert st.axis_source in (0, 3)


def test_learned_mode_requires_model(dev):
    s = G.Sap(axis_mode=G.AXIS_LEARNED)
    with pytest.raises(G.GsapError) as e:
        s.count(to_dev(W.small(100), dev))
    assert e.value.code == G.ERR_UNSUPPORTED
