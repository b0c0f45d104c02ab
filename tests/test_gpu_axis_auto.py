"""GSAP_AXIS_AUTO on the GPU: the axis (pair) chosen by k_axis's sweep-test
model equals the oracle's or_axis_auto on the same boxes, the slab sweep's
test count matches a numpy count over the oracle's pair, and the pair lists
are those of brute force / the PCA mode (the axis never changes the pairs)."""
import itertools

import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import pairs_np, to_dev
from tests.test_axis_auto import slab_tests

pytestmark = pytest.mark.gpu


def run(boxes, dev, mode, slabs, kind=G.INPUT_AABB):
    s = G.Sap(input_kind=kind, axis_mode=mode, flags=G.FLAG_SLABS if slabs else 0)
    pairs, _, st = s.collide(to_dev(boxes, dev))
    out = pairs_np(pairs).copy()
    s.close()
    return out, st


def separated(b, slabs, kind=O.KIND_AABB):
    """True when the oracle's best candidate beats the runner-up by > 1e-6
    (so fp64 summation order cannot flip the choice)."""
    mom, (xs, xm) = O.moments(b, kind), O.extents(b, kind)
    n = len(b)
    pca, _ = O.pca_axis(mom, n)
    I = np.eye(3, dtype=np.float32)
    if slabs:
        c = [O.axis_cost(mom, xs, xm, n, pca, O.slab_axis(mom, n, pca))]
        c += [O.axis_cost(mom, xs, xm, n, I[a], I[k]) for a, k in itertools.permutations(range(3), 2)]
    else:
        c = [O.axis_cost(mom, xs, xm, n, pca)] + [O.axis_cost(mom, xs, xm, n, I[a]) for a in range(3)]
    c = np.sort(np.asarray(c))
    return c[1] > c[0] * (1 + 1e-6)


@pytest.mark.parametrize("slabs", [False, True])
def test_small_scene(dev, slabs):
    b = W.small()
    got, st = run(b, dev, G.AXIS_AUTO, slabs)
    assert np.array_equal(got, O.brute(b))
    src, u, v = O.axis_auto(b, slabs=slabs)
    assert separated(b, slabs)
    assert st.axis_source == src
    assert np.array_equal(np.asarray(st.axis, np.float32), u)


@pytest.mark.parametrize("slabs", [False, True])
def test_anisotropic_scene(dev, slabs):
    b = W.aniso(1 << 17, seed=1, sigma=(10.0, 2.5, 0.5))
    ref, st_pca = run(b, dev, G.AXIS_PCA, slabs)
    got, st = run(b, dev, G.AXIS_AUTO, slabs)
    assert np.array_equal(got, ref)
    src, u, v = O.axis_auto(b, slabs=slabs)
    assert separated(b, slabs)
    assert st.axis_source == src == 1
    assert np.array_equal(np.asarray(st.axis, np.float32), u)
    limit = 0.6 if slabs else 0.85  # numpy counts on this scene: 0.42x and 0.76x
    assert st.sweep_tests < limit * st_pca.sweep_tests, (st.sweep_tests, st_pca.sweep_tests)
    if slabs:  # the slab axis the kernel used is the oracle's: same test count
        t = slab_tests(b, u, v)
        assert abs(st.sweep_tests - t) <= 0.01 * t, (st.sweep_tests, t)


def test_spheres_and_degenerate(dev):
    r = np.random.default_rng(4)
    sph = np.concatenate([r.uniform(0, 40, (6000, 3)), r.uniform(0.1, 0.6, (6000, 1))], 1)
    sph = sph.astype(np.float32)
    for slabs in (False, True):
        got, st = run(sph, dev, G.AXIS_AUTO, slabs, kind=G.INPUT_SPHERE)
        assert np.array_equal(got, O.brute(sph, kind=O.KIND_SPHERE))
    same = np.repeat(np.array([[1, 2, 3, 2, 3, 4]], np.float32), 50, axis=0)
    for slabs in (False, True):
        got, st = run(same, dev, G.AXIS_AUTO, slabs)
        assert np.array_equal(got, O.brute(same)) and st.axis_source == 0


def test_batched_scenes(dev):
    import torch
    boxes, off = W.scenes(16, seed=8, lo=300, hi=6000)
    bd = to_dev(boxes, dev)
    refs = [O.brute(boxes[off[k]:off[k + 1]]) + off[k] for k in range(len(off) - 1)]
    total = sum(len(r) for r in refs)
    s = G.Sap(axis_mode=G.AXIS_AUTO)
    pairs = torch.empty((total + 16, 2), dtype=torch.int32, device=dev)
    spo = torch.empty(len(off), dtype=torch.int64, device=dev)
    n, _ = s.collide_batched(bd, off, pairs, spo)
    assert n == total
    assert np.array_equal(pairs_np(pairs[:n]), np.concatenate(refs))
    s.close()
