"""Automatic sweep axis (GSAP_AXIS_AUTO, oracle or_axis_auto): the PCA axis
(pair) unless a coordinate axis (pair) is cheaper by the sweep-test model.

The model is pinned by what it predicts: on each scene, the sweep tests that
the chosen axis (pair) actually costs -- counted here by numpy over sorted
projections, independently of the oracle's C code -- are within a few percent
of the best candidate's and never above the PCA choice's.  The formula is
pinned by closed forms (isotropic spread, coordinate pairs), and the slab
axis by numpy's eigen-decomposition of the projected covariance."""
import itertools

import numpy as np
import pytest

from gsap_amd import workloads as W
from oracle import oracle as O


def _proj(b, u):
    lo, hi = b[:, :3].astype(np.float64), b[:, 3:].astype(np.float64)
    up, un = np.maximum(u, 0), np.minimum(u, 0)
    return lo @ up + hi @ un, hi @ up + lo @ un


def plain_tests(b, u):
    l, h = _proj(b, np.asarray(u, np.float64))
    s = np.sort(l)
    return int((np.searchsorted(s, h, "right") - np.searchsorted(s, l, "left")).sum())


def slab_tests(b, u, v):
    """Window sizes of a sweep along u inside slabs of a second axis v of width
    max v-extent (each box in the one or two slabs it touches); a window
    counts its own box."""
    return slab_windows(b, u, v)[0]


def slab_windows(b, u, v):
    """(slab_tests(b, u, v), the number of slab entries)."""
    l, h = _proj(b, np.asarray(u, np.float64))
    vl, vh = _proj(b, np.asarray(v, np.float64))
    w = (vh - vl).max() * (1 + 2 ** -10)
    v0, span = vl.min(), vh.max() - vl.min()
    w = max(w, span / 254 * (1 + 2 ** -10))
    s0 = np.floor((vl - v0) / w).astype(np.int64)
    s1 = np.floor((vh - v0) / w).astype(np.int64)
    t = e = 0
    for s in range(int(s1.max()) + 1):
        m = (s0 <= s) & (s1 >= s)
        ss = np.sort(l[m])
        t += int((np.searchsorted(ss, h[m], "right") - np.searchsorted(ss, l[m], "left")).sum())
        e += int(m.sum())
    return t, e


def scenes():
    r = np.random.default_rng(11)
    out = [("aniso", W.aniso(1 << 16, seed=1, sigma=(5.0, 1.25, 0.25))),
           ("uniform", W.small(1 << 15, seed=2))]
    for k in range(3):  # rotated anisotropic Gaussians with anisotropic box extents
        R = W.random_rotation(r)
        c = (r.standard_normal((1 << 15, 3)) * r.uniform(0.3, 8.0, 3)) @ R.T
        h = r.uniform(0.02, 0.3, (1 << 15, 3)) * r.uniform(0.2, 2.0, 3)
        out.append((f"gauss{k}", W.boxes_from(c, h)))
    sc, off = W.scenes(2, seed=9)
    out += [(f"scene{s}", sc[off[s]:off[s + 1]]) for s in range(2)]
    return out


@pytest.mark.parametrize("name,b", scenes(), ids=lambda x: x if isinstance(x, str) else "")
def test_plain_choice_is_near_best(name, b):
    src, u, _ = O.axis_auto(b)
    mom = O.moments(b)
    pca, _ = O.pca_axis(mom, len(b))
    cands = [pca] + [np.eye(3, dtype=np.float32)[a] for a in range(3)]
    counts = [plain_tests(b, c) for c in cands]
    got = plain_tests(b, u)
    assert got <= counts[0], (name, got, counts)            # never worse than PCA
    assert got <= 1.05 * min(counts), (name, got, counts)   # within 5% of the best candidate
    assert src == (0 if np.array_equal(u, pca) else 1)


@pytest.mark.parametrize("name,b", scenes(), ids=lambda x: x if isinstance(x, str) else "")
def test_slab_choice_is_near_best(name, b):
    src, u, v = O.axis_auto(b, slabs=True)
    mom = O.moments(b)
    pca, _ = O.pca_axis(mom, len(b))
    pv = O.slab_axis(mom, len(b), pca)
    I = np.eye(3, dtype=np.float32)
    cands = [(pca, pv)] + [(I[a], I[c]) for a, c in itertools.permutations(range(3), 2)]
    counts = [slab_tests(b, *c) for c in cands]
    got = slab_tests(b, u, v)
    assert got <= 1.02 * counts[0], (name, got, counts)
    assert got <= 1.15 * min(counts), (name, got, counts)


def test_headline_scene_prefers_coordinates():
    b = W.aniso(1 << 18, seed=1)
    src, u, v = O.axis_auto(b, slabs=True)
    mom = O.moments(b)
    pca, _ = O.pca_axis(mom, len(b))
    assert src == 1
    assert slab_tests(b, u, v) < 0.5 * slab_tests(b, pca, O.slab_axis(mom, len(b), pca))


def test_cost_closed_forms():
    # centres: an exact isotropic 4-point design (covariance = I / 2 in every
    # coordinate pair); extents 0.2, 0.4, 0.6 per coordinate
    c = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64) / np.sqrt(2)
    h = np.array([0.1, 0.2, 0.3])
    b = W.boxes_from(c, np.tile(h, (4, 1)))
    mom, (xs, xm) = O.moments(b), O.extents(b)
    n = len(b)
    np.testing.assert_allclose(xs / n, 2 * h, rtol=1e-6)
    np.testing.assert_allclose(xm, 2 * h, rtol=1e-6)
    I = np.eye(3, dtype=np.float32)
    var = 0.5
    for a in range(3):
        np.testing.assert_allclose(O.axis_cost(mom, xs, xm, n, I[a]), 2 * h[a] / np.sqrt(var), rtol=1e-6)
        for k in range(3):
            if k != a:
                np.testing.assert_allclose(O.axis_cost(mom, xs, xm, n, I[a], I[k]),
                                           2 * h[a] * 2 * h[k] / var, rtol=1e-6)
    d = np.array([1, 1, 1], np.float32) / np.float32(np.sqrt(3))
    np.testing.assert_allclose(O.axis_cost(mom, xs, xm, n, d), (2 * h).sum() / np.sqrt(3) / np.sqrt(var),
                               rtol=1e-6)
    # no spread: every cost is infinite and PCA is kept
    same = np.repeat(b[:1], 10, axis=0)
    assert O.axis_auto(same)[0] == 0 and O.axis_auto(same, slabs=True)[0] == 0


def test_slab_axis_is_in_plane_principal():
    r = np.random.default_rng(5)
    for _ in range(5):
        R = W.random_rotation(r)
        c = (r.standard_normal((4096, 3)) * r.uniform(0.5, 6.0, 3)) @ R.T
        b = W.boxes_from(c, np.full((4096, 3), 0.1))
        mom = O.moments(b)
        u = np.asarray(r.standard_normal(3), np.float32)
        u /= np.linalg.norm(u)
        v = O.slab_axis(mom, len(b), u)
        assert abs(float(v @ u)) < 1e-6 and abs(np.linalg.norm(v) - 1) < 1e-6
        C = np.cov(0.5 * (b[:, :3].astype(np.float64) + b[:, 3:]).T, bias=True)
        P = np.eye(3) - np.outer(u, u)
        ev, V = np.linalg.eigh(P @ C @ P)
        assert abs(abs(float(V[:, 2] @ v)) - 1) < 1e-5
