"""CPU checks that pin the oracle (oracle/gsap_oracle.c) independently of its
own formulas: exhaustive float-flip order, library sorts, numpy eigh, brute
force, the reference's own pair dump, and closed-form dynamics."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from gsap_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sphere_pairs_2_1024.json")


def test_flip_exhaustive_order():
    """All 2^32 patterns: non-NaN floats in numeric order map to consecutive
    integers 0x007FFFFF..0xFF800000 (so the map is a monotone bijection and
    -0 < +0); NaNs map outside that range."""
    # negative floats in increasing numeric order: -inf (0xFF800000) .. -0 (0x80000000)
    step = 1 << 24
    expect = 0x007FFFFF
    for start in range(0xFF800000, 0x80000000 - 1, -step):
        stop = max(start - step, 0x80000000 - 1)
        u = np.arange(start, stop, -1, dtype=np.int64).astype(np.uint32)
        k = O.flip(u.view(np.float32))
        ref = np.arange(expect, expect + len(u), dtype=np.int64).astype(np.uint32)
        assert np.array_equal(k, ref)
        expect += len(u)
    assert expect == 0x80000000
    for start in range(0, 0x7F800001, step):
        stop = min(start + step, 0x7F800001)
        u = np.arange(start, stop, dtype=np.int64).astype(np.uint32)
        k = O.flip(u.view(np.float32))
        assert np.array_equal(k, (u.astype(np.int64) + 0x80000000).astype(np.uint32))
    nan = np.array([0x7FC00000, 0x7F800001, 0xFFC00000, 0xFFFFFFFF], np.uint32).view(np.float32)
    k = O.flip(nan).astype(np.int64)
    assert np.all((k < 0x007FFFFF) | (k > 0xFF800000))


def test_unflip_inverse():
    rng = np.random.default_rng(1)
    u = rng.integers(0, 2 ** 32, 100000, dtype=np.uint64).astype(np.uint32)
    f = u.view(np.float32)
    f = f[np.isfinite(f)]
    L = O.lib()
    for x in f[:2000]:
        assert np.float32(L.or_unflip(L.or_flip(float(x)))).view(np.uint32) == np.float32(x).view(np.uint32)


@pytest.mark.parametrize("n,digit", [(1, 8), (1000, 8), (65537, 8), (5000, 4), (3000, 11)])
def test_radix_sort_kv_matches_numpy_stable(n, digit):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    k[: n // 3] = k[0]  # many ties: stability matters
    sk, sv = O.radix_sort_kv(k, np.arange(n, dtype=np.uint32), digit)
    o = np.argsort(k, kind="stable")
    assert np.array_equal(sk, k[o]) and np.array_equal(sv, o)


def test_radix_sort_pairs_lexicographic():
    rng = np.random.default_rng(3)
    p = rng.integers(0, 1 << 20, (20000, 2), dtype=np.int64).astype(np.uint32)
    s = O.radix_sort_pairs(p, 20, 40)
    o = np.lexsort((p[:, 1], p[:, 0]))
    assert np.array_equal(s, p[o])


def test_pca_axis_matches_eigh():
    b = W.aniso(20000, seed=5)
    m = O.moments(b)
    ax, ev = O.pca_axis(m, len(b))
    c = 0.5 * (b[:, :3].astype(np.float64) + b[:, 3:].astype(np.float64))
    w, V = np.linalg.eigh(np.cov(c.T, bias=True))
    v = V[:, np.argmax(w)]
    assert abs(abs(float(np.dot(v, ax))) - 1.0) < 1e-6
    assert np.allclose(sorted(ev), sorted(w), rtol=1e-9)
    # closed form: covariance of the generator is R diag(400,25,1) R^T
    r = np.random.Generator(np.random.PCG64(5))
    R = W.random_rotation(r)
    assert abs(abs(float(np.dot(R[:, 0], ax))) - 1.0) < 2e-3


def test_coord_maxvar_rule():
    # reference axisVecMultiple: ties go to the lower index (>=)
    def mom_for(var):
        # moments of a zero-mean set with the given diagonal variances
        return np.array([0, 0, 0, var[0], 0, 0, var[1], 0, var[2]], np.float64)
    for var, want in [((3, 2, 1), 0), ((1, 3, 2), 1), ((1, 2, 3), 2), ((2, 2, 1), 0),
                      ((1, 2, 2), 1), ((2, 1, 2), 0), ((1, 1, 1), 0)]:
        a, ax, lam = O.coord_axis(mom_for(var), 1)
        assert a == want
        assert list(lam) == [want, 1 if want == 0 else 0, 3 - want - (1 if want == 0 else 0)]


def test_projection_contains_box():
    b = W.aniso(5000, seed=7)
    rng = np.random.default_rng(2)
    for _ in range(5):
        u = rng.standard_normal(3).astype(np.float32)
        u /= np.linalg.norm(u)
        keys, hik = O.project(b, u)
        lo = O.lib()
        corners = []
        for mask in range(8):
            sel = [(mask >> k) & 1 for k in range(3)]
            p = np.stack([b[:, 3 + k] if sel[k] else b[:, k] for k in range(3)], 1).astype(np.float64)
            corners.append(p @ u.astype(np.float64))
        cmin = np.min(corners, 0)
        cmax = np.max(corners, 0)
        lo_f = np.array([lo.or_unflip(int(x)) for x in keys[:500]])
        hi_f = np.array([lo.or_unflip(int(x)) for x in hik[:500]])
        assert np.all(lo_f <= cmin[:500]) and np.all(hi_f >= cmax[:500])


def test_projection_exact_on_coordinate_axis():
    b = W.small(1000)
    keys, hik = O.project(b, np.array([0, 1, 0], np.float32))
    assert np.array_equal(keys, O.flip(b[:, 1])) and np.array_equal(hik, O.flip(b[:, 4]))


def test_aabb_predicate_hand_cases():
    box = lambda lo, hi: np.array(lo + hi, np.float32)
    a = box([0, 0, 0], [1, 1, 1])
    touching = box([1, 0, 0], [2, 1, 1])          # shares a face: inclusive <=
    corner = box([1, 1, 1], [2, 2, 2])            # shares a corner
    apart = box([np.nextafter(np.float32(1), np.float32(2)), 0, 0], [2, 1, 1])
    inside = box([0.2, 0.2, 0.2], [0.3, 0.3, 0.3])
    L = O.lib()
    P = lambda x: x.ctypes.data_as(__import__("ctypes").c_void_p)
    assert L.or_overlap_aabb(P(a), P(touching)) == 1
    assert L.or_overlap_aabb(P(a), P(corner)) == 1
    assert L.or_overlap_aabb(P(a), P(apart)) == 0
    assert L.or_overlap_aabb(P(a), P(inside)) == 1
    s = lambda x, y, z, r: np.array([x, y, z, r], np.float32)
    assert L.or_overlap_sphere(P(s(0, 0, 0, 1)), P(s(2, 0, 0, 1))) == 0   # touching: strict <
    assert L.or_overlap_sphere(P(s(0, 0, 0, 1)), P(s(1.99, 0, 0, 1))) == 1


def _brute_np(b):
    lo, hi = b[:, :3], b[:, 3:]
    ov = np.all((lo[:, None, :] <= hi[None, :, :]) & (lo[None, :, :] <= hi[:, None, :]), axis=2)
    i, j = np.nonzero(np.triu(ov, 1))
    return np.stack([i, j], 1).astype(np.uint32)


def test_brute_matches_numpy_broadcast():
    b = W.small(1500, seed=9)
    assert np.array_equal(O.brute(b), _brute_np(b))


@pytest.mark.parametrize("mode", [O.AXIS_PCA, O.AXIS_COORD_MAXVAR])
def test_sap_equals_brute_small_config(mode):
    b = W.small()
    pairs, tests, ax = O.sap(b, axis_mode=mode)
    ref = O.brute(b)
    assert np.array_equal(pairs, ref)
    assert 500 < len(ref) < 1500  # ~0.10 overlaps per box at 16,384 boxes


@pytest.mark.parametrize("seed", [1, 2])
def test_sap_equals_brute_aniso_small(seed):
    b = W.aniso(6000, seed=seed)
    for mode in (O.AXIS_PCA, O.AXIS_COORD_MAXVAR):
        pairs, _, _ = O.sap(b, axis_mode=mode)
        assert np.array_equal(pairs, O.brute(b))
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(3).astype(np.float32)
    pairs, _, _ = O.sap(b, axis_mode=O.AXIS_FIXED, fixed_axis=u)
    assert np.array_equal(pairs, O.brute(b))


def test_sap_sphere_and_group_filter():
    rng = np.random.default_rng(4)
    s = np.concatenate([rng.uniform(0, 20, (3000, 3)), rng.uniform(0.1, 0.6, (3000, 1))], 1)
    s = s.astype(np.float32)
    pairs, _, _ = O.sap(s, kind=O.KIND_SPHERE)
    assert np.array_equal(pairs, O.brute(s, kind=O.KIND_SPHERE))
    g = (np.arange(3000) // 7).astype(np.uint32)
    pairs, _, _ = O.sap(s, kind=O.KIND_SPHERE, group=g)
    ref = O.brute(s, kind=O.KIND_SPHERE, group=g)
    assert np.array_equal(pairs, ref) and np.all(g[ref[:, 0]] != g[ref[:, 1]])


def test_sap_edge_cases():
    assert len(O.sap(np.zeros((0, 6), np.float32))[0]) == 0
    one = W.small(1)
    assert len(O.sap(one)[0]) == 0
    dup = np.repeat(W.small(3), 4, axis=0)  # duplicates overlap each other
    assert np.array_equal(O.sap(dup)[0], O.brute(dup))
    flat = W.small(500)
    flat[:, 5] = flat[:, 2]  # zero extent in z
    assert np.array_equal(O.sap(flat)[0], O.brute(flat))


def test_golden_2_1024_sphere_pairs():
    """The reference's own CPU broadphase dump: every listed pair satisfies the
    sphere predicate; brute force over the spheres that appear in a frame
    reproduces exactly the frame's pair set; the symmetric-CSR builder
    reproduces its pairEntries."""
    d = json.load(open(GOLD))
    assert len(d["frames"]) == 9
    for fr in d["frames"]:
        ids = sorted(int(k) for k in fr["pos"])
        pos = np.array([fr["pos"][str(i)] for i in ids], np.float32)
        directed = {(a, b) for a, b in fr["pairs"]}
        assert len(directed) == len(fr["pairs"]) == fr["n_pairs_reported"]
        assert all((b, a) in directed for a, b in directed)          # symmetric
        L = O.lib()
        P = lambda x: x.ctypes.data_as(__import__("ctypes").c_void_p)
        idx = {i: k for k, i in enumerate(ids)}
        for a, b in fr["pairs"]:
            assert L.or_overlap_sphere(P(pos[idx[a]]), P(pos[idx[b]])) == 1
        found = O.brute(pos, kind=O.KIND_SPHERE)
        got = {(ids[i], ids[j]) for i, j in found}
        want = {(min(a, b), max(a, b)) for a, b in directed}
        assert got == want
        canon = np.array(sorted(want), np.uint32).reshape(-1, 2)
        sym, ent = O.sym_csr(canon)
        assert sorted(map(tuple, sym.tolist())) == sorted(directed)
        assert list(ent) == fr["entries"]
        assert len(ent) == fr["n_entries_reported"]


def test_sym_csr_format():
    canon = np.array([[0, 3], [0, 5], [2, 3]], np.uint32)
    sym, ent = O.sym_csr(canon)
    assert sym.tolist() == [[0, 3], [0, 5], [2, 3], [3, 0], [3, 2], [5, 0]]
    assert ent.tolist() == [0, 2, 3, 5]


# ---------------------------------------------------------------- dynamics
def _one_body(y=10.0, vy=0.0, quat=(0, 0, 0, 1), npart=1, ref=None):
    ref = np.array([[0, 0, 0, 0.5]], np.float32) if ref is None else ref
    return O.PrbState(pos=[[0, y, 0, 0]], quat=[quat], linvel=[[0, vy, 0, 0]], angvel=[[0, 0, 0, 0]],
                      shape=[0], pstart=[0], npart=[npart], inv_mass=[1.0], ref_pos=ref,
                      ref_mass=np.ones(len(ref), np.float32), max_per_shape=len(ref))


def test_free_fall_closed_form():
    s = _one_body(y=100.0)
    p = O.prb_params(gravity=(0, -9.81, 0))
    dt = np.float32(1e-3)
    v = np.float32(0)
    y = np.float32(100)
    for _ in range(50):
        s.step(float(dt), p)
        v = np.float32(v + np.float32(np.float32(-9.81) * dt))
        y = np.float32(y + np.float32(v * dt))
    assert s.linvel[0, 1] == v and s.pos[0, 1] == y


def test_floor_equilibrium_abs_spring():
    """A resting particle on the floor under k=1e4, c=50: equilibrium depth
    m g / k; the damped system settles there."""
    s = _one_body(y=0.5)
    p = O.prb_params(gravity=(0, -9.81, 0), k_abs=1e4, c_abs=50.0, floor_scale=1.0,
                     flags=O.PRB_ABS_SPRING)
    for _ in range(4000):
        s.step(1e-4, p)
    y = s.pos[0, 1]
    assert abs((0.5 - y) - 9.81 / 1e4) < 1e-5


def test_two_body_momentum_conserved():
    ref = np.array([[0, 0, 0, 0.5]], np.float32)
    s = O.PrbState(pos=[[0, 10, 0, 0], [0.9, 10, 0, 0]], quat=[[0, 0, 0, 1]] * 2,
                   linvel=[[1, 0, 0, 0], [-1, 0, 0, 0]], angvel=[[0, 0, 0, 0]] * 2, shape=[0, 0],
                   pstart=[0, 1], npart=[1, 1], inv_mass=[1.0, 1.0], ref_pos=ref,
                   ref_mass=np.ones(1, np.float32), max_per_shape=1)
    p = O.prb_params(gravity=(0, 0, 0))
    for _ in range(20):
        s.step(1e-3, p)
    assert abs(float(s.linvel[0, 0] + s.linvel[1, 0])) < 1e-6
    assert s.linvel[0, 0] < 1.0  # they pushed each other apart


def test_quaternion_stays_unit_and_torque():
    ref = W.cube_template()
    s = O.PrbState(pos=[[0, 50, 0, 0]], quat=[[0.1, 0.2, 0.3, 0.9]], linvel=[[0, 0, 0, 0]],
                   angvel=[[1.0, 2.0, -0.5, 0]], shape=[0], pstart=[0], npart=[27],
                   inv_mass=[1.0 / 27], ref_pos=ref, ref_mass=np.ones(27, np.float32),
                   max_per_shape=27)
    s.quat[:] = s.quat / np.linalg.norm(s.quat)
    # voxel particles touch exactly; without the same-body filter rounding-level
    # intra-body contacts would inject energy (the reference spaces them 2r*1.001)
    p = O.prb_params(gravity=(0, 0, 0), flags=O.PRB_SKIP_SAME_BODY)
    for _ in range(100):
        s.step(1e-3, p)
        assert abs(float(np.linalg.norm(s.quat[0].astype(np.float64))) - 1.0) < 1e-6
    # torque = sum r x F for an offset force on one particle
    s.update_particles()
    s.pforce[:] = 0
    s.pforce[0, :3] = [0, 1, 0]
    w0 = s.angvel.copy()
    s.update_velocity(1.0, p)
    r = s.ppos[0, :3] - s.pos[0, :3]
    assert np.allclose(s.angvel[0, :3] - w0[0, :3], np.cross(r, [0, 1, 0]), atol=1e-5)


def test_ref_cells_equals_brute_force():
    """The reference's cell + cross pass finds the same pairs as brute force,
    on centred scenes (boxes on both sides of the split planes) for AABBs and
    spheres, and on a scene entirely inside one cell."""
    rng = np.random.default_rng(17)
    c = rng.normal(0.0, 6.0, (6000, 3)).astype(np.float32)
    h = rng.uniform(0.1, 0.9, (6000, 3)).astype(np.float32)
    boxes = np.concatenate([c - h, c + h], 1)
    pairs, st = O.sap_ref_cells(boxes)
    assert np.array_equal(pairs, O.brute(boxes))
    assert st["relevant"] > 0 and st["cross_tests"] > 0 and st["raw_pairs"] > len(pairs)
    sph = np.concatenate([c, rng.uniform(0.2, 0.8, (6000, 1)).astype(np.float32)], 1)
    pairs, st = O.sap_ref_cells(sph, kind=O.KIND_SPHERE)
    assert np.array_equal(pairs, O.brute(sph, kind=O.KIND_SPHERE))
    one = np.abs(boxes) + 1.0  # all centres positive: one cell, no straddlers
    one = np.concatenate([np.minimum(one[:, :3], one[:, 3:]), np.maximum(one[:, :3], one[:, 3:])], 1)
    pairs, st = O.sap_ref_cells(one)
    assert np.array_equal(pairs, O.brute(one))
    # count-only mode reports the same work
    _, st2 = O.sap_ref_cells(boxes, count_only=True)
    _, st1 = O.sap_ref_cells(boxes)
    assert st2["cell_tests"] == st1["cell_tests"] and st2["cross_tests"] == st1["cross_tests"]


def test_ref_cells_straddlers_across_planes():
    """Hand case: two boxes overlapping across a split plane, each in a
    different cell, are found by the cross pass only."""
    b = np.array([[0, -0.3, 1, 1, -0.1, 2],       # centre y < 0
                  [0.5, -0.2, 1.5, 1.5, 0.6, 2.5],  # centre y > 0, straddles y = 0
                  [50, 5, 5, 51, 6, 6]], np.float32)
    # make x the major axis (largest variance)
    far = np.array([[100, 1, 1, 101, 2, 2], [-100, 1, 1, -99, 2, 2]], np.float32)
    boxes = np.concatenate([b, far], 0)
    pairs, st = O.sap_ref_cells(boxes)
    assert pairs.tolist() == [[0, 1]]
    assert st["relevant"] >= 2


def test_cpu_baseline_equals_oracle_and_brute():
    """The timed CPU baseline (oracle/cpu_sap.c, OpenMP) gives the oracle's
    canonical list and sweep-test count on 1 and several threads, with and
    without the group filter, and the brute-force list."""
    for seed, n in ((0, 16384), (3, 40000)):
        b = W.small(n, seed=seed)
        if seed:
            b = W.aniso(n, seed=seed)
        ref, tests, _ = O.sap(b, axis_mode=O.AXIS_COORD_MAXVAR)
        for th in (1, 3):
            p, t, _, _ = O.cpu_sap(b, th)
            assert np.array_equal(p, ref) and t == tests
        if n == 16384:
            assert np.array_equal(p, O.brute(b))
    g = (np.arange(20000) // 27).astype(np.uint32)
    b = W.small(20000, seed=8)
    b[:, 3:] = b[:, :3] + 3.0 * (b[:, 3:] - b[:, :3])
    p, _, _, _ = O.cpu_sap(b, 4, group=g)
    assert np.array_equal(p, O.brute(b, group=g)) and len(p) > 0
    # capacity: a too-small buffer is retried with the exact size
    p2, _, _, _ = O.cpu_sap(b, 2, group=g, cap=16)
    assert np.array_equal(p2, p)
