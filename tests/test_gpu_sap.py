"""GPU parity of the SaP update (libgsap through the C ABI) against the CPU
oracle and brute force.  Integer/index results are compared bit-exactly."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import pairs_np, to_dev

pytestmark = pytest.mark.gpu


def gpu_pairs(boxes, dev, kind=G.INPUT_AABB, axis_mode=G.AXIS_PCA, fixed_axis=None, group=None,
              sort_items=0):
    s = G.Sap(input_kind=kind, axis_mode=axis_mode, fixed_axis=fixed_axis,
              flags=G.FLAG_SKIP_SAME_GROUP if group is not None else 0, sort_items=sort_items)
    b = to_dev(boxes, dev)
    g = None if group is None else to_dev(group.astype(np.uint32).view(np.int32), dev)
    pairs, _, st = s.collide(b, group=g)
    out = pairs_np(pairs)
    s.close()
    return out, st


@pytest.mark.parametrize("mode", [G.AXIS_PCA, G.AXIS_COORD_MAXVAR])
def test_small_config_equals_brute_force(dev, mode):
    """16,384 boxes, seed 0: canonical pair list bit-exact vs O(n^2)."""
    b = W.small()
    got, st = gpu_pairs(b, dev, axis_mode=mode)
    ref = O.brute(b)
    assert np.array_equal(got, ref)
    assert st.num_pairs == len(ref)


def test_axis_and_sweep_tests_match_oracle(dev):
    """PCA axis within fp64-summation-order tolerance; for that exact fp32 axis
    the oracle's sweep-test count (which depends on every key and on the sort
    order) equals the GPU's."""
    b = W.aniso(200_000, seed=11)
    got, st = gpu_pairs(b, dev)
    ax_gpu = np.array(st.axis[:], np.float32)
    m = O.moments(b)
    ax_cpu, _ = O.pca_axis(m, len(b))
    assert np.max(np.abs(ax_gpu - ax_cpu)) < 1e-6
    ref, tests, _ = O.sap(b, axis_mode=O.AXIS_FIXED, fixed_axis=ax_gpu)
    assert st.sweep_tests == tests
    assert np.array_equal(got, ref)


def test_coord_rule_axis_matches_oracle(dev):
    for seed in range(3):
        b = W.aniso(30_000, seed=20 + seed, sigma=(1.0 + seed, 3.0, 2.0))
        got, st = gpu_pairs(b, dev, axis_mode=G.AXIS_COORD_MAXVAR)
        _, ax, _ = O.coord_axis(O.moments(b), len(b))
        assert list(st.axis[:]) == list(ax)
        ref, tests, _ = O.sap(b, axis_mode=O.AXIS_COORD_MAXVAR)
        assert st.sweep_tests == tests and np.array_equal(got, ref)


@pytest.mark.parametrize("items", [4, 8, 16])
def test_fixed_axis_any_direction_same_pairs(dev, items):
    b = W.aniso(40_000, seed=3)
    ref = O.sap(b, axis_mode=O.AXIS_COORD_MAXVAR)[0]
    rng = np.random.default_rng(items)
    for _ in range(2):
        u = rng.standard_normal(3).astype(np.float32)
        got, st = gpu_pairs(b, dev, axis_mode=G.AXIS_FIXED, fixed_axis=u, sort_items=items)
        assert np.array_equal(got, ref)
        assert st.sweep_tests == O.sap(b, axis_mode=O.AXIS_FIXED, fixed_axis=u)[1]


def test_sphere_mode_and_sym_csr(dev):
    rng = np.random.default_rng(5)
    s = np.concatenate([rng.uniform(0, 30, (20000, 3)), rng.uniform(0.1, 0.7, (20000, 1))], 1)
    s = s.astype(np.float32)
    canon = O.brute(s, kind=O.KIND_SPHERE)
    got, _ = gpu_pairs(s, dev, kind=G.INPUT_SPHERE)
    assert np.array_equal(got, canon)
    sym_ref, ent_ref = O.sym_csr(canon)
    sap = G.Sap(input_kind=G.INPUT_SPHERE, axis_mode=G.AXIS_COORD_MAXVAR, output_kind=G.OUT_SYM_CSR)
    pairs, ent, st = sap.collide(to_dev(s, dev))
    assert np.array_equal(pairs.cpu().numpy(), sym_ref)
    assert np.array_equal(ent.cpu().numpy(), ent_ref)
    # the reference-named entry point writes into caller buffers
    import torch
    P = torch.empty((len(sym_ref) + 5, 2), dtype=torch.int32, device=dev)
    E = torch.empty(20000, dtype=torch.int32, device=dev)
    npairs, nent = sap.collision(1.0, to_dev(s, dev), P, E)
    assert npairs == len(sym_ref) and nent == len(ent_ref)
    assert np.array_equal(P[:npairs].cpu().numpy(), sym_ref)


def test_group_filter(dev):
    rng = np.random.default_rng(8)
    b = W.small(20000, seed=8)
    g = (np.arange(20000) // 27).astype(np.uint32)
    got, _ = gpu_pairs(b, dev, group=g)
    assert np.array_equal(got, O.brute(b, group=g))


def test_full_size_aniso_properties(dev):
    """1,048,576 anisotropic boxes: the list is canonical (i<j, strictly
    increasing), every pair overlaps, and for 64 random objects the set of
    partners equals an exhaustive test against all n boxes."""
    import torch
    b = W.aniso()
    n = len(b)
    s = G.Sap()
    bd = to_dev(b, dev)
    pairs, _, st = s.collide(bd)
    p = pairs.to(torch.int64) & 0xFFFFFFFF
    assert st.num_pairs == len(p) and len(p) > 1e8
    assert bool((p[:, 0] < p[:, 1]).all())
    key = p[:, 0] * n + p[:, 1]
    assert bool((key[1:] > key[:-1]).all())
    lo, hi = bd[:, :3], bd[:, 3:]
    ok = ((lo[p[:, 0]] <= hi[p[:, 1]]) & (lo[p[:, 1]] <= hi[p[:, 0]])).all(dim=1)
    assert bool(ok.all())
    rng = np.random.default_rng(0)
    for i in rng.choice(n, 64, replace=False).tolist():
        ov = ((lo <= hi[i]) & (lo[i] <= hi)).all(dim=1)
        ov[i] = False
        want = torch.nonzero(ov).flatten()
        rows = p[p[:, 0] == i, 1]
        cols = p[p[:, 1] == i, 0]
        have = torch.sort(torch.cat([rows, cols])).values
        assert torch.equal(have, want)
    # deterministic: a second update gives the identical list
    pairs2, _, st2 = s.collide(bd)
    assert torch.equal(pairs, pairs2) and st2.sweep_tests == st.sweep_tests


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 2047, 2049, 4097])
def test_ragged_sizes(dev, n):
    b = W.small(n, seed=n)
    b[:, :3] *= 0.05  # crowd them so small n still has pairs
    b[:, 3:] = b[:, :3] + 0.4
    got, _ = gpu_pairs(b, dev)
    assert np.array_equal(got, O.brute(b))


def test_touching_duplicate_and_flat_boxes(dev):
    base = W.small(300, seed=4)
    dup = np.repeat(base, 3, axis=0)
    got, _ = gpu_pairs(dup, dev)
    assert np.array_equal(got, O.brute(dup))
    grid = []
    for i in range(6):
        for j in range(6):
            grid.append([i, j, 0, i + 1, j + 1, 1])  # unit cubes sharing faces/corners
    grid = np.array(grid, np.float32)
    for mode in (G.AXIS_PCA, G.AXIS_COORD_MAXVAR):
        got, _ = gpu_pairs(grid, dev, axis_mode=mode)
        assert np.array_equal(got, O.brute(grid))
    flat = W.small(3000, seed=6)
    flat[:, 5] = flat[:, 2]
    got, _ = gpu_pairs(flat, dev)
    assert np.array_equal(got, O.brute(flat))


def test_empty_input(dev):
    import torch
    s = G.Sap()
    n, st = s.count(torch.empty((0, 6), dtype=torch.float32, device=dev))
    assert n == 0


def test_nonfinite_rejected(dev):
    b = W.small(1000)
    b[17, 2] = np.nan
    s = G.Sap()
    with pytest.raises(G.GsapError) as e:
        s.count(to_dev(b, dev))
    assert e.value.code == G.ERR_NONFINITE
    b[17, 2] = np.inf
    with pytest.raises(G.GsapError) as e:
        s.count(to_dev(b, dev))
    assert e.value.code == G.ERR_NONFINITE
    b[17, 2] = 1.0  # the handle recovers
    assert s.count(to_dev(b, dev))[0] == len(O.brute(b))


def test_capacity_error_then_retry(dev):
    import torch
    b = W.small()
    ref = O.brute(b)
    s = G.Sap()
    bd = to_dev(b, dev)
    small_buf = torch.empty((10, 2), dtype=torch.int32, device=dev)
    with pytest.raises(G.GsapError) as e:
        s.collide_into(bd, small_buf)
    assert e.value.code == G.ERR_CAPACITY
    big = torch.empty((len(ref), 2), dtype=torch.int32, device=dev)
    P, _, _ = s.collide_into(bd, big)
    assert P == len(ref) and np.array_equal(pairs_np(big), ref)


def test_sym_csr_capacity_error_then_retry(dev):
    # an overflowing SYM_CSR update writes no row starts (the pair sort is
    # skipped, and the row-start pass reads no pairs); the retry is exact
    import torch
    rng = np.random.default_rng(9)
    s = np.concatenate([rng.uniform(0, 30, (20000, 3)), rng.uniform(0.1, 0.7, (20000, 1))], 1)
    s = s.astype(np.float32)
    sym_ref, ent_ref = O.sym_csr(O.brute(s, kind=O.KIND_SPHERE))
    sap = G.Sap(input_kind=G.INPUT_SPHERE, axis_mode=G.AXIS_COORD_MAXVAR, output_kind=G.OUT_SYM_CSR)
    sd = to_dev(s, dev)
    small = torch.full((len(sym_ref) // 2, 2), -1, dtype=torch.int32, device=dev)
    E = torch.full((20000,), -7, dtype=torch.int32, device=dev)
    with pytest.raises(G.GsapError) as e:
        sap.collide_into(sd, small, E)
    assert e.value.code == G.ERR_CAPACITY
    assert bool((E == -7).all())
    P = torch.empty((len(sym_ref), 2), dtype=torch.int32, device=dev)
    npairs, nent, _ = sap.collide_into(sd, P, E)
    assert npairs == len(sym_ref) and nent == len(ent_ref)
    assert np.array_equal(P.cpu().numpy(), sym_ref)
    assert np.array_equal(E[:nent].cpu().numpy(), ent_ref)


def test_batched_equals_plain_per_scene(dev):
    import torch
    boxes, off = W.scenes(24, seed=7, lo=200, hi=5000)
    bd = to_dev(boxes, dev)
    s = G.Sap()
    total = 0
    refs = []
    for k in range(len(off) - 1):
        sub = boxes[off[k]:off[k + 1]]
        r = O.brute(sub) + off[k]
        refs.append(r)
        total += len(r)
    pairs = torch.empty((total + 16, 2), dtype=torch.int32, device=dev)
    spo = torch.empty(len(off), dtype=torch.int64, device=dev)
    for rep in range(3):  # first call captures the graph, later calls replay it
        got_n, st = s.collide_batched(bd, off, pairs, spo)
        assert got_n == total
        allp = pairs_np(pairs[:got_n])
        so = spo.cpu().numpy()
        for k in range(len(off) - 1):
            assert np.array_equal(allp[so[k]:so[k + 1]], refs[k])
        assert so[-1] == total


def test_radix_sort_api(dev):
    import torch
    rng = np.random.default_rng(2)
    for n in (1, 1000, 100_003, 1 << 20):
        k = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        k[: n // 4] = k[0]
        kd = to_dev(k.view(np.int32), dev)
        vd = torch.arange(n, dtype=torch.int32, device=dev)
        rs = G.RadixSort(n)
        rs.sort(kd, vd, 32)
        o = np.argsort(k, kind="stable")
        assert np.array_equal(kd.cpu().numpy().view(np.uint32), k[o])
        assert np.array_equal(vd.cpu().numpy(), o.astype(np.int32))
        f = rng.standard_normal(n).astype(np.float32) * 100
        fd = to_dev(f, dev)
        vd = torch.arange(n, dtype=torch.int32, device=dev)
        rs.sort(fd, vd)
        o = np.argsort(f, kind="stable")
        assert np.array_equal(fd.cpu().numpy(), f[o]) and np.array_equal(vd.cpu().numpy(), o)
        # 20-bit keys: three digit passes
        k20 = k & np.uint32((1 << 20) - 1)
        kd = to_dev(k20.view(np.int32), dev)
        vd = torch.arange(n, dtype=torch.int32, device=dev)
        rs.sort(kd, vd, 20)
        o = np.argsort(k20, kind="stable")
        assert np.array_equal(kd.cpu().numpy().view(np.uint32), k20[o])
        assert np.array_equal(vd.cpu().numpy(), o.astype(np.int32))
        rs.close()


def test_batched_empty_and_tiny_scenes(dev):
    # A scene table with empty, one-box and two-box scenes between large ones
    # (enough sort tiles for the reduce-then-scan passes): every scene's pairs
    # equal a single-scene update of it, and empty scenes own no pairs.
    import torch
    sizes = [0, 1, 60_000, 0, 120_000, 1, 0, 90_000, 2, 0]
    rng = np.random.default_rng(17)
    parts = [W.scene(k, rng, "uniform") if k else np.zeros((0, 6), np.float32) for k in sizes]
    parts[8] = np.array([[0, 0, 0, 1, 1, 1], [0.5, 0.5, 0.5, 2, 2, 2]], np.float32)  # one pair
    boxes = np.concatenate(parts).astype(np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    bd = to_dev(boxes, dev)
    refs = []
    for k in range(len(sizes)):
        if sizes[k] == 0:
            refs.append(np.zeros((0, 2), np.uint32))
            continue
        got, _, _ = G.Sap().collide(to_dev(boxes[off[k]:off[k + 1]], dev))
        refs.append(pairs_np(got) + off[k])
    total = sum(len(r) for r in refs)
    assert len(refs[8]) == 1
    for slabs in (0, 1):
        s = G.Sap()
        s.set_tune(slabs=slabs, pair_regions=8, sort_items=8)
        P, _ = s.collide_batched(bd, off, None, None)
        assert P == total
        pairs = torch.empty((P, 2), dtype=torch.int32, device=dev)
        po = torch.empty(len(off), dtype=torch.int64, device=dev)
        for _ in range(2):  # capture, then replay
            s.collide_batched(bd, off, pairs, po)
            allp, so = pairs_np(pairs), po.cpu().numpy()
            for k in range(len(sizes)):
                assert np.array_equal(allp[so[k]:so[k + 1]], refs[k]), (slabs, k)
