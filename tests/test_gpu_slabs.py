"""Slab mode (GSAP_FLAG_SLABS): the sweep runs inside slabs of a second axis
with each box keyed into the one or two slabs it touches.  The pair list must
be bit-identical to the plain sweep's and to brute force; only the number of
sweep tests changes."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import d2h, pairs_np, to_dev

pytestmark = pytest.mark.gpu


def run(boxes, dev, slabs, kind=G.INPUT_AABB, axis_mode=G.AXIS_PCA, group=None,
        output_kind=G.OUT_CANON):
    flags = (G.FLAG_SLABS if slabs else 0) | (G.FLAG_SKIP_SAME_GROUP if group is not None else 0)
    s = G.Sap(input_kind=kind, axis_mode=axis_mode, flags=flags, output_kind=output_kind)
    g = None if group is None else to_dev(group.astype(np.uint32).view(np.int32), dev)
    pairs, ent, st = s.collide(to_dev(boxes, dev), group=g)
    out = pairs.cpu().numpy().view(np.uint32).copy()
    e = None if ent is None else ent.cpu().numpy()
    s.close()
    return out, e, st


@pytest.mark.parametrize("mode", [G.AXIS_PCA, G.AXIS_COORD_MAXVAR])
def test_small_equals_brute_force(dev, mode):
    b = W.small()
    got, _, st = run(b, dev, True, axis_mode=mode)
    assert np.array_equal(got, O.brute(b))


def test_spread_scene_fewer_tests_same_pairs(dev):
    b = W.aniso(200_000, seed=7)
    ref, _, st0 = run(b, dev, False)
    got, _, st1 = run(b, dev, True)
    assert np.array_equal(got, ref)
    assert st1.sweep_tests < 0.5 * st0.sweep_tests, (st1.sweep_tests, st0.sweep_tests)


@pytest.mark.parametrize("n", [1, 2, 3, 255, 257, 2049, 4097])
def test_ragged_sizes(dev, n):
    b = W.small(n, seed=n)
    b[:, :3] *= 0.05
    b[:, 3:] = b[:, :3] + 0.4
    got, _, _ = run(b, dev, True)
    assert np.array_equal(got, O.brute(b))


def test_touching_duplicate_flat_and_degenerate(dev):
    grid = np.array([[i, j, k, i + 1, j + 1, k + 1] for i in range(8) for j in range(8)
                     for k in range(3)], np.float32)  # unit cubes sharing faces, edges, corners
    for mode in (G.AXIS_PCA, G.AXIS_COORD_MAXVAR):
        got, _, _ = run(grid, dev, True, axis_mode=mode)
        assert np.array_equal(got, O.brute(grid))
    dup = np.repeat(W.small(400, seed=4), 3, axis=0)
    assert np.array_equal(run(dup, dev, True)[0], O.brute(dup))
    flat = W.small(3000, seed=6)
    flat[:, 5] = flat[:, 2]
    assert np.array_equal(run(flat, dev, True)[0], O.brute(flat))
    same = np.repeat(np.array([[1, 2, 3, 2, 3, 4]], np.float32), 50, axis=0)  # no spread at all
    assert np.array_equal(run(same, dev, True)[0], O.brute(same))
    line = np.zeros((500, 6), np.float32)  # boxes on a line: no spread across the sweep axis
    line[:, 0] = np.arange(500) * 0.5
    line[:, 3] = line[:, 0] + 0.7
    line[:, 4] = line[:, 5] = 1.0
    assert np.array_equal(run(line, dev, True)[0], O.brute(line))


def test_many_slabs_and_one_huge_box(dev):
    rng = np.random.default_rng(3)
    c = rng.uniform(0, 1000, (30000, 3)).astype(np.float32)
    c[:, 2] *= 0.01
    b = W.boxes_from(c, np.full((30000, 3), 0.3, np.float32))  # range / extent >> 255
    ref = O.brute(b)
    assert np.array_equal(run(b, dev, True)[0], ref)
    big = b.copy()
    big[0, :3] = -1.0
    big[0, 3:] = 1001.0  # one box overlapping everything: one slab
    assert np.array_equal(run(big, dev, True)[0], O.brute(big))


def test_spheres_groups_sym_csr(dev):
    rng = np.random.default_rng(9)
    s = np.concatenate([rng.uniform(0, 40, (30000, 3)), np.full((30000, 1), 0.5)], 1).astype(np.float32)
    grp = (np.arange(30000) // 27).astype(np.uint32)
    ref, eref, _ = run(s, dev, False, kind=G.INPUT_SPHERE, group=grp, output_kind=G.OUT_SYM_CSR)
    got, e, _ = run(s, dev, True, kind=G.INPUT_SPHERE, group=grp, output_kind=G.OUT_SYM_CSR)
    assert np.array_equal(got, ref) and np.array_equal(e, eref)
    canon = O.brute(s, kind=O.KIND_SPHERE, group=grp)
    assert np.array_equal(O.sym_csr(canon)[0].view(np.uint32), got)


def test_full_size_aniso_identical_to_plain(dev):
    import torch
    bd = to_dev(W.aniso(), dev)
    p0, _, st0 = G.Sap().collide(bd)
    p1, _, st1 = G.Sap(flags=G.FLAG_SLABS).collide(bd)
    assert torch.equal(p0, p1)
    assert st1.num_pairs == st0.num_pairs and st1.sweep_tests < st0.sweep_tests / 3


def test_particle_step_identical(dev):
    """Ten rigid-body steps with the slab broadphase give the same state, bit
    for bit, as with the plain one (same pair lists)."""
    nb = 400
    sc = W.bodies(nb, seed=2, box=30.0, height=10.0)
    flags = G.PRB_ABS_SPRING | G.PRB_WALLS | G.PRB_SKIP_SAME_BODY
    out = []
    for slabs in (False, True):
        prb = G.ParticleRigidBody(nb, nb * 27, nb * 27 * 40)
        prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"],
                        sc["pstart"], sc["npart"], sc["inv_mass"], sc["inv_inertia"])
        prb.set_particle_data(sc["ref_pos"], sc["ref_mass"], sc["max_per_shape"])
        prb.set_particle_shape_inv_inertia(sc["shape_inv_inertia"])
        prb.set_physics(gravity=(0, -9.81, 0), k_abs=1e4, c_abs=50.0, floor_scale=1.0,
                        wall_lo=(0, 0), wall_hi=(30.0, 30.0), flags=flags,
                        broadphase_kind=G.INPUT_AABB)
        sap = G.Sap(input_kind=G.INPUT_AABB, output_kind=G.OUT_SYM_CSR,
                    flags=G.FLAG_SKIP_SAME_GROUP | (G.FLAG_SLABS if slabs else 0))
        for _ in range(10):
            prb.step(sap, 1e-4)
        p = prb.ptrs()
        out.append((d2h(p.rb_pos, (nb, 4), np.float32), d2h(p.rb_quat, (nb, 4), np.float32),
                    p.num_pairs))
    assert out[0][2] == out[1][2] and out[0][2] > 0
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("mode", [G.AXIS_AUTO, G.AXIS_PCA])
def test_batched_slabs_equal_brute_per_scene(dev, mode):
    """Batched slab sweep (each scene with its own slab axis and parameters):
    every scene's pairs equal brute force, on the captured update and its
    replays, with fewer sweep tests than the batched plain sweep; scenes of
    1 and 2 objects and a scene that fills one slab included."""
    import torch
    boxes, off = W.scenes(24, seed=7, lo=200, hi=6000)
    extra = [W.small(1, seed=1), W.small(2, seed=2)]
    one = W.small(300, seed=3)
    one[:, :3] *= 0.01
    one[:, 3:] = one[:, :3] + 0.5  # all overlapping: one slab
    parts = [boxes] + extra + [one]
    boxes = np.concatenate(parts).astype(np.float32)
    off = np.concatenate([off, off[-1] + np.cumsum([len(p) for p in parts[1:]])]).astype(np.uint32)
    bd = to_dev(boxes, dev)
    refs = [O.brute(boxes[off[k]:off[k + 1]]) + off[k] for k in range(len(off) - 1)]
    total = sum(len(r) for r in refs)
    plain = G.Sap(axis_mode=mode)
    _, st_plain = plain.collide_batched(bd, off, None, None)
    s = G.Sap(axis_mode=mode, flags=G.FLAG_SLABS)
    pairs = torch.empty((total + 16, 2), dtype=torch.int32, device=dev)
    spo = torch.empty(len(off), dtype=torch.int64, device=dev)
    for rep in range(3):
        got_n, st = s.collide_batched(bd, off, pairs, spo)
        assert got_n == total
        allp = pairs_np(pairs[:got_n])
        so = spo.cpu().numpy()
        for k in range(len(off) - 1):
            assert np.array_equal(allp[so[k]:so[k + 1]], refs[k]), (rep, k)
    assert st.sweep_tests < st_plain.sweep_tests


def test_full_size_bench_config_identical(dev):
    """The configuration the benchmark times, at its own size: 1,048,576
    anisotropic boxes (seed 1), automatic axis, slab sweep, 8 pair regions,
    XCD chunk 16, the per-part second pass with 4-byte items and the local
    pair stage.  Its list equals the default path's (plain sweep, one region,
    whole-key pair sort), which test_full_size_aniso_properties pins."""
    import torch
    bd = to_dev(W.aniso(), dev)
    p0, _, st0 = G.Sap().collide(bd)
    s = G.Sap(max_n=len(bd), axis_mode=G.AXIS_AUTO, flags=G.FLAG_SLABS)
    s.set_tune(slabs=1, pair_local=1, pair_regions=8, sweep_chunk=16, pair_items=16, pair_bits=8,
               sort_items=8, sort_bits=8, pair_grid=2048)
    p1, _, st1 = s.collide(bd)
    assert st1.num_pairs == st0.num_pairs
    assert torch.equal(p0, p1)
    del p1
    p2, _, _ = s.collide(bd)  # and again on the same handle (epochs, counters)
    assert torch.equal(p0, p2)
