"""Temporal coherence (GSAP_FLAG_COHERENT): successive updates of one object
set keep the key function and sort through buckets picked from the previous
update's order.  Every update's pair list must equal a from-scratch update's
(which brute force and the oracle pin elsewhere), including updates where
objects jump far, buckets overflow (the device falls back to the regular
sort) and the key function is refreshed; rigid-body trajectories must be
bit-identical to the from-scratch broadphase's."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import d2h, pairs_np, to_dev

pytestmark = pytest.mark.gpu


def _moving(n, seed, steps):
    """n AABBs drifting by small random steps with 20 jumping far each step;
    at the middle step the whole scene is shrunk 50x about the origin (its
    keys then fall between two of the previous update's splitters, so a
    bucket overflows and the device falls back to the regular sort), and the
    step after it is back at full size (the same, the other way)."""
    rng = np.random.default_rng(seed)
    b = W.aniso(n, seed=seed).astype(np.float32)
    c, h = 0.5 * (b[:, :3] + b[:, 3:]), 0.5 * (b[:, 3:] - b[:, :3])
    v = rng.normal(0, 0.02, (n, 3)).astype(np.float32)
    for t in range(steps):
        c = c + v
        jump = rng.choice(n, 20, replace=False)
        c[jump] = rng.normal(0, 10, (20, 3))
        s = np.float32(0.02) if t == steps // 2 else np.float32(1.0)
        yield np.concatenate([s * (c - h), s * (c + h)], 1).astype(np.float32)


@pytest.mark.parametrize("slabs,mode", [(True, G.AXIS_AUTO), (False, G.AXIS_PCA)])
def test_coherent_updates_equal_from_scratch(dev, slabs, mode):
    import torch
    flags = G.FLAG_SLABS if slabs else 0
    coh = G.Sap(axis_mode=mode, flags=flags | G.FLAG_COHERENT)
    if slabs:
        coh.set_tune(slabs=1, pair_local=1, pair_regions=8)
    steps = 70  # past one refresh of the key function (every 64 updates)
    for t, boxes in enumerate(_moving(120_000, 5 if slabs else 6, steps)):
        x = to_dev(boxes, dev)
        ref, _, st0 = G.Sap(axis_mode=mode, flags=flags).collide(x)
        got, _, st1 = coh.collide(x)  # pre-count + update, both coherent
        assert st1.num_pairs == st0.num_pairs, t
        assert torch.equal(got, ref), t
        if t in (0, 1, steps // 2, steps // 2 + 1, steps - 1):
            want, _, _, _ = O.cpu_sap(boxes, 8)  # the oracle's list (pinned in the CPU tests)
            assert np.array_equal(pairs_np(got), want), t


def test_coherent_particle_trajectories_bit_identical(dev):
    """100 rigid-body steps (400 cubes, slab broadphase on particle AABBs,
    same-body pairs skipped, symmetric CSR) with a coherent broadphase handle
    and with a from-scratch one: body positions, quaternions and velocities
    identical bit for bit after every 10 steps."""
    nb, box = 400, 30.0
    sc = W.bodies(nb, seed=2, box=box, height=10.0)
    flags = G.PRB_ABS_SPRING | G.PRB_WALLS | G.PRB_SKIP_SAME_BODY
    prbs, saps = [], []
    for coherent in (False, True):
        prb = G.ParticleRigidBody(nb, nb * 27, nb * 27 * 40)
        prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"],
                        sc["pstart"], sc["npart"], sc["inv_mass"], sc["inv_inertia"])
        prb.set_particle_data(sc["ref_pos"], sc["ref_mass"], sc["max_per_shape"])
        prb.set_particle_shape_inv_inertia(sc["shape_inv_inertia"])
        prb.set_physics(gravity=(0, -9.81, 0), k_abs=1e4, c_abs=50.0, floor_scale=1.0,
                        wall_lo=(0, 0), wall_hi=(box, box), flags=flags,
                        broadphase_kind=G.INPUT_AABB)
        sap = G.Sap(input_kind=G.INPUT_AABB, output_kind=G.OUT_SYM_CSR, axis_mode=G.AXIS_AUTO,
                    flags=G.FLAG_SKIP_SAME_GROUP | G.FLAG_SLABS |
                    (G.FLAG_COHERENT if coherent else 0))
        prbs.append(prb)
        saps.append(sap)
    for k in range(10):
        for _ in range(10):
            for prb, sap in zip(prbs, saps):
                prb.step(sap, 1e-4)
        st = []
        for prb in prbs:
            p = prb.ptrs()
            st.append([d2h(x, (nb, 4), np.float32) for x in (p.rb_pos, p.rb_quat, p.rb_linvel,
                                                              p.rb_angvel)] + [p.num_pairs])
        assert st[0][4] == st[1][4], k
        for a, b in zip(st[0][:4], st[1][:4]):
            assert np.isfinite(a).all() and np.array_equal(a, b), k
    assert st[0][4] > 0  # bodies were in contact


def _drift(n, seed, steps, jumpers):
    """n random AABBs drifting by small steps, `jumpers` of them moved far
    each step."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 40.0, (n, 3)).astype(np.float32)
    h = rng.uniform(0.1, 0.6, (n, 3)).astype(np.float32)
    for _ in range(steps):
        c = c + rng.normal(0.0, 0.05, (n, 3)).astype(np.float32)
        if jumpers:
            j = rng.choice(n, min(n, jumpers), replace=False)
            c[j] = rng.uniform(0.0, 40.0, (len(j), 3)).astype(np.float32)
        yield np.concatenate([c - h, c + h], 1).astype(np.float32)


# sizes around one block of the previous order (2,048 entries; two per box
# with slabs)
@pytest.mark.parametrize("n,slabs", [(1, True), (700, True), (1024, True), (1025, True),
                                     (2048, False), (2049, False), (5000, True)])
def test_coherent_small_and_ragged_sizes(dev, n, slabs):
    import torch
    flags = G.FLAG_SLABS if slabs else 0
    coh = G.Sap(axis_mode=G.AXIS_AUTO, flags=flags | G.FLAG_COHERENT)
    for t, boxes in enumerate(_drift(n, n, 6, 3)):
        x = to_dev(boxes, dev)
        ref, _, _ = G.Sap(axis_mode=G.AXIS_AUTO, flags=flags).collide(x)
        got, _, _ = coh.collide(x)
        assert torch.equal(got, ref), t
        if t == 5:
            assert np.array_equal(pairs_np(got), O.brute(boxes)), t


@pytest.mark.parametrize("slabs", [True, False])
def test_coherent_side_list_overflow_falls_back(dev, slabs):
    # 8,000 of 60,000 boxes jump each step: more far movers than the side
    # list holds, so the device runs the regular sort; pairs stay equal
    import torch
    flags = G.FLAG_SLABS if slabs else 0
    coh = G.Sap(axis_mode=G.AXIS_AUTO, flags=flags | G.FLAG_COHERENT)
    for t, boxes in enumerate(_drift(60_000, 7, 5, 8000)):
        x = to_dev(boxes, dev)
        ref, _, _ = G.Sap(axis_mode=G.AXIS_AUTO, flags=flags).collide(x)
        got, _, _ = coh.collide(x)
        assert torch.equal(got, ref), t
