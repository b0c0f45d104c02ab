"""Pair sort by rows (gsap_tune.pair_local = 2): a counting sort on the first
index, then every row ordered by the second index (rows of up to 64 pairs by
comparison counts, longer rows by the local stage's big-window kernel: an
LDS sort, or LSD passes through global memory past its LDS capacity).  The
canonical pair list, the symmetric output's row starts and the batched scene
offsets must be bit-identical to brute force and to the whole-key sort,
across object counts (8 to 24 index bits; below 8 the whole-key passes run),
row shapes, pair regions, slabs, a capacity overflow and rigid-body steps."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import d2h, pairs_np, to_dev

pytestmark = pytest.mark.gpu


def _pairs(b, dev, **tune):
    s = G.Sap()
    s.set_tune(**tune)
    pairs, _, _ = s.collide(to_dev(b, dev))
    return pairs_np(pairs)


# n -> index bits: 100 -> 7 (whole-key passes), 200 -> 8, 3000 -> 12, 20000 -> 15
@pytest.mark.parametrize("n", [100, 200, 3000, 20000])
@pytest.mark.parametrize("regions", [1, 8])
def test_rows_equals_brute(dev, n, regions):
    b = W.scene(n, np.random.default_rng(n + 1), "cluster")
    ref = O.brute(b)
    for slabs in (0, 1):
        got = _pairs(b, dev, pair_local=2, pair_regions=regions, slabs=slabs)
        assert np.array_equal(got, ref), slabs


def test_rows_long_rows_equal_brute(dev):
    # Four boxes span the scene (rows of ~20,000 pairs: past the big-window
    # kernel's LDS sort, so LSD passes through global memory), and 40 boxes
    # of a tenth of the scene's width (rows of a few hundred: its LDS sort),
    # among 20,000 small ones (rows of a few pairs: comparison counts).
    r = np.random.default_rng(21)
    b = W.scene(20000, r, "uniform").copy()
    lo, hi = b[:, :3].min(0), b[:, 3:].max(0)
    for i in (0, 1, 2, 19999):
        b[i, :3] = lo - 1.0
        b[i, 3:] = hi + 1.0
    w = 0.1 * (hi - lo)
    for i in range(100, 140):
        c = r.uniform(lo, hi)
        b[i, :3] = c - w
        b[i, 3:] = c + w
    ref = O.brute(b)
    rows = np.bincount(ref[:, 0], minlength=len(b))
    assert rows.max() > 16384 and ((rows > 64) & (rows < 4096)).sum() >= 10
    for slabs in (0, 1):
        for regions in (1, 8):
            got = _pairs(b, dev, pair_local=2, slabs=slabs, pair_regions=regions)
            assert np.array_equal(got, ref), (slabs, regions)


@pytest.mark.parametrize("n,sigma", [(300_000, (20.0, 5.0, 1.0)),      # 19 index bits
                                     (1_100_000, (60.0, 30.0, 20.0))])  # 21 index bits
def test_rows_equals_whole_key_sort(dev, n, sigma):
    import torch
    b = W.aniso(n, seed=3, sigma=sigma)
    bd = to_dev(b, dev)
    ref, _, _ = G.Sap().collide(bd)
    for slabs in (0, 1):
        for regions in (1, 8):
            s = G.Sap()
            s.set_tune(pair_local=2, slabs=slabs, pair_regions=regions)
            got, _, _ = s.collide(bd)
            assert torch.equal(got, ref), (slabs, regions)


def test_rows_sparse_700k(dev):
    # 700,000 tiny boxes: about one pair per box, the case the row sort is for
    import torch
    r = np.random.default_rng(9)
    c = r.uniform(0.0, 400.0, (700_000, 3))
    h = r.uniform(0.2, 0.6, (700_000, 3))
    bd = to_dev(W.boxes_from(c, h), dev)
    ref, _, _ = G.Sap().collide(bd)
    assert 0 < ref.shape[0] < 4 * 700_000
    for regions in (1, 8):
        s = G.Sap()
        s.set_tune(pair_local=2, pair_regions=regions)
        got, _, _ = s.collide(bd)
        assert torch.equal(got, ref), regions


def test_rows_sym_csr_grouped(dev):
    # symmetric output: directed pairs and the row starts of every non-empty row
    import torch
    sph = np.concatenate([np.random.default_rng(1).uniform(0, 30, (20000, 3)),
                          np.full((20000, 1), 0.5)], 1).astype(np.float32)
    g = to_dev((np.arange(20000) // 27).astype(np.int32), dev)
    out = []
    for loc in (0, 2):
        h = G.Sap(input_kind=G.INPUT_SPHERE, output_kind=G.OUT_SYM_CSR,
                  flags=G.FLAG_SKIP_SAME_GROUP)
        h.set_tune(pair_local=loc)
        out.append(h.collide(to_dev(sph, dev), group=g))
    assert out[0][0].shape[0] > 0
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


def test_rows_batched(dev):
    import torch
    boxes, off = W.scenes(48, seed=8, lo=1000, hi=20000)
    bd = to_dev(boxes, dev)
    res = []
    for loc, slabs in ((0, 0), (2, 0), (2, 1)):
        s = G.Sap()
        s.set_tune(pair_local=loc, pair_regions=8, slabs=slabs)
        P, _ = s.collide_batched(bd, off, None, None)
        pairs = torch.empty((P, 2), dtype=torch.int32, device=dev)
        po = torch.empty(len(off), dtype=torch.int64, device=dev)
        for _ in range(2):  # capture, then replay
            s.collide_batched(bd, off, pairs, po)
            res.append((pairs.clone(), po.clone()))
    for p, o in res[1:]:
        assert torch.equal(p, res[0][0]) and torch.equal(o, res[0][1])


@pytest.mark.parametrize("regions", [1, 8])
def test_rows_capacity_error_then_retry(dev, regions):
    # an overflowing update sorts nothing (its regions are emptied); the row
    # counters stay zero, so the retry into a large enough buffer is exact
    import torch
    b = W.small()
    ref = O.brute(b)
    s = G.Sap()
    s.set_tune(pair_local=2, pair_regions=regions)
    bd = to_dev(b, dev)
    small_buf = torch.empty((100, 2), dtype=torch.int32, device=dev)
    with pytest.raises(G.GsapError) as e:
        s.collide_into(bd, small_buf)
    assert e.value.code == G.ERR_CAPACITY
    big = torch.empty((len(ref), 2), dtype=torch.int32, device=dev)
    for _ in range(2):
        P, _, _ = s.collide_into(bd, big)
        assert P == len(ref) and np.array_equal(pairs_np(big), ref)


def test_rows_particle_trajectories_bit_identical(dev):
    """50 rigid-body steps (400 cubes, slab broadphase on particle AABBs,
    same-body pairs skipped, symmetric CSR) with the row sort and with the
    whole-key sort: body states identical bit for bit."""
    nb, box = 400, 30.0
    sc = W.bodies(nb, seed=2, box=box, height=10.0)
    flags = G.PRB_ABS_SPRING | G.PRB_WALLS | G.PRB_SKIP_SAME_BODY
    prbs, saps = [], []
    for loc in (0, 2):
        prb = G.ParticleRigidBody(nb, nb * 27, nb * 27 * 40)
        prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"],
                        sc["pstart"], sc["npart"], sc["inv_mass"], sc["inv_inertia"])
        prb.set_particle_data(sc["ref_pos"], sc["ref_mass"], sc["max_per_shape"])
        prb.set_particle_shape_inv_inertia(sc["shape_inv_inertia"])
        prb.set_physics(gravity=(0, -9.81, 0), k_abs=1e4, c_abs=50.0, floor_scale=1.0,
                        wall_lo=(0, 0), wall_hi=(box, box), flags=flags,
                        broadphase_kind=G.INPUT_AABB)
        sap = G.Sap(input_kind=G.INPUT_AABB, output_kind=G.OUT_SYM_CSR, axis_mode=G.AXIS_AUTO,
                    flags=G.FLAG_SKIP_SAME_GROUP | G.FLAG_SLABS)
        sap.set_tune(pair_local=loc, slabs=1, pair_regions=8)
        prbs.append(prb)
        saps.append(sap)
    for k in range(5):
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


def test_rows_sym_csr_capacity_error_then_retry(dev):
    # the sweep counts the rows as it writes: after an overflowing update its
    # counters are cleared and the row starts are not written; the retry
    # matches the whole-key sort, pairs and row starts
    import torch
    sph = np.concatenate([np.random.default_rng(4).uniform(0, 30, (20000, 3)),
                          np.full((20000, 1), 0.5)], 1).astype(np.float32)
    g = to_dev((np.arange(20000) // 27).astype(np.int32), dev)
    x = to_dev(sph, dev)
    ref = G.Sap(input_kind=G.INPUT_SPHERE, output_kind=G.OUT_SYM_CSR,
                flags=G.FLAG_SKIP_SAME_GROUP)
    ref.set_tune(pair_local=0)
    rp, re_, _ = ref.collide(x, group=g)
    h = G.Sap(input_kind=G.INPUT_SPHERE, output_kind=G.OUT_SYM_CSR, flags=G.FLAG_SKIP_SAME_GROUP)
    h.set_tune(pair_local=2, pair_regions=8)
    small = torch.empty((rp.shape[0] // 3, 2), dtype=torch.int32, device=dev)
    ent = torch.empty(20000, dtype=torch.int32, device=dev)
    with pytest.raises(G.GsapError) as e:
        h.collide_into(x, small, group=g, entries=ent)
    assert e.value.code == G.ERR_CAPACITY
    for _ in range(2):
        got, ge, _ = h.collide(x, group=g)
        assert torch.equal(got, rp) and torch.equal(ge, re_)
