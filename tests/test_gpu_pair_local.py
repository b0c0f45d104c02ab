"""Pair sort with the local stage: onesweep passes over the high key bits only,
then every run of equal high bits ("bucket") sorted by the low bits in LDS, or
through global memory when a window's buckets do not fit.  The canonical pair
list must be bit-identical to the brute-force oracle and to the whole-key
sort, across key widths (local bits L = 8 ... 24, one to three onesweep
passes), bucket shapes (one bucket larger than a workgroup's LDS sort, windows
with more buckets than a 32-bit LDS key can rank), output kinds, batched
scenes, pair regions with padding, a capacity overflow, and the two-pass sort
whose second pass runs in per-XCD parts (34- to 40-bit keys, all keys in one
part, sparse lists whose tiles span many first-digit ranges, no pairs)."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import pairs_np, to_dev

pytestmark = pytest.mark.gpu


def _pairs(b, dev, **tune):
    s = G.Sap()
    s.set_tune(**tune)
    pairs, _, _ = s.collide(to_dev(b, dev))
    return pairs_np(pairs)


# n -> key bits K = 2 * bits(n): 200 -> 16 (L = 8), 3000 -> 24 (L = 16),
# 20000 -> 30 (L = 22, one onesweep pass)
@pytest.mark.parametrize("n", [200, 3000, 20000])
@pytest.mark.parametrize("regions", [1, 8])
def test_local_equals_brute(dev, n, regions):
    b = W.scene(n, np.random.default_rng(n), "uniform")
    ref = O.brute(b)
    for items, grid in ((8, 512), (16, 4096)):
        got = _pairs(b, dev, pair_local=1, pair_items=items, pair_grid=grid,
                     pair_regions=regions)
        assert np.array_equal(got, ref), (items, grid)


def test_local_oversized_bucket_equals_brute(dev):
    # boxes 0..2 span the whole scene: their buckets (ids 0..2 are the smaller
    # index of ~20000 pairs each) exceed the LDS sort and take the global path
    b = W.scene(20000, np.random.default_rng(5), "cluster").copy()
    lo, hi = b[:, :3].min(0), b[:, 3:].max(0)
    for i in (0, 1, 2, 19999):
        b[i, :3] = lo - 1.0
        b[i, 3:] = hi + 1.0
    ref = O.brute(b)
    for slabs in (0, 1):
        for regions in (1, 8):
            got = _pairs(b, dev, pair_local=1, slabs=slabs, pair_regions=regions)
            assert np.array_equal(got, ref), (slabs, regions)


@pytest.mark.parametrize("n,sigma", [(300_000, (20.0, 5.0, 1.0)),    # K = 38, two passes, L = 22
                                     (700_000, (20.0, 5.0, 1.0)),    # K = 40, L = 24
                                     (1_100_000, (60.0, 30.0, 20.0))])  # K = 42, three passes
def test_local_equals_whole_key_sort(dev, n, sigma):
    import torch
    b = W.aniso(n, seed=3, sigma=sigma)
    bd = to_dev(b, dev)
    ref, _, _ = G.Sap().collide(bd)
    for slabs in (0, 1):
        for regions in (1, 8):
            s = G.Sap()
            s.set_tune(pair_local=1, slabs=slabs, pair_regions=regions)
            got, _, _ = s.collide(bd)
            assert torch.equal(got, ref), (slabs, regions)


def test_local_sparse_windows_many_buckets(dev):
    # 700,000 tiny boxes (K = 40, L = 24): about one pair per box, so a window
    # of 2048 pairs spans more buckets than a 32-bit LDS key can rank (256)
    import torch
    r = np.random.default_rng(9)
    c = r.uniform(0.0, 400.0, (700_000, 3))
    h = r.uniform(0.2, 0.6, (700_000, 3))
    bd = to_dev(W.boxes_from(c, h), dev)
    ref, _, _ = G.Sap().collide(bd)
    assert 0 < ref.shape[0] < 4 * 700_000
    s = G.Sap()
    s.set_tune(pair_local=1)
    got, _, _ = s.collide(bd)
    assert torch.equal(got, ref)


def test_local_sym_csr_grouped(dev):
    import torch
    sph = np.concatenate([np.random.default_rng(1).uniform(0, 30, (20000, 3)),
                          np.full((20000, 1), 0.5)], 1).astype(np.float32)
    g = to_dev((np.arange(20000) // 27).astype(np.int32), dev)
    out = []
    for loc in (0, 1):
        h = G.Sap(input_kind=G.INPUT_SPHERE, output_kind=G.OUT_SYM_CSR,
                  flags=G.FLAG_SKIP_SAME_GROUP)
        h.set_tune(pair_local=loc)
        out.append(h.collide(to_dev(sph, dev), group=g))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


def test_local_batched(dev):
    import torch
    boxes, off = W.scenes(48, seed=8, lo=1000, hi=20000)
    bd = to_dev(boxes, dev)
    res = []
    for loc in (0, 1):
        s = G.Sap()
        s.set_tune(pair_local=loc, pair_regions=8)
        P, _ = s.collide_batched(bd, off, None, None)
        pairs = torch.empty((P, 2), dtype=torch.int32, device=dev)
        po = torch.empty(len(off), dtype=torch.int64, device=dev)
        for _ in range(2):  # capture, then replay
            s.collide_batched(bd, off, pairs, po)
            res.append((pairs.clone(), po.clone()))
    for p, o in res[1:]:
        assert torch.equal(p, res[0][0]) and torch.equal(o, res[0][1])


@pytest.mark.parametrize("loc,regions", [(0, 8), (1, 1), (1, 8)])
def test_local_capacity_error_then_retry(dev, loc, regions):
    import torch
    b = W.small()
    ref = O.brute(b)
    s = G.Sap()
    s.set_tune(pair_local=loc, pair_regions=regions)
    bd = to_dev(b, dev)
    small_buf = torch.empty((100, 2), dtype=torch.int32, device=dev)  # about 875 pairs
    with pytest.raises(G.GsapError) as e:
        s.collide_into(bd, small_buf)
    assert e.value.code == G.ERR_CAPACITY
    big = torch.empty((len(ref), 2), dtype=torch.int32, device=dev)
    P, _, _ = s.collide_into(bd, big)
    assert P == len(ref) and np.array_equal(pairs_np(big), ref)


@pytest.mark.parametrize("regions", [1, 8])
def test_local_two_pass_capacity_error_then_retry(dev, regions):
    # K = 38 (two onesweep passes; with 8 regions the first one claims its
    # slots per region): an overflowing call reports the capacity and the
    # retry into a large enough buffer is exact
    import torch
    b = W.aniso(300_000, seed=4, sigma=(20.0, 5.0, 1.0))
    bd = to_dev(b, dev)
    ref, _, _ = G.Sap().collide(bd)
    s = G.Sap()
    s.set_tune(pair_local=1, pair_regions=regions)
    small = torch.empty((ref.shape[0] // 3, 2), dtype=torch.int32, device=dev)
    with pytest.raises(G.GsapError) as e:
        s.collide_into(bd, small)
    assert e.value.code == G.ERR_CAPACITY
    big = torch.empty_like(ref)
    P, _, _ = s.collide_into(bd, big)
    assert P == ref.shape[0] and torch.equal(big, ref)


def test_local_40bit_keys_big_windows(dev):
    # 600,000 boxes: K = 40 (a per-region first pass, then one onesweep pass
    # before the local stage).  Three boxes span the scene: their buckets hold
    # ~600,000 pairs, far over the LDS sorts, so they take the global-memory
    # path.  Reference: the whole-key sort.
    import torch
    b = W.scene(600_000, np.random.default_rng(12), "uniform").copy()
    lo, hi = b[:, :3].min(0), b[:, 3:].max(0)
    for i in (0, 1, 2):
        b[i, :3] = lo - 1.0
        b[i, 3:] = hi + 1.0
    bd = to_dev(b, dev)
    s0 = G.Sap()
    s0.set_tune(pair_local=0)
    ref, _, _ = s0.collide(bd)
    for slabs in (0, 1):
        s = G.Sap()
        s.set_tune(pair_local=1, pair_regions=8, slabs=slabs)
        got, _, _ = s.collide(bd)
        assert torch.equal(got, ref), slabs


def test_local_40bit_keys_capacity_error_then_retry(dev):
    # an overflowing update skips the pair sort (its digit counts include the
    # pairs that were not written); the retry is exact
    import torch
    b = W.aniso(700_000, seed=6, sigma=(20.0, 5.0, 1.0))
    bd = to_dev(b, dev)
    ref, _, _ = G.Sap().collide(bd)
    s = G.Sap()
    s.set_tune(pair_local=1, pair_regions=8)
    small = torch.empty((ref.shape[0] // 2, 2), dtype=torch.int32, device=dev)
    with pytest.raises(G.GsapError) as e:
        s.collide_into(bd, small)
    assert e.value.code == G.ERR_CAPACITY
    big = torch.empty_like(ref)
    P, _, _ = s.collide_into(bd, big)
    assert P == ref.shape[0] and torch.equal(big, ref)


# Two onesweep passes with 8 pair regions: the second runs in 8 parts by the
# first digit's top 3 bits.  n -> K: 100,000 -> 34, 200,000 -> 36.
@pytest.mark.parametrize("n", [100_000, 200_000])
def test_local_two_pass_parts_key_widths(dev, n):
    b = W.aniso(n, seed=n % 97, sigma=(12.0, 4.0, 1.0))
    bd = to_dev(b, dev)
    s0 = G.Sap()
    s0.set_tune(pair_local=0)
    ref, _, _ = s0.collide(bd)
    for slabs in (0, 1):
        s = G.Sap()
        s.set_tune(pair_local=1, pair_regions=8, slabs=slabs)
        got, _, _ = s.collide(bd)
        assert np.array_equal(pairs_np(got), pairs_np(ref)), slabs


def test_local_two_pass_one_part(dev):
    # 256 identical boxes overlap each other and nothing else; 300,000 small
    # boxes on a grid overlap nothing.  Every pair has i < j < 256, so all
    # keys fall in the first part (seven parts empty); 32,640 pairs.
    n = 300_000
    g = np.arange(n, dtype=np.float64)
    lo = np.stack([(g % 100) * 10.0, ((g // 100) % 100) * 10.0, (g // 10000) * 10.0], 1)
    b = np.concatenate([lo, lo + 1.0], 1).astype(np.float32)
    b[:256, :3] = -50.0
    b[:256, 3:] = -40.0
    i, j = np.triu_indices(256, 1)
    expect = np.stack([i, j], 1).astype(np.uint32)
    for slabs in (0, 1):
        s = G.Sap()
        s.set_tune(pair_local=1, pair_regions=8, slabs=slabs)
        got, _, _ = s.collide(to_dev(b, dev))
        assert np.array_equal(pairs_np(got), expect), slabs


@pytest.mark.parametrize("n", [300_000, 1_000_000])  # K = 38 and 40
def test_local_two_pass_sparse_and_empty(dev, n):
    # Boxes on a grid overlap nothing; then box i + 1 is moved onto box i for
    # a few thousand even i spread over the whole index range: pairs (i, i+1)
    # only, a few per first-digit range, so one tile of the second pass
    # spans many first-digit ranges (and some ranges are empty).  The grid
    # alone gives no pairs at all.
    import torch
    g = np.arange(n, dtype=np.float64)
    lo = np.stack([(g % 100) * 10.0, ((g // 100) % 100) * 10.0, (g // 10000) * 10.0], 1)
    b = np.concatenate([lo, lo + 1.0], 1).astype(np.float32)
    for slabs in (0, 1):
        s = G.Sap()
        s.set_tune(pair_local=1, pair_regions=8, slabs=slabs)
        got, _, _ = s.collide(to_dev(b, dev))
        assert got.shape[0] == 0, slabs
    i = np.sort(np.random.default_rng(n).choice(n // 2 - 1, 3000, replace=False)) * 2
    b[i + 1] = b[i] + np.float32(0.5)
    expect = np.stack([i, i + 1], 1).astype(np.uint32)
    for slabs in (0, 1):
        for items in (8, 16):
            s = G.Sap()
            s.set_tune(pair_local=1, pair_regions=8, slabs=slabs, pair_items=items)
            got, _, _ = s.collide(to_dev(b, dev))
            assert np.array_equal(pairs_np(got), expect), (slabs, items)


def test_local_two_pass_random_sizes_and_tunes(dev):
    # Seeded draws over the two-pass range (17 to 20 index bits): scene kind,
    # size, tile width, grid and sweep kind at random; the canonical list must
    # equal the whole-key sort's every time.
    import torch
    r = np.random.default_rng(2024)
    for case in range(8):
        n = int(r.integers(70_000, 1_048_577))
        kind = ("uniform", "cluster", "aniso")[case % 3]
        if kind == "aniso":
            b = W.aniso(n, seed=int(r.integers(1 << 30)), sigma=(20.0, 5.0, 1.0))
        else:
            b = W.scene(n, np.random.default_rng(int(r.integers(1 << 30))), kind)
        bd = to_dev(b, dev)
        s0 = G.Sap()
        s0.set_tune(pair_local=0)
        ref, _, _ = s0.collide(bd)
        tune = dict(pair_local=1, pair_regions=8, slabs=int(r.integers(2)),
                    pair_items=int(r.choice([8, 16])), pair_grid=int(r.choice([512, 1024, 2048, 4096])))
        s = G.Sap()
        s.set_tune(**tune)
        got, _, _ = s.collide(bd)
        assert got.shape == ref.shape and torch.equal(got, ref), (case, n, kind, tune)
