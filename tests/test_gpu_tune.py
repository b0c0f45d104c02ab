"""Autotuner and launch configurations: every digit width / tile / grid /
sweep variant gives the identical pair list; the tuner returns a valid
configuration, keeps it, and round-trips through its cache file."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import pairs_np, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sort_bits,pair_bits", [(6, 6), (7, 7), (8, 8), (6, 8), (8, 7)])
@pytest.mark.parametrize("slabs", [0, 1])
def test_digit_widths_same_pairs(dev, sort_bits, pair_bits, slabs):
    b = W.small(20000, seed=11)
    ref = O.brute(b)
    for sort_items, pair_items, grid, chunk, reg in ((4, 8, 512, 0, 1), (12, 16, 4096, 16, 8),
                                                     (16, 8, 256, 64, 8), (8, 16, 1024, 4, 1)):
        s = G.Sap()
        s.set_tune(sort_items=sort_items, sort_bits=sort_bits, pair_items=pair_items,
                   pair_bits=pair_bits, pair_grid=grid, slabs=slabs, sweep_chunk=chunk,
                   pair_regions=reg)
        pairs, _, _ = s.collide(to_dev(b, dev))
        assert np.array_equal(pairs_np(pairs), ref), (sort_items, pair_items, grid)


def test_digit_widths_batched(dev):
    import torch
    boxes, off = W.scenes(24, seed=3, lo=500, hi=6000)
    bd = to_dev(boxes, dev)
    outs = []
    for bits in (6, 7, 8):
        s = G.Sap()
        s.set_tune(sort_bits=bits, pair_bits=bits, sort_items=12, pair_items=8,
                   pair_regions=8 if bits == 7 else 1)
        P, _ = s.collide_batched(bd, off, None, None)
        pairs = torch.empty((P, 2), dtype=torch.int32, device=dev)
        po = torch.empty(len(off), dtype=torch.int64, device=dev)
        s.collide_batched(bd, off, pairs, po)
        outs.append((pairs.cpu(), po.cpu()))
    for p, o in outs[1:]:
        assert torch.equal(p, outs[0][0]) and torch.equal(o, outs[0][1])


def test_set_tune_rejects_invalid(dev):
    s = G.Sap()
    for bad in (dict(sort_items=5), dict(sort_bits=5), dict(pair_bits=9), dict(pair_items=12),
                dict(pair_grid=100), dict(slabs=2), dict(sweep_chunk=5), dict(pair_regions=4),
                dict(pair_local=3)):
        with pytest.raises(G.GsapError) as e:
            s.set_tune(**bad)
        assert e.value.code == G.ERR_INVALID_ARG
    assert s.tune["sort_items"] == 8  # unchanged


def test_autotune_single_and_cache(dev, tmp_path):
    import torch
    b = W.aniso(300_000, seed=2)
    bd = to_dev(b, dev)
    ref, _, _ = G.Sap().collide(bd)
    cache = str(tmp_path / "tune.txt")
    s = G.Sap()
    t = s.autotune(bd, cache=cache)
    assert t["from_cache"] == 0 and t["ms"] > 0
    assert t["sort_items"] in (4, 8, 12, 16) and 6 <= t["sort_bits"] <= 8
    assert s.tune["pair_grid"] == t["pair_grid"]
    got, _, _ = s.collide(bd)
    assert torch.equal(got, ref)
    s2 = G.Sap()
    t2 = s2.autotune(bd, cache=cache)
    assert t2["from_cache"] == 1
    for k in ("sort_items", "sort_bits", "pair_items", "pair_bits", "pair_grid", "slabs",
              "sweep_chunk", "pair_regions", "pair_local"):
        assert t2[k] == t[k]
    with open(cache) as f:
        assert len(f.read().splitlines()) == 1


def test_autotune_batched_and_grouped(dev):
    import torch
    boxes, off = W.scenes(48, seed=8, lo=1000, hi=20000)
    bd = to_dev(boxes, dev)
    base = G.Sap()
    P, _ = base.collide_batched(bd, off, None, None)
    p0 = torch.empty((P, 2), dtype=torch.int32, device=dev)
    o0 = torch.empty(len(off), dtype=torch.int64, device=dev)
    base.collide_batched(bd, off, p0, o0)
    s = G.Sap()
    t = s.autotune(bd, scene_offsets=off)  # may pick the per-scene slab sweep
    assert t["from_cache"] == 0
    p1 = torch.empty_like(p0)
    o1 = torch.empty_like(o0)
    s.collide_batched(bd, off, p1, o1)
    assert torch.equal(p0, p1) and torch.equal(o0, o1)
    # a grouped, symmetric (particle-style) handle tunes with its group array
    sph = np.concatenate([np.random.default_rng(1).uniform(0, 30, (20000, 3)),
                          np.full((20000, 1), 0.5)], 1).astype(np.float32)
    g = to_dev((np.arange(20000) // 27).astype(np.int32), dev)
    h = G.Sap(input_kind=G.INPUT_SPHERE, output_kind=G.OUT_SYM_CSR, flags=G.FLAG_SKIP_SAME_GROUP)
    ref, ent, _ = h.collide(to_dev(sph, dev), group=g)
    h.autotune(to_dev(sph, dev), group=g)
    got, ent2, _ = h.collide(to_dev(sph, dev), group=g)
    assert torch.equal(got, ref) and torch.equal(ent, ent2)


def test_digit_widths_reduce_then_scan(dev):
    # Above 128 sort tiles the object sort and (up to 4,096 tiles of the
    # capacity) the pair sort's look-back passes run by reduce-then-scan:
    # every digit width gives the identical list, one scene and batched.
    import torch
    b = W.aniso(300_000, seed=5, sigma=(20.0, 10.0, 5.0))
    bd = to_dev(b, dev)
    ref, _, _ = G.Sap().collide(bd)
    # 2,048-pair tiles of the capacity (the pair count plus padding) within
    # [128, 4096]: the pair passes take the reduce-then-scan path too
    assert 128 * 2048 <= ref.shape[0] <= 4000 * 2048
    for sort_bits, pair_bits, regions, slabs in ((6, 6, 1, 1), (7, 7, 8, 1), (6, 8, 8, 0), (8, 7, 1, 0)):
        s = G.Sap()
        s.set_tune(sort_items=8, sort_bits=sort_bits, pair_items=8, pair_bits=pair_bits, pair_local=0,
                   pair_regions=regions, slabs=slabs)
        got, _, _ = s.collide(bd)
        assert torch.equal(got, ref), (sort_bits, pair_bits, regions, slabs)
    # a capacity of four times the pairs (the passes' grid follows the
    # capacity; the tiles past the pair count return at once)
    s = G.Sap()
    s.set_tune(sort_items=8, pair_items=16, pair_local=0, pair_regions=1)
    big = torch.full((4 * ref.shape[0], 2), -1, dtype=torch.int32, device=dev)
    P, _, _ = s.collide_into(bd, big)
    assert P == ref.shape[0] and torch.equal(big[:P], ref)
    assert bool((big[P:] == -1).all())
    boxes, off = W.scenes(48, seed=9, lo=4000, hi=30000)
    bd = to_dev(boxes, dev)
    outs = []
    for bits in (6, 7, 8):
        s = G.Sap()
        s.set_tune(sort_bits=bits, pair_bits=bits, sort_items=8, pair_items=8, pair_local=0,
                   pair_regions=8, slabs=1)
        P, _ = s.collide_batched(bd, off, None, None)
        pairs = torch.empty((P, 2), dtype=torch.int32, device=dev)
        po = torch.empty(len(off), dtype=torch.int64, device=dev)
        s.collide_batched(bd, off, pairs, po)
        outs.append((pairs.cpu(), po.cpu()))
    for p, o in outs[1:]:
        assert torch.equal(p, outs[0][0]) and torch.equal(o, outs[0][1])
