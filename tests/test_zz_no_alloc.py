"""The allocation-free entry point (gsap_sap_collide_noalloc) against the
existing stream-ordered one (gsap_sap_collide_async) on the same handle
configuration: identical counts, pairs and row starts, on two calls in a row,
the second with a smaller input; sizes beyond the static scratch are
rejected before any device work."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from tests.gpu_util import to_dev

pytestmark = pytest.mark.gpu


def _both(sap_a, sap_b, x, cap, sym, dev):
    import torch
    n = x.shape[0]
    res = []
    for s, noalloc in ((sap_a, False), (sap_b, True)):
        pairs = torch.full((cap, 2), -1, dtype=torch.int32, device=dev)
        counts = torch.zeros(8, dtype=torch.int64, device=dev)
        ent = torch.full((n,), -1, dtype=torch.int32, device=dev) if sym else None
        if noalloc:
            s.collide_noalloc(x, pairs, counts, entries=ent)
        else:
            s.collide_async(x, pairs, counts, entries=ent)
        torch.cuda.synchronize()
        c = counts.cpu().numpy()
        res.append((c, pairs.cpu().numpy(), None if ent is None else ent.cpu().numpy()))
    return res


@pytest.mark.parametrize("case", ["aabb_canon", "sphere_sym_slabs"])
def test_noalloc_equals_async(dev, case):
    if case == "aabb_canon":
        kw = dict(input_kind=G.INPUT_AABB, axis_mode=G.AXIS_PCA, output_kind=G.OUT_CANON)
        b0 = W.small(60_000, seed=31)
        c, h = 0.3 * (b0[:, :3] + b0[:, 3:]), 0.5 * (b0[:, 3:] - b0[:, :3])  # centres x0.6: denser
        big = np.concatenate([c - h, c + h], 1).astype(np.float32)
        small = big[:17_001].copy()
        sym = False
    else:
        kw = dict(input_kind=G.INPUT_SPHERE, axis_mode=G.AXIS_AUTO, output_kind=G.OUT_SYM_CSR,
                  flags=G.FLAG_SLABS)
        rng = np.random.default_rng(32)
        big = np.concatenate([rng.uniform(0, 40, (50_000, 3)), rng.uniform(0.2, 0.7, (50_000, 1))],
                             1).astype(np.float32)
        small = big[:9_999].copy()
        sym = True
    a = G.Sap(**kw)  # the existing entry point
    b = G.Sap(**kw)  # the allocation-free one
    for boxes in (big, small):
        x = to_dev(boxes, dev)
        P, _ = a.count(x)
        cap = P + 64
        assert b.query_workspace(len(boxes), cap) <= G.NOALLOC_SCRATCH_BYTES
        (ca, pa, ea), (cb, pb, eb) = _both(a, b, x, cap, sym, dev)
        assert ca[0] == P and ca[3] == 0
        assert np.array_equal(ca[:5], cb[:5])  # pairs, entries, sweep tests, flags, executed
        assert np.array_equal(pa, pb)          # every slot, the unused tail included
        if sym:
            ne = int(ca[1])
            assert ne > 0 and np.array_equal(ea[:ne], eb[:ne])
        assert P > 1000


def test_noalloc_rejects_beyond_scratch(dev):
    import torch
    s = G.Sap()
    x = to_dev(W.small(1000, seed=1), dev)
    cap = G.NOALLOC_SCRATCH_BYTES // 8  # pair buffers alone exceed the scratch
    pairs = torch.empty((16, 2), dtype=torch.int32, device=dev)
    counts = torch.zeros(8, dtype=torch.int64, device=dev)
    with pytest.raises(G.GsapError) as e:
        G.check(G.lib().gsap_sap_collide_noalloc(s.h, G._stream(), G._ptr(x), 1000, None,
                                                  G._ptr(pairs), cap, None, G._ptr(counts)))
    assert e.value.code == G.ERR_CAPACITY and "static scratch" in str(e.value)
    # the handle still works through the allocation-free entry point
    n_ok, _ = s.count(x)
    p2 = torch.empty((max(n_ok, 1), 2), dtype=torch.int32, device=dev)
    s.collide_noalloc(x, p2, counts)
    torch.cuda.synchronize()
    assert int(counts[0]) == n_ok and int(counts[3]) == 0
