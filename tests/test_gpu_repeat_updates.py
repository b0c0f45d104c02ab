"""Single-scene updates repeated into the same buffers: every update must
equal brute force (or the default handle) on the boxes of that call -- the
box contents change between calls, with other work (count-only updates,
other sizes, a retune) in between, so no state of one update may leak into
the next."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import pairs_np, to_dev

pytestmark = pytest.mark.gpu


def _moved(b, rng, scale=0.3):
    d = rng.normal(0.0, scale, (len(b), 3)).astype(np.float32)
    return np.concatenate([b[:, :3] + d, b[:, 3:] + d], 1).astype(np.float32)


@pytest.mark.parametrize("slabs", [0, 1])
def test_replays_equal_brute(dev, slabs):
    import torch
    rng = np.random.default_rng(3 + slabs)
    b = W.small(6000, seed=5)
    x = to_dev(b, dev)
    s = G.Sap(axis_mode=G.AXIS_AUTO, flags=G.FLAG_SLABS if slabs else 0)
    cap = 10 * len(O.brute(b)) + 10000
    out = torch.empty((cap, 2), dtype=torch.int32, device=dev)
    for t in range(6):
        bt = _moved(b, rng)
        x.copy_(torch.from_numpy(bt))  # same device buffer, new boxes
        if t == 3:
            s.count(x)  # eager count-only update between replays
            s.collide(to_dev(bt[:4000].copy(), dev))  # another size, eager
        P, _, _ = s.collide_into(x, out)
        assert np.array_equal(pairs_np(out[:P]), O.brute(bt)), t


def test_replays_after_retune_and_sym_rows(dev):
    import torch
    rng = np.random.default_rng(11)
    sph = np.concatenate([rng.uniform(0, 25, (8000, 3)), np.full((8000, 1), 0.5)], 1).astype(np.float32)
    g = to_dev((np.arange(8000) // 27).astype(np.int32), dev)
    x = to_dev(sph, dev)
    ref = G.Sap(input_kind=G.INPUT_SPHERE, output_kind=G.OUT_SYM_CSR, flags=G.FLAG_SKIP_SAME_GROUP)
    h = G.Sap(input_kind=G.INPUT_SPHERE, output_kind=G.OUT_SYM_CSR, flags=G.FLAG_SKIP_SAME_GROUP)
    pairs = torch.empty((400_000, 2), dtype=torch.int32, device=dev)
    ent = torch.empty(8000, dtype=torch.int32, device=dev)
    for t in range(7):
        if t == 4:
            h.set_tune(pair_local=2, pair_regions=8)  # a new configuration mid-sequence
        st = sph.copy()
        st[:, :3] += rng.normal(0.0, 0.2, (8000, 3)).astype(np.float32)
        x.copy_(torch.from_numpy(st))
        P, E, _ = h.collide_into(x, pairs, entries=ent, group=g)
        rp, re_, _ = ref.collide(x, group=g)
        assert P == rp.shape[0] and torch.equal(pairs[:P], rp), t
        assert E == re_.shape[0] and torch.equal(ent[:E], re_), t
