"""Buffer lifetime of a SaP handle: the captured batched update after the
handle's buffers moved, and updates carved from a caller-provided workspace
(the scratch the reference's Sap::Sap allocates once, Sap.cpp:12-114)."""
import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import pairs_np, to_dev

pytestmark = pytest.mark.gpu


def _scene_refs(boxes, off):
    refs = [O.brute(boxes[off[k]:off[k + 1]]) + off[k] for k in range(len(off) - 1)]
    return refs, sum(len(r) for r in refs)


def _check_batched(s, bd, off, pairs, spo, refs, total):
    got_n, _ = s.collide_batched(bd, off, pairs, spo)
    assert got_n == total
    allp = pairs_np(pairs[:got_n])
    so = spo.cpu().numpy()
    for k in range(len(off) - 1):
        assert np.array_equal(allp[so[k]:so[k + 1]], refs[k]), k


def test_batched_replay_after_buffers_move(dev):
    """Batched (graph captured) -> a larger single-scene update on the same
    handle (every object and pair buffer is reallocated) -> the judge's axis
    tests on another scene table (scene tables rewritten) -> the same batched
    call again: it must not replay the graph recorded over the old buffers."""
    import torch
    boxes, off = W.scenes(16, seed=3, lo=300, hi=3000)
    bd = to_dev(boxes, dev)
    refs, total = _scene_refs(boxes, off)
    pairs = torch.empty((total + 16, 2), dtype=torch.int32, device=dev)
    spo = torch.empty(len(off), dtype=torch.int64, device=dev)
    s = G.Sap()
    for _ in range(2):  # capture, then replay
        _check_batched(s, bd, off, pairs, spo, refs, total)
    big = W.aniso(400_000, seed=5)
    bb = to_dev(big, dev)
    P, _ = s.count(bb)
    bp = torch.empty((P + 1000, 2), dtype=torch.int32, device=dev)
    n2, _, _ = s.collide_into(bb, bp)
    assert n2 == P
    small, soff = W.scenes(5, seed=4, lo=100, hi=2000)
    tests, _ = s.axis_tests_batched(to_dev(small, dev), soff, G.axis_candidates()[:3])
    assert int(tests.min()) >= 0
    for _ in range(2):  # recapture, then replay
        _check_batched(s, bd, off, pairs, spo, refs, total)


@pytest.mark.parametrize("slabs", [0, 1])
def test_caller_workspace_equals_internal(dev, slabs):
    import torch
    b = W.aniso(150_000, seed=8)
    ref = O.brute_or_sap(b, O.KIND_AABB, None)
    bd = to_dev(b, dev)
    s = G.Sap(axis_mode=G.AXIS_AUTO, flags=G.FLAG_SLABS if slabs else 0)
    if slabs:
        s.set_tune(slabs=1, pair_local=1, pair_regions=8, sweep_chunk=16)
    need = s.query_workspace(len(b), len(ref))
    assert need > 0
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    s.set_workspace(ws)
    pairs = torch.empty((len(ref), 2), dtype=torch.int32, device=dev)
    for _ in range(2):
        n, _, _ = s.collide_into(bd, pairs)
        assert n == len(ref) and np.array_equal(pairs_np(pairs[:n]), ref)
    # a smaller update keeps the carve
    sub = b[:50_000]
    rs = O.brute_or_sap(sub, O.KIND_AABB, None)
    n, _, _ = s.collide_into(to_dev(sub, dev), pairs)
    assert n == len(rs) and np.array_equal(pairs_np(pairs[:n]), rs)
    # one byte short of the requirement: a capacity error, nothing allocated
    s2 = G.Sap(axis_mode=G.AXIS_AUTO, flags=G.FLAG_SLABS if slabs else 0)
    if slabs:
        s2.set_tune(slabs=1, pair_local=1, pair_regions=8, sweep_chunk=16)
    assert s2.query_workspace(len(b), len(ref)) == need
    s2.set_workspace(ws[:need - 256])
    with pytest.raises(G.GsapError) as e:
        s2.collide_into(bd, pairs)
    assert e.value.code == G.ERR_CAPACITY and "workspace" in str(e.value)
    # back to internal allocation
    s2.set_workspace(None)
    n, _, _ = s2.collide_into(bd, pairs)
    assert n == len(ref) and np.array_equal(pairs_np(pairs[:n]), ref)


def test_batched_in_caller_workspace(dev):
    import torch
    boxes, off = W.scenes(12, seed=11, lo=500, hi=4000)
    bd = to_dev(boxes, dev)
    refs, total = _scene_refs(boxes, off)
    s = G.Sap()
    ws = torch.empty(64 << 20, dtype=torch.uint8, device=dev)
    s.set_workspace(ws)
    pairs = torch.empty((total + 16, 2), dtype=torch.int32, device=dev)
    spo = torch.empty(len(off), dtype=torch.int64, device=dev)
    for _ in range(3):
        _check_batched(s, bd, off, pairs, spo, refs, total)
