"""Training and evaluation of the learned sweep-axis selector.

Everything that touches data runs in libgsap's HIP kernels: per-scene features
(gsap_sap_features_batched), the labels -- the exact sweep-test count of every
scene on each of the 64 candidates, the judge (gsap_sap_axis_tests_batched) --
and the MLP's forward, backward and Adam steps (gsap_mlp_train).  numpy
generates the synthetic scenes; torch only holds device memory.

Two models, one per sweep kind:
  plain sweep  the network names the sweep axis; labels from the judge's
               in-LDS sort (scenes of at most 2,048 boxes)
  slab sweep   the network names the normal of the sweep plane (the sweep and
               slab axes span the plane normal to it); labels are the slab
               sweep's own per-scene test counts (count-only batched updates,
               scenes of any size)

    python -m gsap_amd.learn --scenes 100000 --out gsap_amd/data/axis_mlp.f32
    python -m gsap_amd.learn --slabs --scenes 16384 --lo 256 --hi 65536 \
        --out gsap_amd/data/axis_mlp_slabs.f32

The selector can only change the sweep work, never the pair set: every axis
(pair) yields the same pairs (the projection is conservative, the final test
exact).
"""
import argparse
import json
import os
import time

import numpy as np

from . import AXIS_AUTO, AXIS_LEARNED, AXIS_PCA, FLAG_SLABS, AxisModel, Sap, axis_candidates
from . import workloads as W

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(_DATA, "axis_mlp.f32")
DEFAULT_MODEL_SLABS = os.path.join(_DATA, "axis_mlp_slabs.f32")


def default_model(slabs=False):
    return DEFAULT_MODEL_SLABS if slabs else DEFAULT_MODEL


def make_dataset(sap, boxes_d, offsets, axes, regret_scale=1.0, slabs=False):
    """Features (ns x 64), targets (ns x 65: the regret of each candidate,
    (tests / min tests - 1) x regret_scale clipped at 4, then log(1 + pairs)),
    the full test matrix (ns x 64) and the best candidate of packed scenes
    (plain-sweep judge: scenes of at most 2,048 boxes; slab handles: any size).
    Slab mode: only the three coordinate normals are sweep planes (a tilted
    plane lengthens every box's projection; the inference snaps to them), so
    only they are judged, and every candidate carries the counts of the
    coordinate normal nearest to it -- the loss's softmax over directions then
    weighs the three planes, not the tilted directions around them (which are
    not spread evenly over the three)."""
    import torch
    feats = sap.features_batched(boxes_d, offsets)
    if slabs:
        ax = np.asarray(axes)
        near = torch.from_numpy(np.argmax(np.abs(ax), axis=1)).to(boxes_d.device)
        t3, pairs = sap.axis_tests_batched(boxes_d, offsets, ax[:3])
        tests = t3[:, near].contiguous()
    else:
        tests, pairs = sap.axis_tests_batched(boxes_d, offsets, axes)
    best = torch.argmin(tests, dim=1)  # first minimum on ties
    tmin = tests.min(dim=1, keepdim=True).values.clamp(min=1).to(torch.float32)
    regret = ((tests.to(torch.float32) / tmin - 1.0) * regret_scale).clamp(max=4.0)
    tgt = torch.cat([regret, torch.log1p(pairs.to(torch.float32))[:, None]], dim=1).contiguous()
    return feats, tgt, tests, best


def train_selector(n_scenes=100_000, seed=7, lo=128, hi=2048, epochs=40, batch=1000, lr=2e-3,
                   lambda_count=0.05, model_seed=0, slabs=False, chunk=4096, regret_scale=1.0,
                   log=print):
    """Generate labelled scenes (in chunks of `chunk` scenes) and train.  The
    slab sweep's candidates differ by a few percent, so its regrets are
    scaled up (regret_scale) for the loss to tell them apart."""
    import torch
    dev = torch.device("cuda:0")
    t0 = time.time()
    sap = Sap(axis_mode=AXIS_PCA, flags=FLAG_SLABS if slabs else 0)
    axes = axis_candidates()
    parts, nboxes, t_lab = [], 0, 0.0
    for c0 in range(0, n_scenes, chunk):
        ns = min(chunk, n_scenes - c0)
        boxes, off = W.scenes(ns, seed=seed * 100003 + c0, lo=lo, hi=hi)
        nboxes += len(boxes)
        bd = torch.from_numpy(boxes).to(dev)
        t1 = time.time()
        parts.append(make_dataset(sap, bd, off, axes, regret_scale, slabs))
        torch.cuda.synchronize()
        t_lab += time.time() - t1
        del bd
    feats, tgt, tests, best = (torch.cat([p[i] for p in parts]) for i in range(4))
    log(f"{n_scenes} scenes, {nboxes} boxes; features + labels (64 candidates) {t_lab:.1f}s "
        f"(total {time.time() - t0:.1f}s)")
    perm = torch.from_numpy(np.random.default_rng(seed).permutation(n_scenes)).to(dev)
    x, t = feats[perm].contiguous(), tgt[perm].contiguous()
    model = AxisModel(model_seed)
    t2 = time.time()
    losses = model.train(x, t, epochs=epochs, batch=batch, lr=lr, lambda_count=lambda_count)
    log(f"trained {epochs} epochs in {time.time() - t2:.1f}s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    # held-out check on fresh small scenes: judge tests of the chosen candidate vs PCA's
    info = {"n_scenes": n_scenes, "boxes": int(nboxes), "slabs": bool(slabs),
            "scene_sizes": [lo, hi], "regret_scale": regret_scale, "epochs": epochs,
            "loss_first": float(losses[0]),
            "loss_last": float(losses[-1]),
            "label_axis_hist": np.bincount(best.cpu().numpy(), minlength=64).tolist()}
    return model, info


def sweep_tests_batched(sap, boxes_d, offsets):
    """Total sweep tests of a batched update (count-only) and its pair count."""
    n, st = sap.collide_batched(boxes_d, offsets, None, None)
    return int(st.sweep_tests), n


def evaluate(model, boxes_d, offsets, conf_min=0.95, check_pairs=True, slabs=False):
    """Sweep tests with PCA, the learned axis (slab mode: plane) and the
    closed-form automatic rule on a packed workload, and whether the PCA and
    learned pair lists are identical."""
    import torch
    fl = FLAG_SLABS if slabs else 0
    sp = Sap(axis_mode=AXIS_PCA, flags=fl)
    sl = Sap(axis_mode=AXIS_LEARNED, flags=fl)
    sa = Sap(axis_mode=AXIS_AUTO, flags=fl)
    sl.set_learned_model(model, conf_min)
    tp, n_p = sweep_tests_batched(sp, boxes_d, offsets)
    tl, n_l = sweep_tests_batched(sl, boxes_d, offsets)
    ta, n_a = sweep_tests_batched(sa, boxes_d, offsets)
    same = n_p == n_l == n_a
    if check_pairs and same:
        dev = boxes_d.device
        pp = torch.empty((max(n_p, 1), 2), dtype=torch.int32, device=dev)
        pl = torch.empty((max(n_l, 1), 2), dtype=torch.int32, device=dev)
        op = torch.empty(len(offsets), dtype=torch.int64, device=dev)
        ol = torch.empty(len(offsets), dtype=torch.int64, device=dev)
        sp.collide_batched(boxes_d, offsets, pp, op)
        sl.collide_batched(boxes_d, offsets, pl, ol)
        same = bool(torch.equal(pp[:n_p], pl[:n_l]) and torch.equal(op, ol))
    return {"tests_pca": tp, "tests_learned": tl, "tests_auto": ta, "reduction": 1.0 - tl / max(tp, 1),
            "reduction_vs_auto": 1.0 - tl / max(ta, 1), "pairs": n_p, "pairs_identical": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=100_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--slabs", action="store_true", help="train the slab-sweep (plane) selector")
    ap.add_argument("--lo", type=int, default=128, help="smallest training scene")
    ap.add_argument("--hi", type=int, default=2048, help="largest training scene")
    ap.add_argument("--regret-scale", type=float, default=None,
                    help="regret multiplier of the targets (default: 4 plain, 20 slabs)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval-scenes", type=int, default=2048)
    args = ap.parse_args()
    import torch
    model, info = train_selector(n_scenes=args.scenes, seed=args.seed, lo=args.lo, hi=args.hi,
                                 epochs=args.epochs, slabs=args.slabs,
                                 regret_scale=args.regret_scale or (20.0 if args.slabs else 4.0))
    out = args.out or default_model(args.slabs)
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    model.save(out)
    boxes, off = W.scenes(args.eval_scenes, seed=5, lo=1024, hi=65536)
    ev = evaluate(model, torch.from_numpy(boxes).to("cuda:0"), off, slabs=args.slabs)
    info["eval"] = ev
    print(json.dumps(info))


if __name__ == "__main__":
    main()
