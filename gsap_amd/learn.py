"""Training and evaluation of the learned sweep-axis selector.

Everything that touches data runs in libgsap's HIP kernels: per-scene features
(gsap_sap_features_batched), the labels -- the exact sweep-test count of every
scene on each of the 64 candidate axes, the judge (gsap_sap_axis_tests_batched)
-- and the MLP's forward, backward and Adam steps (gsap_mlp_train).  numpy
generates the synthetic scenes; torch only holds device memory.

    python -m gsap_amd.learn --scenes 100000 --out gsap_amd/data/axis_mlp.f32

The selector can only change the sweep work, never the pair set: every axis
yields the same pairs (the projection is conservative, the final test exact).
"""
import argparse
import json
import os
import time

import numpy as np

from . import AXIS_LEARNED, AXIS_PCA, AxisModel, Sap, axis_candidates
from . import workloads as W

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "axis_mlp.f32")


def make_dataset(sap, boxes_d, offsets, axes):
    """Features (ns x 64), targets (ns x 65: the regret of each candidate axis,
    tests / min tests - 1 clipped at 4, then log(1 + pairs)), the full test
    matrix (ns x 64) and the best candidate of packed scenes of <= 2048 boxes."""
    import torch
    feats = sap.features_batched(boxes_d, offsets)
    tests, pairs = sap.axis_tests_batched(boxes_d, offsets, axes)
    best = torch.argmin(tests, dim=1)  # first minimum on ties
    tmin = tests.min(dim=1, keepdim=True).values.clamp(min=1).to(torch.float32)
    regret = (tests.to(torch.float32) / tmin - 1.0).clamp(max=4.0)
    tgt = torch.cat([regret, torch.log1p(pairs.to(torch.float32))[:, None]], dim=1).contiguous()
    return feats, tgt, tests, best


def train_selector(n_scenes=100_000, seed=7, lo=128, hi=2048, epochs=40, batch=1000, lr=2e-3,
                   lambda_count=0.05, model_seed=0, log=print):
    import torch
    dev = torch.device("cuda:0")
    t0 = time.time()
    boxes, off = W.scenes(n_scenes, seed=seed, lo=lo, hi=hi)
    log(f"generated {n_scenes} scenes, {len(boxes)} boxes ({time.time() - t0:.1f}s)")
    bd = torch.from_numpy(boxes).to(dev)
    sap = Sap(axis_mode=AXIS_PCA)
    axes = axis_candidates()
    t1 = time.time()
    feats, tgt, tests, best = make_dataset(sap, bd, off, axes)
    torch.cuda.synchronize()
    log(f"features + labels (64 axes x {n_scenes} scenes) {time.time() - t1:.1f}s")
    perm = torch.from_numpy(np.random.default_rng(seed).permutation(n_scenes)).to(dev)
    x, t = feats[perm].contiguous(), tgt[perm].contiguous()
    model = AxisModel(model_seed)
    t2 = time.time()
    losses = model.train(x, t, epochs=epochs, batch=batch, lr=lr, lambda_count=lambda_count)
    log(f"trained {epochs} epochs in {time.time() - t2:.1f}s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    # held-out check on fresh small scenes: judge tests of the chosen candidate vs PCA's
    info = {"n_scenes": n_scenes, "boxes": int(len(boxes)), "loss_first": float(losses[0]),
            "loss_last": float(losses[-1]),
            "label_axis_hist": np.bincount(best.cpu().numpy(), minlength=64).tolist()}
    return model, info


def sweep_tests_batched(sap, boxes_d, offsets):
    """Total sweep tests of a batched update (count-only) and its pair count."""
    n, st = sap.collide_batched(boxes_d, offsets, None, None)
    return int(st.sweep_tests), n


def evaluate(model, boxes_d, offsets, conf_min=0.95, check_pairs=True):
    """Sweep tests with PCA vs the learned axis on a packed workload, and
    whether the two pair lists are identical."""
    import torch
    sp = Sap(axis_mode=AXIS_PCA)
    sl = Sap(axis_mode=AXIS_LEARNED)
    sl.set_learned_model(model, conf_min)
    tp, n_p = sweep_tests_batched(sp, boxes_d, offsets)
    tl, n_l = sweep_tests_batched(sl, boxes_d, offsets)
    same = n_p == n_l
    if check_pairs and same:
        dev = boxes_d.device
        pp = torch.empty((max(n_p, 1), 2), dtype=torch.int32, device=dev)
        pl = torch.empty((max(n_l, 1), 2), dtype=torch.int32, device=dev)
        op = torch.empty(len(offsets), dtype=torch.int64, device=dev)
        ol = torch.empty(len(offsets), dtype=torch.int64, device=dev)
        sp.collide_batched(boxes_d, offsets, pp, op)
        sl.collide_batched(boxes_d, offsets, pl, ol)
        same = bool(torch.equal(pp[:n_p], pl[:n_l]) and torch.equal(op, ol))
    return {"tests_pca": tp, "tests_learned": tl, "reduction": 1.0 - tl / max(tp, 1),
            "pairs": n_p, "pairs_identical": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=100_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--eval-scenes", type=int, default=2048)
    args = ap.parse_args()
    import torch
    model, info = train_selector(n_scenes=args.scenes, seed=args.seed, epochs=args.epochs)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    model.save(args.out)
    boxes, off = W.scenes(args.eval_scenes, seed=5, lo=1024, hi=65536)
    ev = evaluate(model, torch.from_numpy(boxes).to("cuda:0"), off)
    info["eval"] = ev
    print(json.dumps(info))


if __name__ == "__main__":
    main()
