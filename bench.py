"""Benchmark: exact-pair SaP updates/s at 1,048,576 AABBs on one MI355X.

One "step" is one complete broadphase update of the anisotropic 1M-box
workload: device-resident boxes in, device-resident canonical pair list
(i<j, lexicographic) out with its count known on the host.  Inputs are
generated once (seeded) and uploaded before timing.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

Prints ONE JSON line.  `roofline` is for the dominant kernel of the update,
timed with hipEvents recorded by the library on the launch stream between
stages, over the timed steps themselves.  `cpu_baseline` is the CPU
restatement of the reference algorithm (oracle/) timed on this host on a
bounded sample: full axis/projection/sort plus the sweep of every 64th sorted
object, scaled by 64.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--axis", default="pca", choices=["pca", "coord", "learned"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gsap_amd as G
    from gsap_amd import workloads as W

    dev = torch.device("cuda:0")
    boxes = W.aniso(args.n, seed=1)
    n = len(boxes)
    bd = torch.from_numpy(boxes).to(dev)
    mode = {"pca": G.AXIS_PCA, "coord": G.AXIS_COORD_MAXVAR, "learned": G.AXIS_LEARNED}[args.axis]
    sap = G.Sap(max_n=n, axis_mode=mode)
    P, st0 = sap.count(bd)  # pre-count pass sizes the 64-bit pair buffer
    pairs = torch.empty((max(P, 1), 2), dtype=torch.int32, device=dev)
    sap.set_timing(True)
    for _ in range(args.warmup):
        sap.collide_into(bd, pairs)
    torch.cuda.synchronize()
    stage_sum = None
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        P2, _, st = sap.collide_into(bd, pairs)
        s = np.array(sap.stage_times(), np.float64)
        stage_sum = s if stage_sum is None else stage_sum + s
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert P2 == P
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    value = n / (ms_per_step / 1e3)

    names = ["moments+project", "object_sort", "gather", "sweep", "pair_sort", "finish"]
    stages = dict(zip(names, (stage_sum / args.steps).tolist()))
    # algorithmic bytes per launch (per-unit figures x units)
    sort_passes = 4
    alg = {
        "sweep": 24.0 * n + 8.0 * P,                  # boxes streamed + pairs written
        "object_sort": 16.0 * n * sort_passes,         # key+value read+write per pass
        "pair_sort": 16.0 * P * 5 + 8.0 * P,           # 5 passes + histogram read
    }
    dom = max(("sweep", "pair_sort"), key=lambda k: stages[k])
    launches = {"sweep": 1, "pair_sort": 6}[dom]
    dom_ms = stages[dom] / launches
    dom_bytes = alg[dom] / launches
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_sweep" if dom == "sweep" else "k_onesweep<pair>",
            "bytes_per_launch": dom_bytes, "ms_per_launch": round(dom_ms, 4)}
    sort_gbs = alg["object_sort"] / (stages["object_sort"] / 1e3) / 1e9
    sort_roof = {"achieved": round(sort_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(sort_gbs / HBM_PEAK_GBS, 4), "ms": round(stages["object_sort"], 4),
                 "bytes": alg["object_sort"], "passes": sort_passes}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        stride = args.cpu_stride
        t = time.perf_counter()
        _, tests_s, _ = O.sap(boxes, axis_mode={"pca": O.AXIS_PCA}.get(args.axis, O.AXIS_PCA),
                              sample_stride=stride, cap=max(16, 4 * P // stride + 1024))
        setup, sweep, canon = O.last_times()
        est = setup + stride * (sweep + canon)
        cpu = {"value": round(n / est, 1), "unit": "updates/s", "cores": 1, "kind": "port",
               "sample": f"axis+projection+radix sort of all {n} boxes ({setup:.2f}s), sweep and "
                         f"canonical sort of every {stride}th sorted object ({sweep + canon:.2f}s), "
                         f"extrapolated x{stride}: {est:.1f}s per update; "
                         f"{time.perf_counter() - t:.1f}s measured"}

    out = {
        "metric": "exact-pair SaP updates/s @1M AABBs",
        "value": round(value, 1),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"aniso-{n}: AABBs, centres R diag(20,5,1) z, half-extents "
                               "log-U[0.05,0.8], seed 1, canonical uint32 pair list",
                   "n": n, "pairs": P, "axis": args.axis,
                   "sweep_tests": int(st.sweep_tests), "axis_used": [round(x, 6) for x in st.axis]},
        "roofline": roof,
        "sort_roofline": sort_roof,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "wall_s": round(wall, 3),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
