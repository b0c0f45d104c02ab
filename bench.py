"""Benchmark: exact-pair SaP updates/s at 1,048,576 AABBs on one MI355X.

One "step" is one complete broadphase update of the anisotropic 1M-box
workload: device-resident boxes in, device-resident canonical pair list
(i<j, lexicographic) out with its count known on the host.  Inputs are
generated once (seeded) and uploaded before timing.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

Prints ONE JSON line.  `roofline` is for the dominant kernel of the update,
timed with hipEvents recorded by the library on the launch stream between
stages, over as many steps again right after the timed ones (the timed steps
run without those events).  `cpu_baseline` is the reference's
SaP method restated on the CPU (oracle/cpu_sap.c, OpenMP, built with
-march=native on this host): full updates of the same boxes on all threads
and on one thread, with the host's core count and CPU model.  The timed GPU
list is compared with the default path's and with the CPU list (SHA-256).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def stage_dict(t, local=False):
    """Library stage times [moments+project, object_sort, gather, sweep,
    pair_hist, one per pair-sort onesweep pass..., pair_local (with the local
    stage), pair_tail, finish] -> (named stages, per-pass array).  With the
    sort by rows (local == 2) the stages after the sweep are [rows,
    pair_tail, finish]; the rows stage is reported as pair_local."""
    import numpy as np
    t = np.asarray(t, np.float64)
    if local == 2:
        st = {"moments+project": t[0], "object_sort": t[1], "gather": t[2], "sweep": t[3],
              "pair_hist": 0.0, "pair_passes": 0.0, "pair_local": t[4], "pair_tail": t[5],
              "finish": t[6]}
        st["pair_sort"] = st["pair_local"] + st["pair_tail"]
        return {k: float(v) for k, v in st.items()}, np.zeros(0)
    passes = t[5:len(t) - (3 if local else 2)]
    st = {"moments+project": t[0], "object_sort": t[1], "gather": t[2], "sweep": t[3],
          "pair_hist": t[4], "pair_passes": float(passes.sum()),
          "pair_local": t[-3] if local else 0.0, "pair_tail": t[-2], "finish": t[-1]}
    st["pair_sort"] = st["pair_hist"] + st["pair_passes"] + st["pair_local"] + st["pair_tail"]
    return {k: float(v) for k, v in st.items()}, passes


def traffic_for(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the same counters on a
    copy of known size), or None when none is committed."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def issue_for(kernel):
    """Instruction-issue fractions of `kernel` from the committed counter
    summary (rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS, wave64
    instructions per dispatch, over the same dispatches' average duration):
    each count / (CUs x shader clock x duration), i.e. against one wave64
    instruction of that kind per CU per cycle.  None when none is committed."""
    path = os.path.join(ROOT, "profiles", "issue_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    if k is None:
        return None
    cyc = d["cus"] * d["clock_mhz"] * 1e6 * k["avg_ns"] * 1e-9
    return {"valu": round(k["valu"] / cyc, 4), "salu": round(k["salu"] / cyc, 4),
            "lds": round(k["lds"] / cyc, 4), "avg_ns": k["avg_ns"],
            "source": "profiles/issue_counters.json"}


def copy_floor(torch, dev, nbytes, reps=50):
    """A plain device copy of nbytes (half read, half written), timed with
    hipEvents: the fraction of HBM peak a pass that moves nbytes could reach
    at best at this size (launch and ramp included)."""
    n = max(1, int(nbytes // 8))
    src = torch.ones(n, dtype=torch.int32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 8.0 * n / (ms / 1e3) / 1e9
    return {"bytes": 8 * n, "ms": round(ms, 5), "achieved": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def _roof(alg_bytes, ms):
    """Roofline fraction of algorithmic bytes moved in `ms` against HBM peak."""
    gbs = alg_bytes / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes": alg_bytes}


def _time_ms(torch, fn, steps, warmup):
    """Mean ms per call of fn over `steps` calls after `warmup`, from events on
    torch's current stream (the stream every library call here launches on)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def host_info():
    """Cores and CPU model of the measuring host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = int(omp) if omp and omp.isdigit() else aff
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": omp,
            "threads_used": max(1, min(threads, aff)), "cpu_model": model}


_NATIVE = {}


def cpu_lib():
    """The CPU baseline library built for this host (-O3 -march=native
    -ffp-contract=off -fopenmp), else the portable prebuilt one."""
    if "path" not in _NATIVE:
        import tempfile
        from oracle import oracle as O
        d = tempfile.mkdtemp(prefix="gsap_cpu_")
        _NATIVE["path"] = O.build_native_cpu_sap(d)
        _NATIVE["flags"] = ("-O3 -march=native -ffp-contract=off -fopenmp" if _NATIVE["path"] else
                            "-O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (prebuilt)")
    return _NATIVE["path"], _NATIVE["flags"]


def _sha(a):
    import hashlib
    import numpy as np
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8)).hexdigest()


def cpu_baseline_c2(boxes, gpu_pairs, P, runs):
    """The reference's SaP method restated on the CPU (oracle/cpu_sap.c:
    coordinate max-variance axis, flipped keys, stable 8-bit LSD radix sort,
    sweep with early exit, canonical sort), full updates of the same boxes:
    all threads (median of `runs` after one warm-up) and one thread (one
    update).  The CPU list's SHA-256 is compared with the GPU's."""
    import numpy as np
    from oracle import oracle as O
    lib, flags = cpu_lib()
    info = host_info()
    nt = info["threads_used"]
    n = len(boxes)
    O.cpu_sap(boxes, nt, cap=P, lib_path=lib)  # warm-up (page faults, thread pool)
    ts, phases = [], None
    for _ in range(runs):
        t = time.perf_counter()
        cp, tests, ax, ph = O.cpu_sap(boxes, nt, cap=P, lib_path=lib)
        ts.append(time.perf_counter() - t)
        phases = ph
    t_all = float(np.median(ts))
    same = len(cp) == P and _sha(cp) == _sha(gpu_pairs.cpu().numpy().view(np.uint32))
    del cp
    t = time.perf_counter()
    _, _, _, ph1 = O.cpu_sap(boxes, 1, cap=P, lib_path=lib)
    t_one = time.perf_counter() - t
    # the reference's own sweep work on this input: its 4-cell pass plus
    # cross pass (window sizes only), for comparison with sweep_tests
    _, ref_work = O.sap_ref_cells(boxes, count_only=True)
    return {"value": round(n / t_all, 1), "unit": "updates/s", "cores": nt, "kind": "port",
            "ms_per_update": round(t_all * 1e3, 1),
            "single_thread": {"value": round(n / t_one, 1), "ms_per_update": round(t_one * 1e3, 1),
                              "phases_s": {k: round(v, 3) for k, v in ph1.items()}},
            "phases_s": {k: round(v, 3) for k, v in phases.items()},
            "sweep_tests": tests, "axis": "xyz"[ax], "pairs_sha256_equal_gpu": same,
            "host": info, "build": flags,
            "reference_sweep_tests": ref_work["cell_tests"] + ref_work["cross_tests"],
            "reference_sweep_work": ref_work,
            "sample": f"full updates of all {n} boxes (no sampling): median of {runs} on {nt} "
                      f"threads after one warm-up, and one update on 1 thread"}


def bench_c1(G, W, torch, dev, steps, warmup, with_cpu, cache=None):
    """16,384 boxes (centres U[0,64)^3, half-extents U[0.1,0.5] per axis, seed
    0), PCA axis, autotuned handle (as the headline's): GPU updates/s, the CPU
    baseline on 1 and all threads (median of 10 after 2 warm-ups) and the
    O(n^2) brute force the lists are checked against."""
    import numpy as np
    from oracle import oracle as O
    boxes = W.small()
    n = len(boxes)
    bd = torch.from_numpy(boxes).to(dev)
    sap = G.Sap(max_n=n, axis_mode=G.AXIS_PCA)
    tune = sap.autotune(bd, cache=cache)
    P, st = sap.count(bd)
    pairs = torch.empty((max(P, 1), 2), dtype=torch.int32, device=dev)
    ms = _time_ms(torch, lambda: sap.collide_into(bd, pairs), steps, warmup)
    got = pairs[:P].cpu().numpy().view(np.uint32)
    t = time.perf_counter()
    ref = O.brute(boxes)
    t_brute = time.perf_counter() - t
    out = {"workload": "16384 AABBs, centres U[0,64)^3, half-extents U[0.1,0.5] per axis, seed 0, "
                       "PCA axis", "n": n, "pairs": P, "sweep_tests": int(st.sweep_tests),
           "ms_per_update": round(ms, 4), "updates_per_s": round(n / (ms / 1e3), 1),
           "identical_to_brute_force": bool(np.array_equal(got, ref)),
           "brute_force_s": round(t_brute, 3),
           "tune": {k: v for k, v in tune.items() if k != "from_cache"}}
    if with_cpu:
        lib, flags = cpu_lib()
        info = host_info()
        res = {}
        for name, nt in (("cpu_1thread", 1), ("cpu_all_threads", info["threads_used"])):
            ts = []
            for k in range(12):
                t = time.perf_counter()
                cp, _, _, _ = O.cpu_sap(boxes, nt, lib_path=lib)
                if k >= 2:
                    ts.append(time.perf_counter() - t)
            med = float(np.median(ts))
            res[name] = {"threads": nt, "ms_per_update": round(med * 1e3, 3),
                         "updates_per_s": round(n / med, 1),
                         "identical": bool(np.array_equal(cp, ref))}
        out.update(res)
        out["cpu_build"] = flags
    return out


def sap_bytes(n, P):
    """Algorithmic bytes of one SaP update (every stage streamed once):
    moments 24n, projection 32n, radix sort 68n (8-bit digits), sorted-box
    gather 52n, sweep 24n + 8P, canonical CSR 8n + 16P -> 208n + 24P."""
    return 208.0 * n + 24.0 * P


def particle_step_bytes(nb, npart, directed, P):
    """Algorithmic bytes of one rigid-body step: per body 64 B read, 48 B
    written, 48 B inverse inertia, 16 B force/torque read; per particle 36 B
    written and 52 B read (update), 16 B written and 16 B read (forces); per
    directed pair 8 B + 36 B partner state; plus the SaP update of the
    particles."""
    return nb * (64.0 + 48 + 48 + 16) + npart * (36.0 + 52 + 16 + 16) + directed * 44.0 + \
        sap_bytes(npart, P)


def bench_batched(G, W, torch, dev, ns, steps, warmup, cache=None):
    """512 packed scenes (sizes log-uniform in [1024, 65536], seed 4) in one
    graph-replayed update, against the same kernels run scene by scene."""
    import numpy as np
    boxes, off = W.scenes(ns, seed=4)
    ntot = int(off[-1])
    bd = torch.from_numpy(boxes).to(dev)
    sb = G.Sap(max_n=ntot, axis_mode=G.AXIS_AUTO)
    tune = sb.autotune(bd, scene_offsets=off, cache=cache)
    P, st = sb.collide_batched(bd, off, None, None)
    pairs = torch.empty((max(P, 1), 2), dtype=torch.int32, device=dev)
    poff = torch.empty(ns + 1, dtype=torch.int64, device=dev)
    ms_b = _time_ms(torch, lambda: sb.collide_batched(bd, off, pairs, poff), steps, warmup)
    po = poff.cpu().numpy()
    sp = G.Sap(max_n=int(np.max(np.diff(off))), axis_mode=G.AXIS_AUTO)
    lp = torch.empty_like(pairs)
    counts = torch.zeros((ns, 8), dtype=torch.int64, device=dev)
    views = [(bd[off[s]:off[s + 1]], lp[po[s]:po[s + 1]]) for s in range(ns)]

    def loop_async():
        for s in range(ns):
            sp.collide_async(views[s][0], views[s][1], counts[s])

    def loop_sync():
        for s in range(ns):
            sp.collide_into(views[s][0], views[s][1])

    ms_la = _time_ms(torch, loop_async, max(1, steps // 4), 1)
    ms_ls = _time_ms(torch, loop_sync, max(1, steps // 4), 1)
    loop_async()
    torch.cuda.synchronize()
    shift = torch.repeat_interleave(torch.from_numpy(off[:-1].astype(np.int32)).to(dev),
                                    torch.from_numpy(np.diff(po)).to(dev))
    same = bool(torch.equal(counts[:, 0].cpu(), torch.from_numpy(np.diff(po))) and
                torch.equal(lp[:P] + shift[:, None], pairs[:P]))
    return {"workload": f"{ns} scenes, n log-U[1024,65536] seed 4, uniform or 8-cluster centres, "
                        "half-extents U[0.1,0.6]",
            "boxes": ntot, "pairs": P, "sweep_tests": int(st.sweep_tests),
            "ms_batched": round(ms_b, 4), "updates_per_s": round(ntot / (ms_b / 1e3), 1),
            "ms_per_scene_loop_async": round(ms_la, 4),
            "ms_per_scene_loop_sync": round(ms_ls, 4),
            "speedup_vs_loop_async": round(ms_la / ms_b, 2),
            "speedup_vs_loop_sync": round(ms_ls / ms_b, 2),
            "pairs_identical": same,
            # the whole batched update against the SaP byte model
            "roofline": _roof(sap_bytes(ntot, P), ms_b),
            "tune": {k: v for k, v in tune.items() if k != "from_cache"}}


def bench_learned(G, W, torch, dev, ns):
    """2048 packed scenes (seed 5), plain sweep and slab sweep: sweep tests and
    update time with the PCA axis (pair), the trained selector (plain: the
    sweep axis; slabs: the sweep plane) and the closed-form automatic rule; the
    pair lists must be identical."""
    from gsap_amd import learn as L
    boxes, off = W.scenes(ns, seed=5)
    bd = torch.from_numpy(boxes).to(dev)
    out = {"workload": f"{ns} scenes, n log-U[1024,65536] seed 5", "boxes": int(off[-1])}
    for slabs in (False, True):
        path = L.default_model(slabs)
        if not os.path.exists(path):
            rec = {"skipped": "no trained model at " + os.path.relpath(path, ROOT)}
        else:
            model = G.AxisModel.load(path)
            ev = L.evaluate(model, bd, off, conf_min=0.95, slabs=slabs)
            fl = G.FLAG_SLABS if slabs else 0
            ms = {}
            for name, mode in (("pca", G.AXIS_PCA), ("learned", G.AXIS_LEARNED), ("auto", G.AXIS_AUTO)):
                s = G.Sap(max_n=int(off[-1]), axis_mode=mode, flags=fl)
                if mode == G.AXIS_LEARNED:
                    s.set_learned_model(model, 0.95)
                P, _ = s.collide_batched(bd, off, None, None)
                pairs = torch.empty((max(P, 1), 2), dtype=torch.int32, device=dev)
                po = torch.empty(ns + 1, dtype=torch.int64, device=dev)
                ms[name] = _time_ms(torch, lambda: s.collide_batched(bd, off, pairs, po), 5, 2)
                s.close()
            rec = {"pairs": ev["pairs"], "sweep_tests_pca": ev["tests_pca"],
                   "sweep_tests_learned": ev["tests_learned"], "sweep_tests_auto": ev["tests_auto"],
                   "reduction": round(ev["reduction"], 4),
                   "reduction_vs_auto": round(ev["reduction_vs_auto"], 4),
                   "pairs_identical": ev["pairs_identical"], "ms_batched_pca": round(ms["pca"], 4),
                   "ms_batched_learned": round(ms["learned"], 4), "ms_batched_auto": round(ms["auto"], 4)}
        if slabs:
            out["slabs"] = rec
        else:
            out.update(rec)
    return out


def _d2h(ptr, nfloat):
    """Copy nfloat float32 values at a device address to the host."""
    import ctypes
    import numpy as np
    out = np.empty(nfloat, np.float32)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ptr, out.nbytes, 2) == 0
    return out


def _tune_particles(G, torch, sap, pp, npart, cache=None):
    """Autotune the particle broadphase on the step's own particle AABBs and
    body ids (device buffers owned by the rigid-body handle)."""
    import ctypes
    aabb = torch.empty((npart, 6), dtype=torch.float32, device="cuda")
    pos = torch.empty((npart, 4), dtype=torch.float32, device="cuda")
    body = torch.empty(npart, dtype=torch.int32, device="cuda")
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    assert hip.hipMemcpy(pos.data_ptr(), pp.p_pos, npart * 16, 3) == 0
    assert hip.hipMemcpy(body.data_ptr(), pp.p_body, npart * 4, 3) == 0
    r = pos[:, 3:4]
    aabb[:, :3] = pos[:, :3] - r
    aabb[:, 3:] = pos[:, :3] + r
    return sap.autotune(aabb, group=body, cache=cache)


def bench_particles(G, W, torch, dev, nb, steps, with_cpu, cache=None):
    """Rigid cubes (3x3x3 particles, radius 0.5, spacing 1) dropped into a
    256x256 walled box from heights U[0,200]: spring-dashpot k=1e4, c=50,
    g=-9.81, dt=1e-4, per-step SaP on the particle AABBs (same-body pairs
    skipped), symmetric CSR pairs feeding the force kernel."""
    import numpy as np
    sc = W.bodies(nb, seed=3)
    npart = nb * 27
    flags = G.PRB_ABS_SPRING | G.PRB_WALLS | G.PRB_SKIP_SAME_BODY
    kw = dict(gravity=(0, -9.81, 0), k_abs=1e4, c_abs=50.0, floor_scale=1.0, wall_lo=(0, 0),
              wall_hi=(256.0, 256.0), flags=flags)
    dt = 1e-4
    runs = {}
    # The same scene stepped with a from-scratch broadphase handle and with a
    # coherent one (GSAP_FLAG_COHERENT: key function kept, sort buckets from
    # the previous step's order); both end in the same state.
    for name, coh in (("from_scratch", 0), ("coherent", G.FLAG_COHERENT)):
        prb = G.ParticleRigidBody(nb, npart, npart * 8)
        prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"],
                        sc["pstart"], sc["npart"], sc["inv_mass"], sc["inv_inertia"])
        prb.set_particle_data(sc["ref_pos"], sc["ref_mass"], sc["max_per_shape"])
        prb.set_particle_shape_inv_inertia(sc["shape_inv_inertia"])
        prb.set_physics(broadphase_kind=G.INPUT_AABB, **kw)
        sap = G.Sap(max_n=npart, input_kind=G.INPUT_AABB, output_kind=G.OUT_SYM_CSR,
                    axis_mode=G.AXIS_AUTO, flags=G.FLAG_SKIP_SAME_GROUP | G.FLAG_SLABS | coh)
        prb.step(sap, dt)  # particle AABBs of the first step, to tune the broadphase on
        tune = _tune_particles(G, torch, sap, prb.ptrs(), npart, cache)
        ms = _time_ms(torch, lambda: prb.step(sap, dt), steps, 10)
        p = prb.ptrs()  # raises on pair-buffer overflow
        hip_pos = _d2h(p.rb_pos, nb * 4)
        runs[name] = (ms, tune, int(p.num_pairs), hip_pos)
        prb.close()
        sap.close()
    ms, tune, directed, pos_s = runs["from_scratch"]
    ms_c, tune_c, _, pos_c = runs["coherent"]
    out = {"workload": f"{nb} cubes x 27 particles, {steps} steps, dt 1e-4, AABB broadphase "
                       "(slab sweep)",
           "particles": npart, "steps": steps, "ms_per_step": round(ms, 4),
           "steps_per_s": round(1e3 / ms, 1), "particle_updates_per_s": round(npart * 1e3 / ms, 1),
           "ms_per_step_coherent": round(ms_c, 4),
           "coherent_final_state_identical": bool(np.array_equal(pos_c, pos_s)),
           "directed_pairs_last_step": directed,
           "roofline": _roof(particle_step_bytes(nb, npart, directed, directed // 2), ms),
           "tune": {k: v for k, v in tune.items() if k != "from_cache"},
           "tune_coherent": {k: v for k, v in tune_c.items() if k != "from_cache"}}
    if with_cpu:
        # Full-size CPU steps of the same scene: the oracle's dynamics (one
        # thread) and the CPU-baseline broadphase on the particle AABBs (all
        # threads, then one), in the DemoParticles::step order.
        from oracle import oracle as O
        lib, flags = cpu_lib()
        info = host_info()
        nt = info["threads_used"]
        ref = O.PrbState(sc["pos"], sc["quat"] / np.linalg.norm(sc["quat"], axis=1, keepdims=True),
                         sc["linvel"], sc["angvel"], sc["shape"], sc["pstart"], sc["npart"],
                         sc["inv_mass"], sc["ref_pos"], sc["ref_mass"], sc["max_per_shape"],
                         sc["shape_inv_inertia"])
        params = O.prb_params(**kw)
        grp = ref.pbody.astype(np.uint32)

        def cpu_step(threads):
            t = time.perf_counter()
            ref.integrate(dt, params)
            ref.update_particles()
            canon, _, _, _ = O.cpu_sap(ref.particle_aabbs(), threads, group=grp, lib_path=lib)
            sym, _ = O.sym_csr(canon)
            ref.prepare()
            ref.forces(sym, dt, params)
            ref.update_velocity(dt, params)
            return time.perf_counter() - t

        cpu_step(nt)  # warm-up
        t_all = float(np.median([cpu_step(nt) for _ in range(3)]))
        t_one = cpu_step(1)
        out["cpu_baseline"] = {"value": round(npart / t_all, 1), "unit": "particle updates/s",
                               "cores": nt, "kind": "port", "ms_per_step": round(t_all * 1e3, 1),
                               "single_thread": {"value": round(npart / t_one, 1),
                                                 "ms_per_step": round(t_one * 1e3, 1)},
                               "build": flags,
                               "sample": f"full steps of all {npart} particles: median of 3 with the "
                                         f"broadphase on {nt} threads after one warm-up, then one "
                                         "step on 1 thread; dynamics single-threaded"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--axis", default="auto", choices=["auto", "pca", "coord", "learned"],
                    help="sweep axis rule (auto: PCA or coordinate axes by the sweep-test model)")
    ap.add_argument("--no-slabs", action="store_true",
                    help="plain sweep along one axis (no slab split on a second axis)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="timed full CPU-baseline updates at 1M boxes on all threads")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched, learned-axis and particle-step sections")
    ap.add_argument("--scenes", type=int, default=512, help="batched workload size")
    ap.add_argument("--learned-scenes", type=int, default=2048)
    ap.add_argument("--prb-bodies", type=int, default=65536)
    ap.add_argument("--prb-steps", type=int, default=1000)
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file (a hit skips the measurement of configurations)")
    ap.add_argument("--only", choices=["batched", "learned", "particles", "c1"], default=None,
                    help="run one extra section alone and print its JSON (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {args.gpus} requested); "
              "the broadphase runs on a single GPU and its sharding is not implemented",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gsap_amd as G
    from gsap_amd import workloads as W

    dev = torch.device("cuda:0")
    if args.only:
        if args.only == "batched":
            r = bench_batched(G, W, torch, dev, args.scenes, args.steps, args.warmup, args.tune_cache)
        elif args.only == "learned":
            r = bench_learned(G, W, torch, dev, args.learned_scenes)
        elif args.only == "c1":
            r = bench_c1(G, W, torch, dev, args.steps, args.warmup, not args.no_cpu_baseline,
                         args.tune_cache)
        else:
            r = bench_particles(G, W, torch, dev, args.prb_bodies, args.prb_steps,
                                not args.no_cpu_baseline, args.tune_cache)
        print(json.dumps({"section": args.only, **r}))
        return
    boxes = W.aniso(args.n, seed=1)
    n = len(boxes)
    bd = torch.from_numpy(boxes).to(dev)
    mode = {"auto": G.AXIS_AUTO, "pca": G.AXIS_PCA, "coord": G.AXIS_COORD_MAXVAR,
            "learned": G.AXIS_LEARNED}[args.axis]
    slabs = not args.no_slabs
    sap = G.Sap(max_n=n, axis_mode=mode, flags=G.FLAG_SLABS if slabs else 0)
    tune = sap.autotune(bd, cache=args.tune_cache)  # measured tiles, digit widths, sweep kind
    P, st0 = sap.count(bd)  # pre-count pass sizes the 64-bit pair buffer
    pairs = torch.empty((max(P, 1), 2), dtype=torch.int32, device=dev)
    # The timed steps run without the library's per-stage events (a
    # diagnostic: each costs a few microseconds at its stage boundary); the
    # stage breakdown comes from as many steps again with them on.
    sap.set_timing(False)
    for _ in range(args.warmup):
        sap.collide_into(bd, pairs)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        P2, _, st = sap.collide_into(bd, pairs)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert P2 == P
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    value = n / (ms_per_step / 1e3)
    sap.set_timing(True)
    stage_sum = None
    for _ in range(args.steps):
        sap.collide_into(bd, pairs)
        s = np.array(sap.stage_times(), np.float64)
        stage_sum = s if stage_sum is None else stage_sum + s

    tn = sap.tune
    stages, pass_ms = stage_dict(stage_sum / args.steps, tn["pair_local"])
    # Algorithmic bytes per launch: per-unit figures x the units one launch
    # processes.  Object sort: 8 B key+value read and written per entry and
    # pass (entries = 2n in slab mode); pair-sort pass: 8 B pair read and
    # written, P pairs (the local stage likewise: each pair read and written
    # once); sweep: 24 B per box streamed + 8 B per pair written.
    entries = 2 * n if tn["slabs"] else n
    obj_passes = -(-(24 if tn["slabs"] else 32) // tn["sort_bits"])  # slab keys: 24 bits
    alg_sweep = 24.0 * n + 8.0 * P
    alg_pass = 16.0 * P
    # Keys of 34 to 40 bits sorted in two passes with 8 pair regions: both
    # passes write 4-byte items (first pass: 8 B read + 4 B written per pair;
    # second: 4 B + 4 B) and the local stage reads them (4 B read + 8 B
    # written).
    x32 = tn["pair_local"] == 1 and tn["pair_regions"] == 8 and 2 * max(1, (n - 1).bit_length()) <= 40 \
        and 2 * max(1, (n - 1).bit_length()) > 32
    pass_bytes = [12.0 * P if x32 else alg_pass, 8.0 * P if x32 else alg_pass]
    alg_local = 12.0 * P if x32 else alg_pass
    # the dominant kernel: the longest single launch (pair-sort passes are
    # separate kernels: k_msplit, then k_xcount + k_xstarts + k_xsplit after
    # the sweep's regions; k_onesweep otherwise)
    if tn["pair_local"] == 1 and tn["pair_regions"] == 8 and len(pass_ms) == 2:
        pass_names = ["k_msplit", "k_xcount+k_xsplit"]
    else:
        pass_names = ["k_onesweep<pair>"] * len(pass_ms)
    cands = [("k_sweep", stages["sweep"], alg_sweep)]
    for i, (nm, t) in enumerate(zip(pass_names, pass_ms)):
        cands.append((nm, float(t), pass_bytes[min(i, 1)] if len(pass_ms) == 2 else alg_pass))
    if tn["pair_local"]:
        cands.append(("k_pair_local", stages["pair_local"], alg_local))
    roof_k, roof_ms, roof_b = max(cands, key=lambda c: c[1])
    achieved = roof_b / (roof_ms / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(roof_k),
            "kernel": roof_k, "bytes_per_launch": roof_b, "ms_per_launch": round(roof_ms, 4)}
    # The sweep is priced against HBM here, but what limits it is
    # instruction issue (its exact tests, hit masks and staging): the
    # counter-based issue fractions sit beside the HBM fraction.
    iss = issue_for(roof_k)
    if iss is not None:
        roof["issue"] = iss
        roof["limiter"] = "issue" if max(iss["valu"], iss["salu"]) > roof["frac"] else "hbm"
    obj_bytes = 16.0 * entries * obj_passes
    sort_gbs = obj_bytes / (stages["object_sort"] / 1e3) / 1e9
    sort_roof = {"achieved": round(sort_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(sort_gbs / HBM_PEAK_GBS, 4), "ms": round(stages["object_sort"], 4),
                 "bytes": obj_bytes, "passes": obj_passes, "entries": entries,
                 # what a plain device copy of one pass's bytes reaches at
                 # this size, and at the pair sort's size (1.2 GB)
                 "copy_floor_per_pass": copy_floor(torch, dev, 16.0 * entries),
                 "copy_large": copy_floor(torch, dev, 8.0 * P, reps=10)}
    pb = [pass_bytes[min(i, 1)] if len(pass_ms) == 2 else alg_pass for i in range(len(pass_ms))]
    pair_gbs = sum(pb) / (float(pass_ms.sum()) / 1e3) / 1e9
    pair_roof = {"achieved": round(pair_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(pair_gbs / HBM_PEAK_GBS, 4), "ms_per_pass": round(float(pass_ms.mean()), 4),
                 "ms_passes": [round(float(x), 4) for x in pass_ms],
                 "frac_passes": [round(b / (float(t) / 1e3) / 1e9 / HBM_PEAK_GBS, 4) for b, t in zip(pb, pass_ms)],
                 "passes": len(pass_ms), "bytes_passes": pb}
    if tn["pair_local"]:
        loc_gbs = alg_local / (stages["pair_local"] / 1e3) / 1e9
        pair_roof["local"] = {"achieved": round(loc_gbs, 1), "frac": round(loc_gbs / HBM_PEAK_GBS, 4),
                              "ms": round(stages["pair_local"], 4), "bytes": alg_local}
        pair_roof["ms_whole_sort"] = round(stages["pair_sort"], 4)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_c2(boxes, pairs[:P], P, args.cpu_runs)
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
                   "sweep": "slabs" if sap.tune["slabs"] else "plain",
                   "tune": {k: v for k, v in tune.items() if k != "from_cache"},
                   "sweep_tests": int(st.sweep_tests), "axis_used": [round(x, 6) for x in st.axis]},
        "roofline": roof,
        "sort_roofline": sort_roof,
        "pair_sort_roofline": pair_roof,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "wall_s": round(wall, 3),
        "cpu_baseline": cpu,
    }
    # The timed configuration's list is checked once, after timing, against
    # the default path on the same input (plain sweep, one pair region,
    # whole-key pair sort); any difference fails the benchmark.
    plain = G.Sap(max_n=n, axis_mode=mode)
    pairs_plain = torch.empty_like(pairs)
    P3, _, _ = plain.collide_into(bd, pairs_plain)
    identical = P3 == P and bool(torch.equal(pairs, pairs_plain))
    out["config"]["pairs_identical_to_default_path"] = identical
    if not identical:
        print(json.dumps(out))
        raise SystemExit("timed pair list differs from the default path's")
    del pairs_plain
    if not args.no_extras and slabs:
        plain.set_timing(True)
        ms_plain = _time_ms(torch, lambda: plain.collide_into(bd, pairs), max(1, args.steps // 2), 1)
        _, _, stp = plain.collide_into(bd, pairs)
        out["plain_sweep"] = {"ms_per_step": round(ms_plain, 4),
                              "updates_per_s": round(n / (ms_plain / 1e3), 1),
                              "sweep_tests": int(stp.sweep_tests),
                              "stages_ms": {k: round(v, 4) for k, v in
                                            stage_dict(np.array(plain.stage_times()),
                                                       plain.tune["pair_local"])[0].items()}}
        plain.close()
        if args.axis != "pca":  # the same slab sweep with the PCA axis pair
            pca = G.Sap(max_n=n, axis_mode=G.AXIS_PCA, flags=G.FLAG_SLABS)
            pca.set_tune(**{k: v for k, v in sap.tune.items() if k not in ("ms", "from_cache")})
            ms_pca = _time_ms(torch, lambda: pca.collide_into(bd, pairs), max(1, args.steps // 2), 1)
            _, _, stq = pca.collide_into(bd, pairs)
            out["pca_axis"] = {"ms_per_step": round(ms_pca, 4),
                               "updates_per_s": round(n / (ms_pca / 1e3), 1),
                               "sweep_tests": int(stq.sweep_tests),
                               "axis_used": [round(x, 6) for x in stq.axis]}
            pca.close()
    if not args.no_extras:
        out["c1"] = bench_c1(G, W, torch, dev, args.steps, args.warmup, not args.no_cpu_baseline,
                             args.tune_cache)
        out["batched"] = bench_batched(G, W, torch, dev, args.scenes, args.steps, args.warmup,
                                       args.tune_cache)
        out["learned_axis"] = bench_learned(G, W, torch, dev, args.learned_scenes)
        out["particle_step"] = bench_particles(G, W, torch, dev, args.prb_bodies, args.prb_steps,
                                               not args.no_cpu_baseline, args.tune_cache)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
