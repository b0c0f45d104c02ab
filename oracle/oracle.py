"""ctypes wrapper of the CPU oracle (oracle/gsap_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker or the timed CPU
baseline -- never the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

KIND_AABB, KIND_SPHERE = 0, 1
AXIS_PCA, AXIS_COORD_MAXVAR, AXIS_FIXED, AXIS_AUTO = 0, 1, 2, 4

PRB_ABS_SPRING, PRB_WALLS, PRB_FLOOR_CLAMP, PRB_CAP_VEL, PRB_SKIP_SAME_BODY, PRB_INERTIA = (
    1, 2, 4, 8, 16, 32)


def build():
    subprocess.run(["make", "-C", _HERE, "-B", "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        u64, i64, sz = ctypes.c_uint64, ctypes.c_int64, ctypes.c_size_t
        L.or_flip.restype = ctypes.c_uint32
        L.or_flip.argtypes = [ctypes.c_float]
        L.or_unflip.restype = ctypes.c_float
        L.or_unflip.argtypes = [ctypes.c_uint32]
        L.or_flip_array.argtypes = [P, P, sz]
        L.or_moments.argtypes = [P, ctypes.c_int, sz, P]
        L.or_pca_axis.argtypes = [P, sz, P, P]
        L.or_axis_coord_maxvar.restype = ctypes.c_int
        L.or_axis_coord_maxvar.argtypes = [P, sz, P, P]
        L.or_extents.argtypes = [P, ctypes.c_int, sz, P, P]
        L.or_slab_axis.argtypes = [P, sz, P, P]
        L.or_axis_cost.restype = ctypes.c_double
        L.or_axis_cost.argtypes = [P, P, P, sz, P, P]
        L.or_axis_auto.restype = ctypes.c_int
        L.or_axis_auto.argtypes = [P, P, P, sz, ctypes.c_int, P, P]
        L.or_project.argtypes = [P, ctypes.c_int, sz, P, P, P]
        L.or_radix_sort_kv.argtypes = [P, P, sz, ctypes.c_int]
        L.or_radix_sort_pairs.argtypes = [P, sz, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_sap.restype = i64
        L.or_sap.argtypes = [P, ctypes.c_int, sz, ctypes.c_int, P, P, P, u64, P, P, ctypes.c_uint32]
        L.or_last_times.argtypes = [P]
        L.or_sweep_tests.restype = u64
        L.or_sweep_tests.argtypes = [P, ctypes.c_int, sz, P]
        L.or_sap_ref_cells.restype = ctypes.c_int64
        L.or_sap_ref_cells.argtypes = [P, ctypes.c_int, sz, ctypes.c_int, P, u64, P]
        L.or_brute.restype = i64
        L.or_brute.argtypes = [P, ctypes.c_int, sz, P, P, u64]
        L.or_canon_sort.argtypes = [P, u64]
        L.or_sym_csr.restype = u64
        L.or_sym_csr.argtypes = [P, u64, P, P, P]
        L.or_prb_integrate.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_float, P]
        L.or_prb_update_particles.argtypes = [P] * 9 + [ctypes.c_int, P, P, P, ctypes.c_int]
        L.or_prb_prepare.argtypes = [P, P, P, P, ctypes.c_int]
        L.or_prb_forces.argtypes = [P, P, P, P, u64, P, P, ctypes.c_int, ctypes.c_float, P]
        L.or_prb_update_velocity.argtypes = [P] * 9 + [ctypes.c_int, P, ctypes.c_int,
                                                       ctypes.c_float, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == cols
    return a


def _boxes(boxes, kind):
    return _f32(boxes, 6 if kind == KIND_AABB else 4)


def flip(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, np.uint32)
    lib().or_flip_array(_p(x), _p(out), x.size)
    return out


def moments(boxes, kind=KIND_AABB):
    b = _boxes(boxes, kind)
    out = np.zeros(9, np.float64)
    lib().or_moments(_p(b), kind, len(b), _p(out))
    return out


def pca_axis(mom, n):
    mom = np.ascontiguousarray(mom, np.float64)
    ax = np.zeros(3, np.float32)
    ev = np.zeros(3, np.float64)
    lib().or_pca_axis(_p(mom), n, _p(ax), _p(ev))
    return ax, ev


def coord_axis(mom, n):
    mom = np.ascontiguousarray(mom, np.float64)
    ax = np.zeros(3, np.float32)
    lam = np.zeros(3, np.int32)
    a = lib().or_axis_coord_maxvar(_p(mom), n, _p(ax), _p(lam))
    return a, ax, lam


def extents(boxes, kind=KIND_AABB):
    """(sum, max) of the per-coordinate box extents, fp64."""
    b = _boxes(boxes, kind)
    xs, xm = np.zeros(3, np.float64), np.zeros(3, np.float64)
    lib().or_extents(_p(b), kind, len(b), _p(xs), _p(xm))
    return xs, xm


def slab_axis(mom, n, u):
    mom = np.ascontiguousarray(mom, np.float64)
    u = np.ascontiguousarray(u, np.float32)
    v = np.zeros(3, np.float32)
    lib().or_slab_axis(_p(mom), n, _p(u), _p(v))
    return v


def axis_cost(mom, xs, xm, n, u, v=None):
    a = [np.ascontiguousarray(x, np.float64) for x in (mom, xs, xm)]
    u = np.ascontiguousarray(u, np.float32)
    v = None if v is None else np.ascontiguousarray(v, np.float32)
    return lib().or_axis_cost(_p(a[0]), _p(a[1]), _p(a[2]), n, _p(u), _p(v))


def axis_auto(boxes, kind=KIND_AABB, slabs=False):
    """(source, u, v): source 0 = PCA kept, 1 = coordinate axis (pair)."""
    mom = moments(boxes, kind)
    xs, xm = extents(boxes, kind)
    u, v = np.zeros(3, np.float32), np.zeros(3, np.float32)
    src = lib().or_axis_auto(_p(mom), _p(xs), _p(xm), len(boxes), int(slabs), _p(u), _p(v))
    return src, u, (v if slabs else None)


def project(boxes, axis, kind=KIND_AABB):
    b = _boxes(boxes, kind)
    ax = np.ascontiguousarray(axis, np.float32)
    keys = np.empty(len(b), np.uint32)
    hik = np.empty(len(b), np.uint32)
    lib().or_project(_p(b), kind, len(b), _p(ax), _p(keys), _p(hik))
    return keys, hik


def radix_sort_kv(keys, vals, digit_bits=8):
    k = np.array(keys, dtype=np.uint32, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    lib().or_radix_sort_kv(_p(k), _p(v), len(k), digit_bits)
    return k, v


def radix_sort_pairs(pairs, bbits, key_bits, digit_bits=8):
    p = np.array(pairs, dtype=np.uint32, copy=True).reshape(-1, 2)
    lib().or_radix_sort_pairs(_p(p), len(p), bbits, key_bits, digit_bits)
    return p


def sap(boxes, kind=KIND_AABB, axis_mode=AXIS_PCA, fixed_axis=None, group=None, cap=None,
        sample_stride=1):
    """Full reference-semantics SaP update: (pairs[P,2] canonical, tests, axis)."""
    b = _boxes(boxes, kind)
    n = len(b)
    fa = None if fixed_axis is None else np.ascontiguousarray(fixed_axis, np.float32)
    g = None if group is None else np.ascontiguousarray(group, np.uint32)
    if cap is None:
        cap = max(16, 64 * n)
    while True:
        pairs = np.empty((cap, 2), np.uint32)
        tests = ctypes.c_uint64(0)
        ax = np.zeros(3, np.float32)
        P = lib().or_sap(_p(b), kind, n, axis_mode, _p(fa), _p(g), _p(pairs), cap,
                         ctypes.byref(tests), _p(ax), sample_stride)
        if P <= cap:
            return pairs[:P].copy(), int(tests.value), ax
        cap = int(P)


def last_times():
    """(setup_s, sweep_s, canon_s) of the last sap() call."""
    t = np.zeros(3, np.float64)
    lib().or_last_times(_p(t))
    return tuple(float(x) for x in t)


def sap_ref_cells(boxes, kind=KIND_AABB, count_only=False, cap=None):
    """The reference's 4-cell sweep + cross pass (Sap.cpp:211-295): (canonical
    pairs or None when count_only, {cell_tests, cross_tests, relevant,
    raw_pairs})."""
    b = _boxes(boxes, kind)
    n = len(b)
    stats = np.zeros(4, np.uint64)
    if count_only:
        lib().or_sap_ref_cells(_p(b), kind, n, 1, None, 0, _p(stats))
        pairs = None
    else:
        cap = max(16, 64 * n) if cap is None else cap
        while True:
            buf = np.empty((cap, 2), np.uint32)
            P = lib().or_sap_ref_cells(_p(b), kind, n, 0, _p(buf), cap, _p(stats))
            if int(stats[3]) <= cap:
                pairs = buf[:P].copy()
                break
            cap = int(stats[3])
    keys = ("cell_tests", "cross_tests", "relevant", "raw_pairs")
    return pairs, {k: int(v) for k, v in zip(keys, stats)}


def sweep_tests(boxes, axis, kind=KIND_AABB):
    b = _boxes(boxes, kind)
    ax = np.ascontiguousarray(axis, np.float32)
    return int(lib().or_sweep_tests(_p(b), kind, len(b), _p(ax)))


def brute(boxes, kind=KIND_AABB, group=None, cap=None):
    b = _boxes(boxes, kind)
    g = None if group is None else np.ascontiguousarray(group, np.uint32)
    if cap is None:
        cap = max(16, 64 * len(b))
    while True:
        pairs = np.empty((cap, 2), np.uint32)
        P = lib().or_brute(_p(b), kind, len(b), _p(g), _p(pairs), cap)
        if P <= cap:
            return pairs[:P].copy()
        cap = int(P)


def canon_sort(pairs):
    p = np.array(pairs, dtype=np.uint32, copy=True).reshape(-1, 2)
    lib().or_canon_sort(_p(p), len(p))
    return p


def sym_csr(canon):
    c = np.ascontiguousarray(canon, np.uint32).reshape(-1, 2)
    P = len(c)
    out = np.empty((max(2 * P, 1), 2), np.int32)
    ent = np.empty(max(2 * P, 1), np.int32)
    ne = ctypes.c_uint32(0)
    lib().or_sym_csr(_p(c), P, _p(out), _p(ent), ctypes.byref(ne))
    return out[:2 * P].copy(), ent[:ne.value].copy()


class PrbParams(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float * 3), ("sp", ctypes.c_float), ("dp", ctypes.c_float),
                ("k_abs", ctypes.c_float), ("c_abs", ctypes.c_float),
                ("floor_scale", ctypes.c_float), ("wall_lo", ctypes.c_float * 2),
                ("wall_hi", ctypes.c_float * 2), ("max_lin_vel", ctypes.c_float),
                ("flags", ctypes.c_uint32)]


def prb_params(gravity=(0.0, -5.0, 0.0), sp=0.5, dp=0.25, k_abs=1e4, c_abs=50.0, floor_scale=0.1,
               wall_lo=(0.0, 0.0), wall_hi=(256.0, 256.0), max_lin_vel=1e10, flags=0):
    p = PrbParams()
    p.gravity[:] = [float(g) for g in gravity]
    p.sp, p.dp, p.k_abs, p.c_abs = sp, dp, k_abs, c_abs
    p.floor_scale = floor_scale
    p.wall_lo[:] = list(wall_lo)
    p.wall_hi[:] = list(wall_hi)
    p.max_lin_vel = max_lin_vel
    p.flags = flags
    return p


class PrbState:
    """Host mirror of the particle rigid-body state, stepped by the oracle."""

    def __init__(self, pos, quat, linvel, angvel, shape, pstart, npart, inv_mass, ref_pos,
                 ref_mass, max_per_shape, shape_inv_inertia=None):
        f = lambda a: np.array(a, dtype=np.float32, copy=True)
        i = lambda a: np.array(a, dtype=np.int32, copy=True)
        self.pos, self.quat, self.linvel, self.angvel = f(pos), f(quat), f(linvel), f(angvel)
        self.shape, self.pstart, self.npart = i(shape), i(pstart), i(npart)
        self.inv_mass = f(inv_mass)
        self.ref_pos, self.ref_mass = f(ref_pos), f(ref_mass)
        self.max_per_shape = int(max_per_shape)
        nb = len(self.pos)
        self.nb = nb
        self.np_ = int(self.npart.sum())
        self.shape_inv_inertia = (np.zeros((max(1, int(self.shape.max()) + 1), 12), np.float32)
                                  if shape_inv_inertia is None else f(shape_inv_inertia))
        self.inv_inertia = np.zeros((nb, 12), np.float32)
        self.ppos = np.zeros((self.np_, 4), np.float32)
        self.pvel = np.zeros((self.np_, 4), np.float32)
        self.pmass = np.zeros(self.np_, np.float32)
        self.pforce = np.zeros((self.np_, 4), np.float32)
        self.pbody = np.repeat(np.arange(nb, dtype=np.int32), self.npart)

    def integrate(self, dt, params):
        lib().or_prb_integrate(_p(self.pos), _p(self.quat), _p(self.linvel), _p(self.angvel),
                               self.nb, dt, ctypes.byref(params))

    def update_particles(self):
        lib().or_prb_update_particles(_p(self.pos), _p(self.quat), _p(self.linvel),
                                      _p(self.angvel), _p(self.shape), _p(self.pstart),
                                      _p(self.npart), _p(self.ref_pos), _p(self.ref_mass),
                                      self.max_per_shape, _p(self.ppos), _p(self.pvel),
                                      _p(self.pmass), self.nb)

    def prepare(self):
        lib().or_prb_prepare(_p(self.quat), _p(self.shape), _p(self.shape_inv_inertia),
                             _p(self.inv_inertia), self.nb)

    def forces(self, sym_pairs, dt, params):
        sp = np.ascontiguousarray(sym_pairs, np.int32).reshape(-1, 2)
        lib().or_prb_forces(_p(self.ppos), _p(self.pvel), _p(self.pmass), _p(sp), len(sp),
                            _p(self.pbody), _p(self.pforce), self.np_, dt, ctypes.byref(params))

    def update_velocity(self, dt, params):
        lib().or_prb_update_velocity(_p(self.linvel), _p(self.angvel), _p(self.quat),
                                     _p(self.inv_mass), _p(self.inv_inertia), _p(self.shape),
                                     _p(self.pstart), _p(self.npart), _p(self.ref_pos),
                                     self.max_per_shape, _p(self.pforce), self.nb, dt,
                                     ctypes.byref(params))

    def particle_aabbs(self):
        r = self.ppos[:, 3:4]
        return np.concatenate([self.ppos[:, :3] - r, self.ppos[:, :3] + r], axis=1)

    def step(self, dt, params, broadphase="sphere", sample_stride=1):
        """DemoParticles::step order (PB/cudaPhysics/demoParticles.h:856-928).
        sample_stride > 1 sweeps only every k-th sorted object (a timing
        sample: the pair list and so the forces are then incomplete)."""
        self.integrate(dt, params)
        self.update_particles()
        grp = self.pbody.astype(np.uint32) if params.flags & PRB_SKIP_SAME_BODY else None
        if broadphase == "sphere":
            canon = brute_or_sap(self.ppos, KIND_SPHERE, grp, sample_stride)
        else:
            canon = brute_or_sap(self.particle_aabbs(), KIND_AABB, grp, sample_stride)
        sym, _ = sym_csr(canon)
        self.prepare()
        self.forces(sym, dt, params)
        self.update_velocity(dt, params)
        return canon


def brute_or_sap(boxes, kind, group, sample_stride=1):
    cap = None if sample_stride == 1 else max(16, 64 * len(boxes) // sample_stride)
    pairs, _, _ = sap(boxes, kind=kind, axis_mode=AXIS_COORD_MAXVAR, group=group, cap=cap,
                      sample_stride=sample_stride)
    return pairs


# ---------------------------------------------------------------- CPU baseline
_CPU_LIBS = {}


def cpu_sap_lib(path=None):
    """The multi-threaded CPU restatement (oracle/cpu_sap.c): the prebuilt
    portable library, or `path` (e.g. a -march=native build of the host)."""
    path = path or os.path.join(_HERE, "_build", "libcpusap.so")
    if path not in _CPU_LIBS:
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.cpu_sap.restype = ctypes.c_int64
        L.cpu_sap.argtypes = [P, ctypes.c_size_t, P, ctypes.c_int, P, ctypes.c_uint64, P, P, P]
        L.cpu_sap_max_threads.restype = ctypes.c_int
        _CPU_LIBS[path] = L
    return _CPU_LIBS[path]


def build_native_cpu_sap(out_dir):
    """Compile oracle/cpu_sap.c for this host (-O3 -march=native
    -ffp-contract=off -fopenmp) into out_dir; returns the library path, or
    None when the compiler is unavailable."""
    out = os.path.join(out_dir, "libcpusap_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
           "-fopenmp", "-fPIC", "-shared", "-o", out, os.path.join(_HERE, "cpu_sap.c"), "-lm"]
    try:
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    except (OSError, subprocess.SubprocessError):
        return None
    return out


def cpu_sap(boxes, threads, group=None, cap=None, lib_path=None):
    """One CPU-baseline update of AABBs: (canonical pairs, sweep tests, axis,
    {axis_project, sort, sweep, canon} seconds)."""
    L = cpu_sap_lib(lib_path)
    b = _boxes(boxes, KIND_AABB)
    g = None if group is None else np.ascontiguousarray(group, np.uint32)
    n = len(b)
    cap = max(16, 64 * n) if cap is None else cap
    tests = np.zeros(1, np.uint64)
    ax = ctypes.c_int(0)
    t = np.zeros(4, np.float64)
    while True:
        pairs = np.empty((cap, 2), np.uint32)
        P = L.cpu_sap(_p(b), n, _p(g), int(threads), _p(pairs), cap, _p(tests), ctypes.byref(ax), _p(t))
        if P <= cap:
            break
        cap = int(P)
    keys = ("axis_project", "sort", "sweep", "canon")
    return pairs[:P], int(tests[0]), int(ax.value), dict(zip(keys, t.tolist()))
