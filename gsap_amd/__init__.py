"""gsap_amd -- Python binding of libgsap (MI355X sweep-and-prune broadphase and
particle rigid-body step).

The product is the C ABI in include/gsap.h implemented by HIP kernels in
gsap_amd/csrc; this module only loads libgsap.so with ctypes and mirrors the
reference's host interface for tests and benchmarks:

  Sap               class Sap (physicboxSap2.16_Fermi/sapDetection/Sap.h:12-84)
  RadixSort         nvRadixSort::RadixSort (sapDetection/radixsort.h:110-278)
  ParticleRigidBody class ParticleRigidBody (ParticleRigidBody/ParticleRigidBody.h:46-165)

Device memory and streams come from PyTorch (plumbing only).  There is no CPU
fallback: every call goes through libgsap.so and raises if it is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsap.so")

OK, ERR_INVALID_ARG, ERR_CAPACITY, ERR_NONFINITE, ERR_HIP, ERR_UNSUPPORTED, ERR_OOM = range(7)
INPUT_AABB, INPUT_SPHERE = 0, 1
AXIS_PCA, AXIS_COORD_MAXVAR, AXIS_FIXED, AXIS_LEARNED, AXIS_AUTO = 0, 1, 2, 3, 4
OUT_CANON, OUT_SYM_CSR = 0, 1
FLAG_SKIP_SAME_GROUP = 1
FLAG_SLABS = 2
FLAG_COHERENT = 4
PRB_ABS_SPRING, PRB_WALLS, PRB_FLOOR_CLAMP, PRB_CAP_VEL, PRB_SKIP_SAME_BODY, PRB_INERTIA = (
    1, 2, 4, 8, 16, 32)

# Every symbol declared in include/gsap.h (checked by tests/test_abi.py).
EXPORTS = [
    "gsap_last_error", "gsap_version", "gsap_device_count",
    "gsap_sap_create", "gsap_sap_destroy", "gsap_sap_collide", "gsap_sap_collide_async",
    "gsap_sap_collide_batched", "gsap_sap_stage_times", "gsap_sap_set_timing",
    "gsap_radix_sort_create", "gsap_radix_sort_destroy", "gsap_radix_sort_u32",
    "gsap_radix_sort_f32",
    "gsap_prb_create", "gsap_prb_destroy", "gsap_prb_set_rb_data", "gsap_prb_set_particle_data",
    "gsap_prb_set_shape_inv_inertia", "gsap_prb_set_params", "gsap_prb_set_max_velocity",
    "gsap_prb_set_physics", "gsap_prb_update_particle_data", "gsap_prb_prepare",
    "gsap_prb_compute_force", "gsap_prb_update_velocity", "gsap_prb_integrate",
    "gsap_prb_collide", "gsap_prb_set_num_pairs", "gsap_prb_step", "gsap_prb_get_ptrs",
    "gsap_mlp_create", "gsap_mlp_destroy", "gsap_mlp_num_params", "gsap_mlp_set_params",
    "gsap_mlp_get_params", "gsap_mlp_forward", "gsap_mlp_loss_grad", "gsap_mlp_train",
    "gsap_axis_candidates", "gsap_sap_set_learned_model", "gsap_sap_features_batched",
    "gsap_sap_axis_tests_batched", "gsap_sap_get_tune", "gsap_sap_set_tune", "gsap_sap_autotune",
    "gsap_sap_query_workspace", "gsap_sap_set_workspace", "gsap_sap_collide_noalloc",
]
N_FEATURES, N_PARAMS, N_CANDIDATES = 64, 25348, 64
NOALLOC_SCRATCH_BYTES = 64 << 20  # GSAP_NOALLOC_SCRATCH_BYTES


class GsapError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gsap status {code}: {msg}")
        self.code = code


class SapDesc(ctypes.Structure):
    _fields_ = [("max_n", ctypes.c_uint32), ("input_kind", ctypes.c_int32),
                ("axis_mode", ctypes.c_int32), ("output_kind", ctypes.c_int32),
                ("fixed_axis", ctypes.c_float * 3), ("flags", ctypes.c_uint32),
                ("digit_bits", ctypes.c_int32), ("sort_items", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("num_pairs", ctypes.c_uint64), ("sweep_tests", ctypes.c_uint64),
                ("axis", ctypes.c_float * 3), ("axis_source", ctypes.c_int32),
                ("ms_total", ctypes.c_float), ("n", ctypes.c_uint32)]


class Tune(ctypes.Structure):
    _fields_ = [("sort_items", ctypes.c_int32), ("sort_bits", ctypes.c_int32),
                ("pair_items", ctypes.c_int32), ("pair_bits", ctypes.c_int32),
                ("pair_grid", ctypes.c_int32), ("slabs", ctypes.c_int32),
                ("sweep_chunk", ctypes.c_int32), ("pair_regions", ctypes.c_int32),
                ("pair_local", ctypes.c_int32), ("ms", ctypes.c_float),
                ("from_cache", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PrbPhysics(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float * 3), ("k_abs", ctypes.c_float),
                ("c_abs", ctypes.c_float), ("floor_scale", ctypes.c_float),
                ("wall_lo", ctypes.c_float * 2), ("wall_hi", ctypes.c_float * 2),
                ("flags", ctypes.c_uint32), ("broadphase_kind", ctypes.c_int32)]


class PrbPtrs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "rb_pos", "rb_quat", "rb_linvel", "rb_angvel", "rb_inv_inertia", "p_pos", "p_vel",
        "p_force", "p_mass", "p_body", "pairs", "pair_entries")] + [
        ("num_bodies", ctypes.c_int32), ("num_particles", ctypes.c_int32),
        ("num_pairs", ctypes.c_int64), ("num_pair_entries", ctypes.c_int64),
        ("rb_shape_idx", ctypes.c_void_p), ("rb_num_particles", ctypes.c_void_p)]


_lib = None


def lib():
    """Load libgsap.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P, i32, u32, u64, i64, f32 = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32,
                                  ctypes.c_uint64, ctypes.c_int64, ctypes.c_float)
    sig = {
        "gsap_last_error": (ctypes.c_char_p, []),
        "gsap_version": (ctypes.c_char_p, []),
        "gsap_device_count": (i32, []),
        "gsap_sap_create": (i32, [P, P]),
        "gsap_sap_destroy": (i32, [P]),
        "gsap_sap_collide": (i32, [P, P, P, u32, P, P, u64, P, P, P, P]),
        "gsap_sap_collide_async": (i32, [P, P, P, u32, P, P, u64, P, P]),
        "gsap_sap_collide_batched": (i32, [P, P, P, P, u32, P, u64, P, P, P]),
        "gsap_sap_stage_times": (i32, [P, P, i32, P]),
        "gsap_sap_set_timing": (i32, [P, i32]),
        "gsap_radix_sort_create": (i32, [u32, i32, P]),
        "gsap_radix_sort_destroy": (i32, [P]),
        "gsap_radix_sort_u32": (i32, [P, P, P, P, u32, u32]),
        "gsap_radix_sort_f32": (i32, [P, P, P, P, u32, u32, i32]),
        "gsap_prb_create": (i32, [i32, i32, i64, P]),
        "gsap_prb_destroy": (i32, [P]),
        "gsap_prb_set_rb_data": (i32, [P, i32] + [P] * 9 + [i32]),
        "gsap_prb_set_particle_data": (i32, [P, P, P, i32, i32]),
        "gsap_prb_set_shape_inv_inertia": (i32, [P, P, i32]),
        "gsap_prb_set_params": (i32, [P, f32, f32]),
        "gsap_prb_set_max_velocity": (i32, [P, f32, f32]),
        "gsap_prb_set_physics": (i32, [P, P]),
        "gsap_prb_update_particle_data": (i32, [P, P]),
        "gsap_prb_prepare": (i32, [P, P]),
        "gsap_prb_compute_force": (i32, [P, P, f32]),
        "gsap_prb_update_velocity": (i32, [P, P, f32]),
        "gsap_prb_integrate": (i32, [P, P, f32]),
        "gsap_prb_collide": (i32, [P, P, P]),
        "gsap_prb_set_num_pairs": (i32, [P, i64, i64]),
        "gsap_prb_step": (i32, [P, P, P, f32]),
        "gsap_prb_get_ptrs": (i32, [P, P]),
        "gsap_mlp_create": (i32, [u64, P]),
        "gsap_mlp_destroy": (i32, [P]),
        "gsap_mlp_num_params": (i32, []),
        "gsap_mlp_set_params": (i32, [P, P, u32]),
        "gsap_mlp_get_params": (i32, [P, P, u32]),
        "gsap_mlp_forward": (i32, [P, P, P, u32, P]),
        "gsap_mlp_loss_grad": (i32, [P, P, P, P, u32, f32, P, P]),
        "gsap_mlp_train": (i32, [P, P, P, P, u32, u32, u32, f32, f32, P]),
        "gsap_axis_candidates": (i32, [P]),
        "gsap_sap_set_learned_model": (i32, [P, P, f32]),
        "gsap_sap_features_batched": (i32, [P, P, P, P, u32, P]),
        "gsap_sap_axis_tests_batched": (i32, [P, P, P, P, u32, P, u32, P, P]),
        "gsap_sap_get_tune": (i32, [P, P]),
        "gsap_sap_set_tune": (i32, [P, P]),
        "gsap_sap_autotune": (i32, [P, P, P, u32, P, P, u32, ctypes.c_char_p, P]),
        "gsap_sap_query_workspace": (i32, [P, u32, u64, P]),
        "gsap_sap_set_workspace": (i32, [P, P, u64]),
        "gsap_sap_collide_noalloc": (i32, [P, P, P, u32, P, P, u64, P, P]),
    }
    for name, (res, args) in sig.items():
        if not hasattr(L, name):
            continue
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise GsapError(rc, lib().gsap_last_error().decode())
    return rc


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def device_count():
    return int(lib().gsap_device_count())


class Sap:
    """Sweep-and-prune broadphase handle (reference class Sap, Sap.h:12-84).

    collide(boxes) runs one complete update on device tensors and returns the
    device pair tensor (CANON: int32 [P,2] holding uint32 (i<j) pairs sorted
    lexicographically; SYM_CSR: [2P,2] directed pairs + row-start entries).
    """

    def __init__(self, max_n=0, input_kind=INPUT_AABB, axis_mode=AXIS_PCA, output_kind=OUT_CANON,
                 fixed_axis=None, flags=0, sort_items=0):
        d = SapDesc()
        d.max_n = max_n
        d.input_kind = input_kind
        d.axis_mode = axis_mode
        d.output_kind = output_kind
        if fixed_axis is not None:
            d.fixed_axis[:] = [float(v) for v in fixed_axis]
        d.flags = flags
        d.digit_bits = 0
        d.sort_items = sort_items
        self.desc = d
        self.h = ctypes.c_void_p()
        check(lib().gsap_sap_create(ctypes.byref(d), ctypes.byref(self.h)))
        self._spacing_seen = False

    def close(self):
        if self.h:
            lib().gsap_sap_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on=True):
        check(lib().gsap_sap_set_timing(self.h, 1 if on else 0))

    def stage_times(self):
        buf = (ctypes.c_float * 32)()
        n = ctypes.c_int32(0)
        check(lib().gsap_sap_stage_times(self.h, buf, 32, ctypes.byref(n)))
        return list(buf[:n.value])

    def count(self, boxes, group=None, stream=None):
        """Pre-count pass: number of pairs (CANON) or directed pairs (SYM_CSR)."""
        n = boxes.shape[0]
        npairs = ctypes.c_uint64(0)
        st = Stats()
        check(lib().gsap_sap_collide(self.h, _stream(stream), _ptr(boxes), n, _ptr(group), None, 0,
                                     None, ctypes.byref(npairs), None, ctypes.byref(st)))
        return int(npairs.value), st

    def collide_into(self, boxes, pairs, entries=None, group=None, stream=None):
        n = boxes.shape[0]
        npairs = ctypes.c_uint64(0)
        nent = ctypes.c_uint32(0)
        st = Stats()
        cap = pairs.shape[0] if pairs is not None else 0
        rc = lib().gsap_sap_collide(self.h, _stream(stream), _ptr(boxes), n, _ptr(group),
                                    _ptr(pairs), cap, _ptr(entries), ctypes.byref(npairs),
                                    ctypes.byref(nent), ctypes.byref(st))
        if rc == ERR_CAPACITY:  # the pair list or the caller's workspace is too small
            raise GsapError(rc, f"{lib().gsap_last_error().decode()} (pair capacity {cap}, "
                                f"pairs reported {npairs.value})")
        check(rc)
        return int(npairs.value), int(nent.value), st

    def collide_async(self, boxes, pairs, counts, entries=None, group=None, stream=None):
        """One update with no host synchronisation: the five counts (pairs,
        entries, sweep tests, flags, executed tests) land in the uint64 device
        tensor `counts` (a view of >= 5 int64 elements)."""
        n = boxes.shape[0]
        cap = pairs.shape[0] if pairs is not None else 0
        check(lib().gsap_sap_collide_async(self.h, _stream(stream), _ptr(boxes), n, _ptr(group),
                                           _ptr(pairs), cap, _ptr(entries), _ptr(counts)))

    def collide_noalloc(self, boxes, pairs, counts, entries=None, group=None, stream=None):
        """collide_async through the allocation-free entry point (scratch from
        the library's static device array; nothing allocated)."""
        n = boxes.shape[0]
        cap = pairs.shape[0] if pairs is not None else 0
        check(lib().gsap_sap_collide_noalloc(self.h, _stream(stream), _ptr(boxes), n, _ptr(group),
                                             _ptr(pairs), cap, _ptr(entries), _ptr(counts)))

    def collide(self, boxes, group=None, stream=None):
        """Pre-count, allocate, update.  Returns (pairs, entries, stats)."""
        import torch
        P, _ = self.count(boxes, group, stream)
        pairs = torch.empty((max(P, 1), 2), dtype=torch.int32, device=boxes.device)
        entries = None
        if self.desc.output_kind == OUT_SYM_CSR:
            entries = torch.empty(max(boxes.shape[0], 1), dtype=torch.int32, device=boxes.device)
        P2, ne, st = self.collide_into(boxes, pairs, entries, group, stream)
        assert P2 == P
        if entries is not None:
            return pairs[:P2], entries[:ne], st
        return pairs[:P2], None, st

    def collision(self, spacing, d_particle_data, pairs, pair_entries):
        """Reference Sap::collision(spacing, float4*, int2*, int*, &NPairs, &NPEntries)
        (Sap.cpp:211-335): spheres in, symmetric sorted pairs + entries out.
        `spacing` only seeded the reference's cell width; the single global
        sweep here has no cells, so it is accepted and unused."""
        del spacing
        P, ne, _ = self.collide_into(d_particle_data, pairs, pair_entries)
        return P, ne

    @property
    def tune(self):
        """The launch configuration (gsap_tune) as a dict."""
        t = Tune()
        check(lib().gsap_sap_get_tune(self.h, ctypes.byref(t)))
        return t.as_dict()

    def set_tune(self, **kw):
        t = Tune()
        check(lib().gsap_sap_get_tune(self.h, ctypes.byref(t)))
        names = {k for k, _ in Tune._fields_}
        for k, v in kw.items():
            if k not in names:
                raise ValueError(f"unknown tune field {k!r}")
            setattr(t, k, v)
        check(lib().gsap_sap_set_tune(self.h, ctypes.byref(t)))

    def autotune(self, boxes, scene_offsets=None, group=None, cache=None, stream=None):
        """Measure candidate configurations on this input and keep the fastest
        (gsap_sap_autotune); `cache` is an optional text file path."""
        import numpy as np
        t = Tune()
        off = None
        ns = 0
        if scene_offsets is not None:
            off = np.ascontiguousarray(scene_offsets, dtype=np.uint32)
            ns = len(off) - 1
        check(lib().gsap_sap_autotune(
            self.h, _stream(stream), _ptr(boxes), boxes.shape[0], _ptr(group),
            None if off is None else off.ctypes.data_as(ctypes.c_void_p), ns,
            None if cache is None else os.fsencode(cache), ctypes.byref(t)))
        return t.as_dict()

    def query_workspace(self, n, pair_capacity=0):
        """Device bytes an update of n objects with room for pair_capacity
        pairs needs under the current configuration."""
        b = ctypes.c_uint64(0)
        check(lib().gsap_sap_query_workspace(self.h, n, pair_capacity, ctypes.byref(b)))
        return int(b.value)

    def set_workspace(self, ws):
        """Carve every later update's scratch from the device tensor `ws`
        (kept referenced here); None returns to internal allocation."""
        self._ws = ws
        nbytes = 0 if ws is None else ws.numel() * ws.element_size()
        check(lib().gsap_sap_set_workspace(self.h, _ptr(ws), nbytes))

    def set_learned_model(self, model, conf_min=0.95):
        check(lib().gsap_sap_set_learned_model(self.h, model.h, conf_min))

    def features_batched(self, boxes, scene_offsets, stream=None):
        import numpy as np
        import torch
        off = np.ascontiguousarray(scene_offsets, dtype=np.uint32)
        ns = len(off) - 1
        f = torch.empty((ns, N_FEATURES), dtype=torch.float32, device=boxes.device)
        check(lib().gsap_sap_features_batched(self.h, _stream(stream), _ptr(boxes),
                                              off.ctypes.data_as(ctypes.c_void_p), ns, _ptr(f)))
        return f

    def axis_tests_batched(self, boxes, scene_offsets, axes, with_pairs=True, stream=None):
        import numpy as np
        import torch
        off = np.ascontiguousarray(scene_offsets, dtype=np.uint32)
        ax = np.ascontiguousarray(axes, dtype=np.float32).reshape(-1, 3)
        ns = len(off) - 1
        tests = torch.empty((ns, len(ax)), dtype=torch.int64, device=boxes.device)
        pairs = torch.empty(ns, dtype=torch.int64, device=boxes.device) if with_pairs else None
        check(lib().gsap_sap_axis_tests_batched(self.h, _stream(stream), _ptr(boxes),
                                                off.ctypes.data_as(ctypes.c_void_p), ns,
                                                ax.ctypes.data_as(ctypes.c_void_p), len(ax),
                                                _ptr(tests), _ptr(pairs)))
        return tests, pairs

    def collide_batched(self, boxes, scene_offsets, pairs, scene_pair_offsets, stream=None):
        import numpy as np
        off = np.ascontiguousarray(scene_offsets, dtype=np.uint32)
        ns = len(off) - 1
        npairs = ctypes.c_uint64(0)
        st = Stats()
        cap = pairs.shape[0] if pairs is not None else 0
        rc = lib().gsap_sap_collide_batched(
            self.h, _stream(stream), _ptr(boxes), off.ctypes.data_as(ctypes.c_void_p), ns,
            _ptr(pairs), cap, _ptr(scene_pair_offsets), ctypes.byref(npairs), ctypes.byref(st))
        if rc == ERR_CAPACITY:
            raise GsapError(rc, f"{lib().gsap_last_error().decode()} (pair capacity {cap}, "
                                f"pairs reported {npairs.value})")
        check(rc)
        return int(npairs.value), st


def axis_candidates():
    """The 64 fixed candidate sweep axes (64 x 3 float32), coordinate axes first."""
    import numpy as np
    a = np.zeros((N_CANDIDATES, 3), np.float32)
    check(lib().gsap_axis_candidates(a.ctypes.data_as(ctypes.c_void_p)))
    return a


class AxisModel:
    """Learned sweep-axis selector: 64-128-128-4 MLP (bf16 weights and
    activations, fp32 accumulation) run and trained by libgsap's HIP kernels."""

    def __init__(self, seed=0):
        self.h = ctypes.c_void_p()
        check(lib().gsap_mlp_create(seed, ctypes.byref(self.h)))

    def params(self):
        import numpy as np
        p = np.zeros(N_PARAMS, np.float32)
        check(lib().gsap_mlp_get_params(self.h, p.ctypes.data_as(ctypes.c_void_p), N_PARAMS))
        return p

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, np.float32)
        check(lib().gsap_mlp_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p), len(p)))

    def save(self, path):
        self.params().tofile(path)

    @classmethod
    def load(cls, path):
        import numpy as np
        m = cls()
        m.set_params(np.fromfile(path, np.float32))
        return m

    def forward(self, x, stream=None):
        import torch
        y = torch.empty((x.shape[0], 4), dtype=torch.float32, device=x.device)
        check(lib().gsap_mlp_forward(self.h, _stream(stream), _ptr(x), x.shape[0], _ptr(y)))
        return y

    def loss_grad(self, x, t, lambda_count=0.05, stream=None):
        import numpy as np
        g = np.zeros(N_PARAMS, np.float32)
        loss = ctypes.c_float(0)
        check(lib().gsap_mlp_loss_grad(self.h, _stream(stream), _ptr(x), _ptr(t), x.shape[0],
                                       lambda_count, ctypes.byref(loss),
                                       g.ctypes.data_as(ctypes.c_void_p)))
        return float(loss.value), g

    def train(self, x, t, epochs=10, batch=1000, lr=1e-3, lambda_count=0.05, stream=None):
        import numpy as np
        losses = np.zeros(epochs, np.float32)
        check(lib().gsap_mlp_train(self.h, _stream(stream), _ptr(x), _ptr(t), x.shape[0], epochs,
                                   batch, lr, lambda_count,
                                   losses.ctypes.data_as(ctypes.c_void_p)))
        return losses

    def close(self):
        if self.h:
            lib().gsap_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RadixSort:
    """nvRadixSort::RadixSort(maxElements, keysOnly) (radixsort.h:110-278)."""

    def __init__(self, max_elements, keys_only=False):
        self.h = ctypes.c_void_p()
        check(lib().gsap_radix_sort_create(max_elements, 1 if keys_only else 0,
                                           ctypes.byref(self.h)))

    def sort(self, keys, values=None, key_bits=32, stream=None):
        import torch
        n = keys.shape[0]
        if keys.dtype == torch.float32:
            check(lib().gsap_radix_sort_f32(self.h, _stream(stream), _ptr(keys), _ptr(values), n,
                                            key_bits, 1))
        else:
            check(lib().gsap_radix_sort_u32(self.h, _stream(stream), _ptr(keys), _ptr(values), n,
                                            key_bits))

    def close(self):
        if self.h:
            lib().gsap_radix_sort_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticleRigidBody:
    """ParticleRigidBody(maxNumBodies, maxNumParticles, maxPairs)
    (ParticleRigidBody.h:46-165) over gsap_prb_*; setters take host numpy
    arrays like the reference's, steps run on the current torch stream."""

    def __init__(self, max_bodies, max_particles, max_pairs):
        self.h = ctypes.c_void_p()
        check(lib().gsap_prb_create(max_bodies, max_particles, max_pairs, ctypes.byref(self.h)))

    @staticmethod
    def _np(a, dt):
        import numpy as np
        return np.ascontiguousarray(a, dtype=dt)

    def set_rb_data(self, pos, quat, linvel, angvel, shape, pstart, npart, inv_mass,
                    inv_inertia=None):
        import numpy as np
        f, i = np.float32, np.int32
        arrs = [self._np(pos, f), self._np(quat, f), self._np(linvel, f), self._np(angvel, f),
                self._np(shape, i), self._np(pstart, i), self._np(npart, i),
                self._np(inv_mass, f)]
        ii = None if inv_inertia is None else self._np(inv_inertia, f)
        self._keep = arrs + [ii]
        p = [a.ctypes.data_as(ctypes.c_void_p) for a in arrs]
        check(lib().gsap_prb_set_rb_data(self.h, len(arrs[0]), *p,
                                         None if ii is None else ii.ctypes.data_as(ctypes.c_void_p),
                                         int(arrs[6].sum())))

    def set_particle_data(self, ref_pos, ref_mass, max_per_shape):
        import numpy as np
        rp, rm = self._np(ref_pos, np.float32), self._np(ref_mass, np.float32)
        check(lib().gsap_prb_set_particle_data(self.h, rp.ctypes.data_as(ctypes.c_void_p),
                                               rm.ctypes.data_as(ctypes.c_void_p), len(rm),
                                               max_per_shape))

    def set_particle_shape_inv_inertia(self, inv_inertia):
        import numpy as np
        a = self._np(inv_inertia, np.float32).reshape(-1, 12)
        check(lib().gsap_prb_set_shape_inv_inertia(self.h, a.ctypes.data_as(ctypes.c_void_p),
                                                   len(a)))

    def set_params(self, sp, dp):
        check(lib().gsap_prb_set_params(self.h, sp, dp))

    def set_max_velocity(self, min_radius, dt):
        check(lib().gsap_prb_set_max_velocity(self.h, min_radius, dt))

    def set_physics(self, gravity=(0, -5, 0), k_abs=1e4, c_abs=50.0, floor_scale=0.1,
                    wall_lo=(0, 0), wall_hi=(256, 256), flags=0, broadphase_kind=INPUT_SPHERE):
        p = PrbPhysics()
        p.gravity[:] = [float(g) for g in gravity]
        p.k_abs, p.c_abs, p.floor_scale = k_abs, c_abs, floor_scale
        p.wall_lo[:] = [float(v) for v in wall_lo]
        p.wall_hi[:] = [float(v) for v in wall_hi]
        p.flags = flags
        p.broadphase_kind = broadphase_kind
        check(lib().gsap_prb_set_physics(self.h, ctypes.byref(p)))

    def update_particle_data(self, stream=None):
        check(lib().gsap_prb_update_particle_data(self.h, _stream(stream)))

    def prepare(self, stream=None):
        check(lib().gsap_prb_prepare(self.h, _stream(stream)))

    def comupteFoceOnParticles(self, dt, stream=None):  # reference spelling
        check(lib().gsap_prb_compute_force(self.h, _stream(stream), dt))

    compute_force = comupteFoceOnParticles

    def update_velocity(self, dt, stream=None):
        check(lib().gsap_prb_update_velocity(self.h, _stream(stream), dt))

    def integrate(self, dt, stream=None):
        check(lib().gsap_prb_integrate(self.h, _stream(stream), dt))

    def collide(self, sap, stream=None):
        check(lib().gsap_prb_collide(self.h, _stream(stream), sap.h))

    def set_num_pairs(self, num_pairs, num_entries):
        check(lib().gsap_prb_set_num_pairs(self.h, num_pairs, num_entries))

    def step(self, sap, dt, stream=None):
        check(lib().gsap_prb_step(self.h, _stream(stream), sap.h, dt))

    def ptrs(self):
        p = PrbPtrs()
        check(lib().gsap_prb_get_ptrs(self.h, ctypes.byref(p)))
        return p

    def close(self):
        if self.h:
            lib().gsap_prb_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
