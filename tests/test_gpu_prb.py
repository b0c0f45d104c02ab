"""GPU parity of the particle rigid-body step against the CPU restatement.
Both sides evaluate the same fp32 operations in the same order with
contraction off, so the expected difference is zero; the stated tolerance is
max |dx| <= 1e-5 * (scene extent) and max |dq| <= 1e-5 per step horizon."""
import ctypes

import numpy as np
import pytest

import gsap_amd as G
from gsap_amd import workloads as W
from oracle import oracle as O
from tests.gpu_util import d2h

pytestmark = pytest.mark.gpu

TOL_X = 1e-5  # relative to the scene extent
TOL_Q = 1e-5


def make(nb, seed, box, height, flags, kind, gravity=(0, -9.81, 0)):
    sc = W.bodies(nb, seed=seed, box=box, height=height)
    prb = G.ParticleRigidBody(nb, nb * 27, nb * 27 * 40)
    prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"], sc["pstart"],
                    sc["npart"], sc["inv_mass"], sc["inv_inertia"])
    prb.set_particle_data(sc["ref_pos"], sc["ref_mass"], sc["max_per_shape"])
    prb.set_particle_shape_inv_inertia(sc["shape_inv_inertia"])
    prb.set_physics(gravity=gravity, k_abs=1e4, c_abs=50.0, floor_scale=1.0, wall_lo=(0, 0),
                    wall_hi=(box, box), flags=flags, broadphase_kind=kind)
    sap = G.Sap(input_kind=kind, axis_mode=G.AXIS_PCA, output_kind=G.OUT_SYM_CSR,
                flags=G.FLAG_SKIP_SAME_GROUP if flags & G.PRB_SKIP_SAME_BODY else 0)
    # host mirror stepped by the oracle; quaternions normalised like setRbData
    q = sc["quat"].copy()
    for i in range(nb):
        v = q[i]
        l = np.float32(np.sqrt(np.float32(np.float32(np.float32(v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3])))
        q[i] = v * np.float32(np.float32(1.0) / l)
    ref = O.PrbState(sc["pos"], q, sc["linvel"], sc["angvel"], sc["shape"], sc["pstart"],
                     sc["npart"], sc["inv_mass"], sc["ref_pos"], sc["ref_mass"],
                     sc["max_per_shape"], sc["shape_inv_inertia"])
    params = O.prb_params(gravity=gravity, k_abs=1e4, c_abs=50.0, floor_scale=1.0,
                          wall_lo=(0, 0), wall_hi=(box, box), flags=flags)
    return prb, sap, ref, params


@pytest.mark.parametrize("kind", [G.INPUT_SPHERE, G.INPUT_AABB])
def test_step_parity_10_steps(dev, kind):
    import torch
    nb, box = 300, 24.0
    flags = G.PRB_ABS_SPRING | G.PRB_WALLS | G.PRB_SKIP_SAME_BODY
    prb, sap, ref, params = make(nb, 5, box, 12.0, flags, kind)
    dt = 1e-4
    bp = "sphere" if kind == G.INPUT_SPHERE else "aabb"
    worst_x = worst_q = 0.0
    for step in range(10):
        prb.step(sap, dt)
        canon = ref.step(dt, params, broadphase=bp)
        torch.cuda.synchronize()
        p = prb.ptrs()
        assert p.num_pairs == 2 * len(canon), (step, p.num_pairs, len(canon))
        pos = d2h(p.rb_pos, (nb, 4), np.float32)
        quat = d2h(p.rb_quat, (nb, 4), np.float32)
        worst_x = max(worst_x, float(np.max(np.abs(pos[:, :3] - ref.pos[:, :3]))) / box)
        worst_q = max(worst_q, float(np.max(np.abs(quat - ref.quat))))
    assert worst_x <= TOL_X and worst_q <= TOL_Q, (worst_x, worst_q)
    assert len(canon) > 0  # bodies were in contact
    print(f"max |dx|/extent {worst_x:.3e}  max |dq| {worst_q:.3e}")


def test_reference_mode_floor_and_clamp(dev):
    """Reference constants: g = -5, mass/dt-derived springs, 0.1 floor scale,
    y clamp; no same-body filter (particles spaced 2r*1.001 like the
    reference's templates)."""
    import torch
    nb = 64
    sc = W.bodies(nb, seed=9, box=40.0, height=6.0)
    ref_pos = sc["ref_pos"].copy()
    ref_pos[:, :3] *= 1.001
    prb = G.ParticleRigidBody(nb, nb * 27, nb * 27 * 40)
    prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"], sc["pstart"],
                    sc["npart"], sc["inv_mass"], sc["inv_inertia"])
    prb.set_particle_data(ref_pos, sc["ref_mass"], 27)
    prb.set_particle_shape_inv_inertia(sc["shape_inv_inertia"])
    prb.set_params(2.1, 0.8)
    flags = G.PRB_FLOOR_CLAMP
    prb.set_physics(gravity=(0, -5, 0), floor_scale=0.1, flags=flags,
                    broadphase_kind=G.INPUT_SPHERE)
    sap = G.Sap(input_kind=G.INPUT_SPHERE, axis_mode=G.AXIS_COORD_MAXVAR, output_kind=G.OUT_SYM_CSR)
    q = sc["quat"] / np.linalg.norm(sc["quat"], axis=1, keepdims=True)
    ref = O.PrbState(sc["pos"], q.astype(np.float32), sc["linvel"], sc["angvel"], sc["shape"],
                     sc["pstart"], sc["npart"], sc["inv_mass"], ref_pos, sc["ref_mass"], 27,
                     sc["shape_inv_inertia"])
    params = O.prb_params(gravity=(0, -5, 0), sp=2.1, dp=0.8, floor_scale=0.1, flags=flags)
    # start both from the GPU's normalised quaternions
    p = prb.ptrs()
    ref.quat[:] = d2h(p.rb_quat, (nb, 4), np.float32)
    for _ in range(5):
        prb.step(sap, 0.025)
        ref.step(0.025, params, broadphase="sphere")
    torch.cuda.synchronize()
    p = prb.ptrs()
    pos = d2h(p.rb_pos, (nb, 4), np.float32)
    assert np.max(np.abs(pos[:, :3] - ref.pos[:, :3])) <= 1e-5 * 40.0
    assert np.all(pos[:, 1] >= 0.25)


def _ref_mode_scene(nb, seed, spread_v):
    """Cubes at the reference's particle spacing 2r*1.001, random initial
    linear and angular velocities (|v| up to spread_v)."""
    sc = W.bodies(nb, seed=seed, box=40.0, height=6.0)
    rng = np.random.default_rng(seed)
    sc["linvel"][:, :3] = rng.uniform(-spread_v, spread_v, (nb, 3)).astype(np.float32)
    sc["angvel"][:, :3] = rng.uniform(-2, 2, (nb, 3)).astype(np.float32)
    ref_pos = sc["ref_pos"].copy()
    ref_pos[:, :3] *= 1.001
    sc["ref_pos"] = ref_pos
    return sc


def _prb_pair(sc, nb):
    prb = G.ParticleRigidBody(nb, nb * 27, nb * 27 * 40)
    prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"], sc["pstart"],
                    sc["npart"], sc["inv_mass"], sc["inv_inertia"])
    prb.set_particle_data(sc["ref_pos"], sc["ref_mass"], 27)
    prb.set_particle_shape_inv_inertia(sc["shape_inv_inertia"])
    ref = O.PrbState(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"],
                     sc["pstart"], sc["npart"], sc["inv_mass"], sc["ref_pos"], sc["ref_mass"], 27,
                     sc["shape_inv_inertia"])
    p = prb.ptrs()
    ref.quat[:] = d2h(p.rb_quat, (nb, 4), np.float32)  # start from the GPU's normalised quaternions
    return prb, ref


def test_reference_demo_config_with_velocity_cap(dev):
    """The demo's DemoParticles::init settings (demoParticles.h:406-407):
    setParams(2.1, 0.8) and setMaxVelocity(smallRadius = 0.25, 0.01), i.e. a
    cap of 25, with initial speeds up to 40 so the cap binds; g = -5, floor
    scale 0.1, y clamp.  Body state after every step equals the oracle's."""
    import torch
    nb = 64
    sc = _ref_mode_scene(nb, 13, 40.0)
    prb, ref = _prb_pair(sc, nb)
    prb.set_params(2.1, 0.8)
    prb.set_max_velocity(0.25, 0.01)
    flags = G.PRB_FLOOR_CLAMP
    prb.set_physics(gravity=(0, -5, 0), floor_scale=0.1, flags=flags, broadphase_kind=G.INPUT_SPHERE)
    sap = G.Sap(input_kind=G.INPUT_SPHERE, axis_mode=G.AXIS_COORD_MAXVAR, output_kind=G.OUT_SYM_CSR)
    vmax = np.float32(np.float32(0.25) / np.float32(0.01))
    params = O.prb_params(gravity=(0, -5, 0), sp=2.1, dp=0.8, floor_scale=0.1, max_lin_vel=float(vmax),
                          flags=flags | O.PRB_CAP_VEL)
    capped = 0
    for step in range(6):
        prb.step(sap, 0.025)
        ref.step(0.025, params, broadphase="sphere")
        torch.cuda.synchronize()
        p = prb.ptrs()
        pos = d2h(p.rb_pos, (nb, 4), np.float32)
        quat = d2h(p.rb_quat, (nb, 4), np.float32)
        lv = d2h(p.rb_linvel, (nb, 4), np.float32)
        av = d2h(p.rb_angvel, (nb, 4), np.float32)
        assert np.max(np.abs(pos[:, :3] - ref.pos[:, :3])) <= TOL_X * 40.0, step
        assert np.max(np.abs(quat - ref.quat)) <= TOL_Q, step
        assert np.max(np.abs(lv - ref.linvel)) <= 1e-5 * vmax, step
        assert np.max(np.abs(av - ref.angvel)) <= 1e-4, step
        capped += int(np.sum(np.abs(np.linalg.norm(ref.linvel[:, :3].astype(np.float64), axis=1) - vmax)
                             < 1e-5 * vmax))
    assert capped > 0  # the cap was exercised
    assert np.all(pos[:, 1] >= 0.25)


def test_inertia_path_and_prepare(dev):
    """PRB_INERTIA: prepare (prepareK, ParticleRigidBodyKernel.h:30-45) makes
    I_w^-1 = R I0^-1 R^T every step and the angular update uses it.  A
    non-diagonal shape inertia; I_w^-1, angular velocity, position and
    quaternion equal the oracle's after every step."""
    import torch
    nb, box = 200, 24.0
    sc = W.bodies(nb, seed=21, box=box, height=10.0)
    rng = np.random.default_rng(21)
    sc["angvel"][:, :3] = rng.uniform(-3, 3, (nb, 3)).astype(np.float32)
    a = rng.standard_normal((3, 3))
    I0 = (a @ a.T + 3 * np.eye(3)) / 30.0
    sii = np.zeros((1, 12), np.float32)
    for i in range(3):
        sii[0, 4 * i:4 * i + 3] = I0[i]
    sc["shape_inv_inertia"] = sii
    flags = G.PRB_ABS_SPRING | G.PRB_WALLS | G.PRB_SKIP_SAME_BODY | G.PRB_INERTIA
    prb, ref = _prb_pair(sc, nb)
    prb.set_physics(gravity=(0, -9.81, 0), k_abs=1e4, c_abs=50.0, floor_scale=1.0, wall_lo=(0, 0),
                    wall_hi=(box, box), flags=flags, broadphase_kind=G.INPUT_AABB)
    sap = G.Sap(input_kind=G.INPUT_AABB, axis_mode=G.AXIS_PCA, output_kind=G.OUT_SYM_CSR,
                flags=G.FLAG_SKIP_SAME_GROUP)
    params = O.prb_params(gravity=(0, -9.81, 0), k_abs=1e4, c_abs=50.0, floor_scale=1.0,
                          wall_lo=(0, 0), wall_hi=(box, box), flags=flags)
    # prepare alone first: against the oracle and the closed form
    prb.prepare()
    ref.prepare()
    torch.cuda.synchronize()
    iw = d2h(prb.ptrs().rb_inv_inertia, (nb, 12), np.float32)
    assert np.array_equal(iw, ref.inv_inertia)
    from scipy.spatial.transform import Rotation
    for b in range(0, nb, 17):
        R = Rotation.from_quat(ref.quat[b].astype(np.float64)).as_matrix()
        want = R @ I0.astype(np.float32).astype(np.float64) @ R.T
        assert np.max(np.abs(iw[b].reshape(3, 4)[:, :3] - want)) < 1e-5 * np.max(np.abs(want))
    for step in range(10):
        prb.step(sap, 1e-4)
        ref.step(1e-4, params, broadphase="aabb")
        torch.cuda.synchronize()
        p = prb.ptrs()
        iw = d2h(p.rb_inv_inertia, (nb, 12), np.float32)
        av = d2h(p.rb_angvel, (nb, 4), np.float32)
        pos = d2h(p.rb_pos, (nb, 4), np.float32)
        quat = d2h(p.rb_quat, (nb, 4), np.float32)
        assert np.max(np.abs(iw - ref.inv_inertia)) <= 1e-6 * np.max(np.abs(I0)), step
        assert np.max(np.abs(av - ref.angvel)) <= 1e-5, step
        assert np.max(np.abs(pos[:, :3] - ref.pos[:, :3])) <= TOL_X * box, step
        assert np.max(np.abs(quat - ref.quat)) <= TOL_Q, step


def test_body_getters(dev):
    """getRbShapeIdx / getRbNumParticles (ParticleRigidBody.h:109-110):
    device arrays of each body's shape and particle count."""
    nb = 10
    sc = W.bodies(nb, seed=2, box=30.0, height=5.0)
    sc["shape"][:] = np.arange(nb) % 2
    sc["shape_inv_inertia"] = np.repeat(sc["shape_inv_inertia"], 2, 0)
    ref_pos = np.concatenate([sc["ref_pos"], sc["ref_pos"]])
    prb = G.ParticleRigidBody(nb, nb * 27, 1000)
    prb.set_rb_data(sc["pos"], sc["quat"], sc["linvel"], sc["angvel"], sc["shape"], sc["pstart"],
                    sc["npart"], sc["inv_mass"], sc["inv_inertia"])
    prb.set_particle_data(ref_pos, np.concatenate([sc["ref_mass"]] * 2), 27)
    p = prb.ptrs()
    assert np.array_equal(d2h(p.rb_shape_idx, nb, np.int32), sc["shape"])
    assert np.array_equal(d2h(p.rb_num_particles, nb, np.int32), sc["npart"])
