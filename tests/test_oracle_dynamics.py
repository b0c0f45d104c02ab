"""Pins the oracle's particle rigid-body functions by closed forms that do not
reuse its own arithmetic: rotations against scipy's quaternion algebra
(float64), the contact laws against the analytic damped-oscillator solution
and the exact discrete recurrences of the reference's spring law
(ParticleRigidBodyKernel.h:13-22, :106-208 with the step order of
demoParticles.h:856-928), prepare against R I0^-1 R^T, and the velocity cap
(:247-256).  A plausible mistake in the oracle (rotation direction, operand
order of the quaternion product, effective mass, a missing 1/dist, the sign
of the damping) fails at least one of them."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O


def _bodies(pos, quat, linvel, angvel, ref, masses, shape=None, inv_mass=None):
    """Single-particle bodies (one template per mass)."""
    nb = len(pos)
    shape = list(range(nb)) if shape is None else shape
    ref_pos = np.array([[*ref[k], r] for k, r in enumerate(masses["radius"])], np.float32)
    ref_mass = np.array(masses["mass"], np.float32)
    inv = [1.0 / ref_mass[shape[b]] for b in range(nb)] if inv_mass is None else inv_mass
    return O.PrbState(pos=[[*p, 0] for p in pos], quat=quat, linvel=[[*v, 0] for v in linvel],
                      angvel=[[*w, 0] for w in angvel], shape=shape, pstart=list(range(nb)),
                      npart=[1] * nb, inv_mass=inv, ref_pos=ref_pos, ref_mass=ref_mass,
                      max_per_shape=1)


def _unit_quat(rng):
    q = rng.standard_normal(4)
    return (q / np.linalg.norm(q)).astype(np.float32)


# ------------------------------------------------------------------ rotation
@pytest.mark.parametrize("axis,v,expect", [
    ((0, 0, 1), (1, 0, 0), (0, 1, 0)),
    ((1, 0, 0), (0, 1, 0), (0, 0, 1)),
    ((0, 1, 0), (0, 0, 1), (1, 0, 0)),
])
def test_quarter_turn_is_right_handed(axis, v, expect):
    """quaternion.h:141-148 rotateVector: q v q^-1 with q = (n sin 45deg,
    cos 45deg) turns v by +90 degrees about n."""
    q = [*(np.sin(np.pi / 4) * np.array(axis)), np.cos(np.pi / 4)]
    s = _bodies([(0, 0, 0)], [q], [(0, 0, 0)], [(0, 0, 0)], [v], {"radius": [0.5], "mass": [1.0]})
    s.update_particles()
    assert np.allclose(s.ppos[0, :3], expect, atol=1e-6)
    assert s.ppos[0, 3] == np.float32(0.5)  # w carries the template radius


def test_rotation_and_particle_velocity_match_scipy():
    """Random unit q and templates: particle position = x + R(q) r and
    velocity = v + w x R(q) r (ParticleRigidBodyKernel.h:71-104), against
    scipy's float64 rotation and numpy's cross product."""
    rng = np.random.default_rng(1)
    nb = 64
    q = np.stack([_unit_quat(rng) for _ in range(nb)])
    ref = rng.uniform(-2, 2, (nb, 3))
    x = rng.uniform(-50, 50, (nb, 3))
    v = rng.uniform(-3, 3, (nb, 3))
    w = rng.uniform(-3, 3, (nb, 3))
    s = _bodies(x, q, v, w, ref, {"radius": [0.5] * nb, "mass": [1.0] * nb})
    s.update_particles()
    r64 = Rotation.from_quat(q.astype(np.float64)).apply(ref.astype(np.float32).astype(np.float64))
    x32, v32, w32 = (a.astype(np.float32).astype(np.float64) for a in (x, v, w))
    assert np.max(np.abs(s.ppos[:, :3] - (x32 + r64))) < 1e-4  # |x| <= 50: fp32 ulp 4e-6
    assert np.max(np.abs(s.pvel[:, :3] - (v32 + np.cross(w32, r64)))) < 1e-5


def test_quaternion_update_is_world_frame():
    """integrateK (ParticleRigidBodyKernel.h:55-58): q += 1/2 (w,0) q dt then
    normalise is the first-order world-frame update R1 = exp([w] dt) R0.  The
    body-frame order (q (w,0)) differs at first order in dt."""
    rng = np.random.default_rng(2)
    dt = 1e-3
    for _ in range(20):
        q0 = _unit_quat(rng)
        w = rng.uniform(-3, 3, 3)
        s = _bodies([(0, 5, 0)], [q0], [(0, 0, 0)], [w], [(0, 0, 0)], {"radius": [0.5], "mass": [1.0]})
        s.integrate(dt, O.prb_params(gravity=(0, 0, 0)))
        w32 = w.astype(np.float32).astype(np.float64)
        world = (Rotation.from_rotvec(w32 * dt) * Rotation.from_quat(q0.astype(np.float64))).as_quat()
        body = (Rotation.from_quat(q0.astype(np.float64)) * Rotation.from_rotvec(w32 * dt)).as_quat()
        got = s.quat[0].astype(np.float64)
        d_world = min(np.max(np.abs(got - world)), np.max(np.abs(got + world)))
        d_body = min(np.max(np.abs(got - body)), np.max(np.abs(got + body)))
        assert d_world < 1e-5, d_world
        if np.linalg.norm(np.cross(w32, Rotation.from_quat(q0).as_rotvec())) > 0.5:
            assert d_body > 1e-4  # the check can tell the two orders apart


# ------------------------------------------------------------------ contacts
def _run_pair(s, dt, params, steps):
    """Step two single-particle bodies with a fixed symmetric pair list (the
    force law itself skips pairs that do not penetrate).  Returns the
    penetration of each step (after integrate, as the force kernel sees it)
    and the relative velocity v1 - v0 along x after each step."""
    sym = np.array([[0, 1], [1, 0]], np.int32)
    pen, rel = [], []
    for _ in range(steps):
        s.integrate(dt, params)
        s.update_particles()
        d = np.linalg.norm(s.ppos[1, :3].astype(np.float64) - s.ppos[0, :3].astype(np.float64))
        pen.append(float(s.ppos[0, 3]) + float(s.ppos[1, 3]) - d)
        s.forces(sym, dt, params)
        s.update_velocity(dt, params)
        rel.append(float(s.linvel[1, 0]) - float(s.linvel[0, 0]))
    return np.array(pen), np.array(rel)


def test_absolute_contact_law_restitution_and_contact_time():
    """Head-on hit of two unit-mass particles (radius 1.5, so |r_ij| = 3 and a
    missing 1/dist would triple the stiffness), k = 1e4, c = 50, g = 0,
    dt = 1e-5.  The linear spring-dashpot F = -k pen n + c (v_j - v_i) gives
    pen'' + (2c/m) pen' + (2k/m) pen = 0: restitution e = exp(-z pi/sqrt(1-z^2))
    and contact time pi/(w0 sqrt(1-z^2)), w0 = sqrt(2k/m), z = c/sqrt(2km)."""
    k, c, m, r, dt = 1e4, 50.0, 1.0, 1.5, 1e-5
    s = _bodies([(0, 10, 0), (2 * r, 10, 0)], [[0, 0, 0, 1]] * 2, [(1, 0, 0), (-1, 0, 0)],
                [(0, 0, 0)] * 2, [(0, 0, 0)] * 2, {"radius": [r, r], "mass": [m, m]})
    p = O.prb_params(gravity=(0, 0, 0), k_abs=k, c_abs=c, flags=O.PRB_ABS_SPRING)
    w0 = np.sqrt(2 * k / m)
    z = c / np.sqrt(2 * k * m)
    e_th = np.exp(-z * np.pi / np.sqrt(1 - z * z))
    t_th = np.pi / (w0 * np.sqrt(1 - z * z))
    pen, rel = _run_pair(s, dt, p, int(1.5 * t_th / dt))
    inside = pen > 0
    assert inside[0] and not inside[-1]
    n_in = int(np.argmin(inside))  # first step out of contact
    assert inside[:n_in].all() and not inside[n_in:].any()
    e = rel[-1] / 2.0  # approach speed 2
    assert abs(e / e_th - 1) < 0.01, (e, e_th)
    assert abs(n_in * dt / t_th - 1) < 0.01, (n_in * dt, t_th)
    # momentum: equal masses, equal and opposite forces
    assert abs(float(s.linvel[0, 0] + s.linvel[1, 0])) < 1e-5


@pytest.mark.parametrize("sp", [0.1, 0.5, 1.0])
def test_reference_spring_law_recurrence(sp):
    """calcSpringCoeff (ParticleRigidBodyKernel.h:13-22) with dp = 1 (no
    damping): k = m_eff sp / dt^2 with m_eff = 1/(1/m_i + 1/m_j), so the
    relative acceleration is sp pen / dt^2 for ANY masses.  With the step
    order integrate -> forces -> update velocity the penetration obeys exactly
    pen[n+2] = (2 - sp) pen[n+1] - pen[n] whenever pen[n+1] > 0, and
    pen[n+2] = 2 pen[n+1] - pen[n] (free flight) otherwise.  Masses 1 and 3
    make a wrong effective mass (e.g. m_i + m_j) fail."""
    dt, r = 1e-3, 0.5
    s = _bodies([(0, 10, 0), (2 * r, 10, 0)], [[0, 0, 0, 1]] * 2, [(1, 0, 0), (-1, 0, 0)],
                [(0, 0, 0)] * 2, [(0, 0, 0)] * 2, {"radius": [r, r], "mass": [1.0, 3.0]})
    p = O.prb_params(gravity=(0, 0, 0), sp=sp, dp=1.0)
    pen, _ = _run_pair(s, dt, p, 40)
    n_contact = 0
    for n in range(len(pen) - 2):
        if pen[n + 1] > 0:
            pred = (2 - sp) * pen[n + 1] - pen[n]
            n_contact += 1
        else:
            pred = 2 * pen[n + 1] - pen[n]
        assert abs(pen[n + 2] - pred) < 2e-6, (n, pen[n:n + 3], pred)
    assert n_contact >= 2
    assert pen[-1] < 0  # they separated


def test_reference_damping_decays_relative_velocity_by_dp():
    """sp = 0: only damping, c = m_eff (1 - dp)/dt, so each step in contact
    multiplies the relative velocity by exactly dp (masses 1 and 3).  A
    reversed damping sign would make it grow."""
    dt, r, dp = 1e-3, 0.5, 0.7
    s = _bodies([(0, 10, 0), (2 * r, 10, 0)], [[0, 0, 0, 1]] * 2, [(1, 0, 0), (-1, 0, 0)],
                [(0, 0, 0)] * 2, [(0, 0, 0)] * 2, {"radius": [r, r], "mass": [1.0, 3.0]})
    p = O.prb_params(gravity=(0, 0, 0), sp=0.0, dp=dp)
    pen, rel = _run_pair(s, dt, p, 12)
    assert (pen > 0).all()
    prev = -2.0
    for k in range(len(rel)):
        assert abs(rel[k] / prev - dp) < 1e-5, (k, rel[k], prev)
        prev = rel[k]


def test_reference_floor_law_recurrence():
    """clearForceAndBoundaryConditionK (:106-140): the floor spring uses
    calcSpringCoeff(1/m, 0.5/m) -> m_eff = m/1.5, scaled by 0.1, damping
    m_eff (1 - dp)/dt.  With dp = 1, g = 0 a particle sinking into y = 0
    obeys pen[n+2] = (2 - 0.1 sp / 1.5) pen[n+1] - pen[n] in contact."""
    dt, r, sp = 1e-3, 0.5, 0.9
    s = _bodies([(0, r, 0)], [[0, 0, 0, 1]], [(0, -1, 0)], [(0, 0, 0)], [(0, 0, 0)],
                {"radius": [r], "mass": [2.0]})
    p = O.prb_params(gravity=(0, 0, 0), sp=sp, dp=1.0, floor_scale=0.1)
    pen = []
    for _ in range(60):
        s.integrate(dt, p)
        s.update_particles()
        pen.append(r - float(s.ppos[0, 1]))
        s.forces(np.zeros((0, 2), np.int32), dt, p)
        s.update_velocity(dt, p)
    a = 0.1 * sp / 1.5
    n_contact = 0
    for n in range(len(pen) - 2):
        if pen[n + 1] > 0:
            assert abs(pen[n + 2] - ((2 - a) * pen[n + 1] - pen[n])) < 2e-6
            n_contact += 1
    assert n_contact >= 10


# ------------------------------------------------------------------ prepare, cap, torque
def test_prepare_is_rotated_inverse_inertia():
    """prepareK (:30-45): I_w^-1 = R I0^-1 R^T, R from scipy's rotation of the
    body quaternion; symmetric with the eigenvalues of I0^-1."""
    rng = np.random.default_rng(4)
    nb = 32
    q = np.stack([_unit_quat(rng) for _ in range(nb)])
    shapes = []
    for _ in range(3):
        a = rng.standard_normal((3, 3))
        m = a @ a.T + np.eye(3)
        shapes.append(m)
    sii = np.zeros((3, 12), np.float32)
    for k, m in enumerate(shapes):
        for i in range(3):
            sii[k, 4 * i:4 * i + 3] = m[i]
    shape = rng.integers(0, 3, nb)
    s = O.PrbState(pos=np.zeros((nb, 4)), quat=q, linvel=np.zeros((nb, 4)), angvel=np.zeros((nb, 4)),
                   shape=shape, pstart=np.arange(nb), npart=np.ones(nb), inv_mass=np.ones(nb),
                   ref_pos=np.array([[0, 0, 0, 0.5]] * 3, np.float32), ref_mass=np.ones(3, np.float32),
                   max_per_shape=1, shape_inv_inertia=sii)
    s.prepare()
    for b in range(nb):
        R = Rotation.from_quat(q[b].astype(np.float64)).as_matrix()
        I0 = sii[shape[b]].reshape(3, 4)[:, :3].astype(np.float64)
        want = R @ I0 @ R.T
        got = s.inv_inertia[b].reshape(3, 4)[:, :3].astype(np.float64)
        assert np.max(np.abs(got - want)) < 1e-5 * np.max(np.abs(want)), b
        assert np.allclose(np.linalg.eigvalsh((got + got.T) / 2), np.linalg.eigvalsh(I0), rtol=1e-5)
        assert np.all(s.inv_inertia[b].reshape(3, 4)[:, 3] == 0)


def test_velocity_cap():
    """updateVelockyK<CAP_VEL> (:247-256): |v| > maxLinVel = minRadius/dt
    (setMaxVelocity, ParticleRigidBody.h:88) is rescaled to exactly maxLinVel
    in the same direction with w = 0; slower bodies are untouched."""
    rng = np.random.default_rng(6)
    nb = 40
    dt, vmax = 0.01, 0.25 / 0.01
    s = _bodies([(0, 10, 0)] * nb, [[0, 0, 0, 1]] * nb, [(0, 0, 0)] * nb, [(0, 0, 0)] * nb,
                [(0, 0, 0)] * nb, {"radius": [0.5] * nb, "mass": [1.0] * nb})
    s.update_particles()
    f = rng.standard_normal((nb, 3)) * rng.uniform(0.5, 6000, (nb, 1))
    s.pforce[:, :3] = f
    p = O.prb_params(gravity=(0, 0, 0), max_lin_vel=vmax, flags=O.PRB_CAP_VEL)
    s.update_velocity(dt, p)
    v_uncapped = f.astype(np.float32).astype(np.float64) * dt
    speed = np.linalg.norm(v_uncapped, axis=1)
    capped = speed > vmax * (1 + 1e-6)
    assert capped.any() and (~capped).any()
    got = s.linvel[:, :3].astype(np.float64)
    assert np.allclose(np.linalg.norm(got[capped], axis=1), vmax, rtol=1e-6)
    u = v_uncapped[capped] / speed[capped, None]
    assert np.allclose(got[capped] / vmax, u, atol=1e-6)
    free = ~capped & (speed < vmax * (1 - 1e-6))
    assert np.allclose(got[free], v_uncapped[free], rtol=1e-6, atol=1e-9)
    assert np.all(s.linvel[capped, 3] == 0)


@pytest.mark.parametrize("inertia", [False, True])
def test_torque_from_known_lever(inertia):
    """updateVelocityK: tau = sum (R r) x F.  A particle at r = (1,0,0) under
    F = (0,1,0) gives tau = (0,0,1); without PRB_INERTIA w += tau dt (the
    reference's :260-261), with it w += I_w^-1 tau dt.  The general case uses
    a lever and force off the axes and a non-symmetric I0^-1 (identity
    rotation, so I_w^-1 = I0^-1), so the matrix-vector product's row/column
    order matters."""
    dt = 0.5
    flags = O.PRB_INERTIA if inertia else 0
    I0 = np.array([[2.0, 0.3, -0.4], [0.1, 3.0, 0.7], [-0.6, 0.2, 4.0]])
    for r, f in (((1, 0, 0), (0, 1, 0)), ((1.0, 2.0, 0.5), (0.3, 1.0, -0.2))):
        s = _bodies([(0, 10, 0)], [[0, 0, 0, 1]], [(0, 0, 0)], [(0, 0, 0)], [r],
                    {"radius": [0.5], "mass": [1.0]})
        s.shape_inv_inertia = np.zeros((1, 12), np.float32)
        for i in range(3):
            s.shape_inv_inertia[0, 4 * i:4 * i + 3] = I0[i]
        s.update_particles()
        s.pforce[0, :3] = f
        if inertia:
            s.prepare()
        s.update_velocity(dt, O.prb_params(gravity=(0, 0, 0), flags=flags))
        tau = np.cross(np.float32(r), np.float32(f)).astype(np.float64)
        want = (I0.astype(np.float32) @ tau if inertia else tau) * dt
        assert np.allclose(s.angvel[0, :3], want, atol=1e-6), (r, f)
        assert np.allclose(s.linvel[0, :3], np.float32(f) * dt, atol=1e-7)
