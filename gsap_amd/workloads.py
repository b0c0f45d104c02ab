"""Seeded synthetic workloads (numpy PCG64; identical on every machine with the
same numpy).  Shapes follow the five benchmark configurations of this project:

  small     16,384 AABBs, centres U[0,64)^3, half-extents U[0.1,0.5] per axis
  aniso     1,048,576 AABBs, centres R diag(20,5,1) z (z ~ N(0,I), R a seeded
            uniform rotation), half-extents log-uniform [0.05,0.8] per axis
  bodies    rigid cubes of 3x3x3 particles (radius 0.5, spacing 1.0) dropped in
            a 256x256 walled box
  scenes    independent scenes of log-uniform size, uniform or 8-cluster
            Gaussian centres, half-extents U[0.1,0.6]
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def random_rotation(rng):
    """Uniform random rotation (Shoemake 1992 unit quaternion)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def boxes_from(c, h):
    c = c.astype(np.float32)
    h = h.astype(np.float32)
    return np.ascontiguousarray(np.concatenate([c - h, c + h], axis=1), dtype=np.float32)


def small(n=16384, seed=0):
    r = _rng(seed)
    c = r.uniform(0.0, 64.0, (n, 3))
    h = r.uniform(0.1, 0.5, (n, 3))
    return boxes_from(c, h)


def aniso(n=1 << 20, seed=1, sigma=(20.0, 5.0, 1.0)):
    r = _rng(seed)
    R = random_rotation(r)
    z = r.standard_normal((n, 3)) * np.asarray(sigma)
    c = z @ R.T
    h = np.exp(r.uniform(np.log(0.05), np.log(0.8), (n, 3)))
    return boxes_from(c, h)


def scene(n, r, kind=None):
    """One scene: uniform in [0,L)^3 with L = 64 (n/16384)^(1/3) (the density of
    `small`) or 8 Gaussian clusters with sigma = L/16; half-extents U[0.1,0.6]."""
    L = 64.0 * (n / 16384.0) ** (1.0 / 3.0)
    if kind is None:
        kind = "uniform" if r.random() < 0.5 else "cluster"
    if kind == "uniform":
        c = r.uniform(0.0, L, (n, 3))
    else:
        centres = r.uniform(0.0, L, (8, 3))
        lab = r.integers(0, 8, n)
        c = centres[lab] + r.standard_normal((n, 3)) * (L / 16.0)
    h = r.uniform(0.1, 0.6, (n, 3))
    return boxes_from(c, h)


def scene_sizes(n_scenes, seed, lo=1024, hi=65536):
    r = _rng(seed)
    return np.floor(np.exp(r.uniform(np.log(lo), np.log(hi), n_scenes))).astype(np.int64)


def scenes(n_scenes=512, seed=4, lo=1024, hi=65536):
    """Packed scenes: (boxes [sum n, 6], offsets [n_scenes+1] uint32)."""
    sizes = scene_sizes(n_scenes, seed, lo, hi)
    r = _rng(seed + 1000)
    parts = [scene(int(s), r) for s in sizes]
    off = np.zeros(n_scenes + 1, np.uint32)
    off[1:] = np.cumsum(sizes)
    return np.concatenate(parts, axis=0), off


def cube_template(spacing=1.0, radius=0.5, per_side=3):
    g = (np.arange(per_side) - (per_side - 1) / 2.0) * spacing
    xs, ys, zs = np.meshgrid(g, g, g, indexing="ij")
    ref = np.stack([xs.ravel(), ys.ravel(), zs.ravel(), np.full(xs.size, radius)], 1)
    return ref.astype(np.float32)


def bodies(nb=65536, seed=3, box=256.0, height=200.0, per_side=3, mass=1.0):
    """Rigid cubes dropped into a walled box.  Returns a dict of host arrays in
    the ParticleRigidBody layout (float4 body SoA, particle templates)."""
    r = _rng(seed)
    ref = cube_template(per_side=per_side)
    k = len(ref)
    half = (per_side - 1) / 2.0 + 0.5
    pos = np.zeros((nb, 4), np.float32)
    pos[:, 0] = r.uniform(half, box - half, nb)
    pos[:, 1] = r.uniform(half, height, nb)
    pos[:, 2] = r.uniform(half, box - half, nb)
    u1, u2, u3 = r.random((3, nb))
    quat = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                     np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)],
                    1).astype(np.float32)
    zeros = np.zeros((nb, 4), np.float32)
    # Solid cube of side per_side * spacing with total mass k*mass.
    side = per_side * 1.0
    I0 = k * mass * side * side / 6.0
    inv_i = np.zeros((1, 12), np.float32)
    inv_i[0, 0] = inv_i[0, 5] = inv_i[0, 10] = 1.0 / I0
    return dict(pos=pos, quat=quat, linvel=zeros.copy(), angvel=zeros.copy(),
                shape=np.zeros(nb, np.int32), pstart=(np.arange(nb) * k).astype(np.int32),
                npart=np.full(nb, k, np.int32), inv_mass=np.full(nb, 1.0 / (k * mass), np.float32),
                inv_inertia=np.repeat(inv_i, nb, 0), ref_pos=ref, ref_mass=np.full(k, mass, np.float32),
                max_per_shape=k, shape_inv_inertia=inv_i)
