// k_prb.hip -- particle rigid-body step on gfx950.
//
// Kernels follow physicboxSap2.16_Fermi/ParticleRigidBody/
// ParticleRigidBodyKernel.h (integrateK :47-69, updateParticleDataK :71-104,
// prepareK :30-45, clearForceAndBoundaryConditionK :106-140,
// comupteFoceOnParticlesK :142-208, updateVelocityK :213-262) with the
// float4 / quaternion arithmetic of PB/include/quaternion.h:90-148 and
// PB/src/cutil_math.h:400-557, evaluated in the same order as the CPU oracle
// with contraction off, so the trajectories agree bit for bit when the
// broadphase agrees.  Differences from the reference, all runtime options:
// gravity, absolute spring constants, walls, the y clamp, same-body pair
// skipping, I^-1 in the angular update; one thread per particle (not per
// body) in updateParticleData; the pair loop bounded by the pair count.
#include "gsap_device.h"

#pragma clang fp contract(off)

namespace gsap {

struct F4 {
    float x, y, z, w;
};
__device__ __forceinline__ F4 f4(float4 v) { return F4{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ float4 t4(F4 v) { return make_float4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ F4 add4(F4 a, F4 b) { return F4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ F4 sub4(F4 a, F4 b) { return F4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
__device__ __forceinline__ F4 mul4s(F4 a, float s) { return F4{a.x * s, a.y * s, a.z * s, a.w * s}; }
__device__ __forceinline__ F4 smul4(float s, F4 a) { return F4{s * a.x, s * a.y, s * a.z, s * a.w}; }
__device__ __forceinline__ float dot4(F4 a, F4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
__device__ __forceinline__ float dot3(F4 a, F4 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ F4 cross4(F4 a, F4 b) {
    return F4{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f};
}
// quaternion.h:130-139
__device__ __forceinline__ F4 qmul(F4 a, F4 b) {
    F4 ans = cross4(a, b);
    ans = add4(ans, add4(smul4(a.w, b), smul4(b.w, a)));
    ans.w = a.w * b.w - dot3(a, b);
    return ans;
}
// quaternion.h:141-148
__device__ __forceinline__ F4 qrot(F4 q, F4 in) {
    F4 qi{-q.x, -q.y, -q.z, q.w};
    return qmul(qmul(q, in), qi);
}
// quaternion.h:116-128
__device__ __forceinline__ F4 qnormalize(F4 q) {
    float l = sqrtf(dot4(q, q));
    float inv = 1.0f / l;
    return mul4s(q, inv);
}

__global__ void k_integrate(PrbDev d, PrbPhys p, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nb) return;
    F4 g{p.gravity[0], p.gravity[1], p.gravity[2], 0.0f};
    F4 v = add4(f4(d.linvel[i]), mul4s(g, dt));
    d.linvel[i] = t4(v);
    F4 q = f4(d.quat[i]);
    F4 w = f4(d.angvel[i]);
    w.w = 0.0f;
    q = add4(q, mul4s(qmul(w, q), dt * 0.5f));
    q = qnormalize(q);
    d.quat[i] = t4(q);
    float4 x = d.pos[i];
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    if (p.flags & 4u) {  // GSAP_PRB_FLOOR_CLAMP (ParticleRigidBodyKernel.h:64-68)
        if ((i / 6) % 2 == 0 && x.y <= 0.25f) x.y = 0.25f;
        else if ((i / 6) % 2 == 1 && x.y <= 0.3f) x.y = 0.3f;
    }
    d.pos[i] = x;
}

// aabb (optional): also the particle's box, as k_particle_aabbs computes it
// from the position written here (the particle step's broadphase input)
__global__ void k_update_particles(PrbDev d, float* aabb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int b = d.pbody[p], ip = d.pip[p];
    const int r = d.max_per_shape * d.shape[b] + ip;
    const F4 x = f4(d.pos[b]), v = f4(d.linvel[b]), q = f4(d.quat[b]), w = f4(d.angvel[b]);
    const F4 ref = f4(d.ref_pos[r]);
    F4 cur = qrot(q, ref);
    const int o = d.pstart[b] + ip;
    d.pvel[o] = t4(add4(v, cross4(w, cur)));
    cur = add4(cur, x);
    const float4 c = make_float4(cur.x, cur.y, cur.z, ref.w);
    d.ppos[o] = c;
    d.pmass[o] = d.ref_mass[r];
    if (aabb) {
        float* a = aabb + 6 * (size_t)o;
        a[0] = c.x - c.w; a[1] = c.y - c.w; a[2] = c.z - c.w;
        a[3] = c.x + c.w; a[4] = c.y + c.w; a[5] = c.z + c.w;
    }
}

__global__ void k_prepare(PrbDev d) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.nb) return;
    const F4 q = f4(d.quat[b]);
    float R[3][3], I[3][3], T[3][3];
    const float x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z;
    R[0][0] = 1 - 2 * y2 - 2 * z2;
    R[0][1] = 2 * q.x * q.y - 2 * q.w * q.z;
    R[0][2] = 2 * q.x * q.z + 2 * q.w * q.y;
    R[1][0] = 2 * q.x * q.y + 2 * q.w * q.z;
    R[1][1] = 1 - 2 * x2 - 2 * z2;
    R[1][2] = 2 * q.y * q.z - 2 * q.w * q.x;
    R[2][0] = 2 * q.x * q.z - 2 * q.w * q.y;
    R[2][1] = 2 * q.y * q.z + 2 * q.w * q.x;
    R[2][2] = 1 - 2 * x2 - 2 * y2;
    const float4* s = d.shape_inv_inertia + 3 * d.shape[b];
    for (int i = 0; i < 3; ++i) {
        const float4 row = s[i];
        I[i][0] = row.x; I[i][1] = row.y; I[i][2] = row.z;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i][j] = (R[i][0] * I[0][j] + R[i][1] * I[1][j]) + R[i][2] * I[2][j];
    for (int i = 0; i < 3; ++i) {
        float o[3];
        for (int j = 0; j < 3; ++j) o[j] = (T[i][0] * R[j][0] + T[i][1] * R[j][1]) + T[i][2] * R[j][2];
        d.inv_inertia[3 * b + i] = make_float4(o[0], o[1], o[2], 0.0f);
    }
}

__device__ __forceinline__ void spring_coeff(float im, float jm, float dt2inv, float dtinv,
                                             const PrbPhys& p, float* s, float* dd) {
    if (p.flags & 1u) { *s = p.k_abs; *dd = p.c_abs; return; }
    float m = im + jm;
    m = 1.0f / m;
    *s = m * dt2inv * p.sp;
    *dd = m * dtinv * (1.0f - p.dp);
}

__device__ __forceinline__ bool plane_force(F4* out, float dist, F4 nrm, F4 pos, F4 vel, float mass,
                                            float dt2inv, float dtinv, const PrbPhys& p) {
    if (!(dist < pos.w)) return false;
    float s, dd;
    spring_coeff(1.0f / mass, 0.5f / mass, dt2inv, dtinv, p, &s, &dd);
    F4 nv{-vel.x, -vel.y, -vel.z, -vel.w};
    F4 sf = mul4s(smul4(-s * (dist - pos.w), nrm), p.floor_scale);
    F4 df = smul4(dd, nv);
    *out = add4(sf, df);
    return true;
}

__global__ void k_boundary(PrbDev d, PrbPhys p, float dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.np) return;
    const float dt2inv = 1.0f / (dt * dt), dtinv = 1.0f / dt;
    const F4 pos = f4(d.ppos[i]), vel = f4(d.pvel[i]);
    const float mass = d.pmass[i];
    F4 force{0, 0, 0, 0}, c;
    if (plane_force(&c, pos.y, F4{0, 1, 0, 0}, pos, vel, mass, dt2inv, dtinv, p)) force = c;
    if (p.flags & 2u) {
        if (plane_force(&c, pos.x - p.wall_lo[0], F4{1, 0, 0, 0}, pos, vel, mass, dt2inv, dtinv, p))
            force = add4(force, c);
        if (plane_force(&c, p.wall_hi[0] - pos.x, F4{-1, 0, 0, 0}, pos, vel, mass, dt2inv, dtinv, p))
            force = add4(force, c);
        if (plane_force(&c, pos.z - p.wall_lo[1], F4{0, 0, 1, 0}, pos, vel, mass, dt2inv, dtinv, p))
            force = add4(force, c);
        if (plane_force(&c, p.wall_hi[1] - pos.z, F4{0, 0, -1, 0}, pos, vel, mass, dt2inv, dtinv, p))
            force = add4(force, c);
    }
    d.pforce[i] = t4(force);
}

// One thread per row of the symmetric sorted pair list (a distinct i), partners
// in list order.  counts[0] = directed pairs, counts[1] = rows.
__global__ void k_pair_forces(PrbDev d, PrbPhys p, const int2* __restrict__ sym,
                              const int32_t* __restrict__ entries,
                              const unsigned long long* __restrict__ counts,
                              unsigned long long max_pairs, float dt) {
    // An overflowed broadphase (count > capacity) leaves no usable list.
    if (counts[0] > max_pairs) return;
    const unsigned long long nrow = counts[1], npair = counts[0];
    const unsigned long long r = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const float dt2inv = 1.0f / (dt * dt), dtinv = 1.0f / dt;
    unsigned long long k = (unsigned long long)entries[r];
    const unsigned long long kend = (r + 1 < nrow) ? (unsigned long long)entries[r + 1] : npair;
    const int i = sym[k].x;
    const F4 ip = f4(d.ppos[i]), iv = f4(d.pvel[i]);
    const float im = d.pmass[i];
    const int bi = d.pbody[i];
    F4 sF{0, 0, 0, 0}, dF{0, 0, 0, 0};
    for (; k < kend; ++k) {
        const int j = sym[k].y;
        if ((p.flags & 16u) && d.pbody[j] == bi) continue;
        const F4 jp = f4(d.ppos[j]), jv = f4(d.pvel[j]);
        const float jm = d.pmass[j];
        F4 rij = sub4(jp, ip);
        rij.w = 0.0f;
        const F4 vij = sub4(jv, iv);
        const float dist = sqrtf(dot4(rij, rij));
        const float orig = ip.w + jp.w;
        float s, dd;
        spring_coeff(1.0f / im, 1.0f / jm, dt2inv, dtinv, p, &s, &dd);
        const float pen = orig - dist;
        if (pen > 0.0f) {
            const float inv = 1.0f / dist;
            sF = sub4(sF, mul4s(smul4(s * pen, rij), inv));
            dF = add4(dF, smul4(dd, vij));
        }
    }
    d.pforce[i] = t4(add4(f4(d.pforce[i]), add4(sF, dF)));
}

__global__ void k_update_velocity(PrbDev d, PrbPhys p, float dt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.nb) return;
    const F4 q = f4(d.quat[b]);
    F4 lf{0, 0, 0, 0}, af{0, 0, 0, 0};
    const int np = d.npart[b], ps = d.pstart[b], sh = d.shape[b];
    for (int ip = 0; ip < np; ++ip) {
        const F4 cur = qrot(q, f4(d.ref_pos[d.max_per_shape * sh + ip]));
        const F4 f = f4(d.pforce[ps + ip]);
        lf = add4(lf, f);
        af = add4(af, cross4(cur, f));
    }
    F4 v = add4(f4(d.linvel[b]), mul4s(mul4s(lf, d.inv_mass[b]), dt));
    if (p.flags & 8u) {
        F4 lv = v;
        lv.w = 0.0f;
        const float len = sqrtf(dot4(lv, lv));
        if (len > p.max_lin_vel) {
            const float inv = 1.0f / sqrtf(dot4(lv, lv));
            lv = mul4s(lv, inv);
            v = smul4(p.max_lin_vel, lv);
        }
    }
    d.linvel[b] = t4(v);
    F4 w = f4(d.angvel[b]);
    if (p.flags & 32u) {
        const float4 r0 = d.inv_inertia[3 * b], r1 = d.inv_inertia[3 * b + 1],
                     r2 = d.inv_inertia[3 * b + 2];
        F4 dw{(r0.x * af.x + r0.y * af.y) + r0.z * af.z, (r1.x * af.x + r1.y * af.y) + r1.z * af.z,
              (r2.x * af.x + r2.y * af.y) + r2.z * af.z, 0.0f};
        w = add4(w, mul4s(dw, dt));
    } else {
        w = add4(w, mul4s(af, dt));
    }
    d.angvel[b] = t4(w);
}

__global__ void k_particle_aabbs(PrbDev d, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.np) return;
    const float4 c = d.ppos[i];
    float* o = out + 6 * (size_t)i;
    o[0] = c.x - c.w; o[1] = c.y - c.w; o[2] = c.z - c.w;
    o[3] = c.x + c.w; o[4] = c.y + c.w; o[5] = c.z + c.w;
}

static inline dim3 grid_for(long long n, int bs) {
    long long g = (n + bs - 1) / bs;
    return dim3((unsigned)(g > 0 ? g : 1));
}

hipError_t prb_integrate(const PrbDev& d, const PrbPhys& p, float dt, hipStream_t st) {
    if (d.nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate, grid_for(d.nb, 256), dim3(256), 0, st, d, p, dt);
    return hipGetLastError();
}
hipError_t prb_update_particles(const PrbDev& d, hipStream_t st, float* aabb) {
    if (d.np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_update_particles, grid_for(d.np, 256), dim3(256), 0, st, d, aabb);
    return hipGetLastError();
}
hipError_t prb_prepare(const PrbDev& d, hipStream_t st) {
    if (d.nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_prepare, grid_for(d.nb, 256), dim3(256), 0, st, d);
    return hipGetLastError();
}
hipError_t prb_forces(const PrbDev& d, const PrbPhys& p, const int2* sym, const int32_t* entries,
                      const unsigned long long* counts, unsigned long long max_pairs, float dt,
                      hipStream_t st) {
    if (d.np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_boundary, grid_for(d.np, 256), dim3(256), 0, st, d, p, dt);
    // rows <= particles
    hipLaunchKernelGGL(k_pair_forces, grid_for(d.np, 256), dim3(256), 0, st, d, p, sym, entries,
                       counts, max_pairs, dt);
    return hipGetLastError();
}
hipError_t prb_update_velocity(const PrbDev& d, const PrbPhys& p, float dt, hipStream_t st) {
    if (d.nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_update_velocity, grid_for(d.nb, 256), dim3(256), 0, st, d, p, dt);
    return hipGetLastError();
}
hipError_t prb_particle_aabbs(const PrbDev& d, float* out, hipStream_t st) {
    if (d.np <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_particle_aabbs, grid_for(d.np, 256), dim3(256), 0, st, d, out);
    return hipGetLastError();
}

}  // namespace gsap
