// gsap_prb_api.cpp -- host side of the gsap_prb_* C ABI (include/gsap.h).
//
// Replaces class ParticleRigidBody (physicboxSap2.16_Fermi/ParticleRigidBody/
// ParticleRigidBody.{h,cpp}) and the PRBWrap launch wrappers
// (ParticleRigidBodyWrapper.h:11-54).  Same state layout (float4 body SoA,
// float4 particles with w = radius, per-shape particle templates) but the
// template table is sized at run time (the reference caps it at
// MAX_NUM_PARTICLES = 10, ParticleRigidBody.h:39), there are no GL buffers, and
// the whole step is stream ordered: the broadphase writes its counts to device
// memory and the force kernel reads them there.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsap.h"
#include "gsap_internal.h"

using namespace gsap;

namespace {
thread_local std::string p_err;
}

struct gsap_prb {
    int32_t max_bodies = 0, max_particles = 0;
    int64_t max_pairs = 0;
    PrbDev d{};
    PrbPhys p{};
    int32_t bp_kind = GSAP_INPUT_SPHERE;
    int2* pairs = nullptr;
    int32_t* entries = nullptr;
    unsigned long long* counts = nullptr;  // device [5]: directed pairs, entries, ...
    float* aabbs = nullptr;
    int32_t ref_capacity = 0, ref_count = 0;
    int32_t n_shapes_cap = 0, n_shapes = 0;
    int32_t max_shape = -1, max_npart = 0;  // host copies for bounds checks
    int64_t num_pairs = 0, num_entries = 0;  // host mirror after set_num_pairs / get_ptrs
};

static int32_t pfail(int32_t code, const std::string& m);

#define PCHK(expr)                                                                 \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return pfail(e_ == hipErrorOutOfMemory ? GSAP_ERR_OOM : GSAP_ERR_HIP,   \
                         std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

// prb failures are reported through gsap_last_error() like every other call.
namespace gsap_err_bridge {
void set(const std::string& m);
}

static int32_t pfail(int32_t code, const std::string& m) {
    p_err = m;
    gsap_err_bridge::set(m);
    return code;
}

extern "C" int32_t gsap_prb_create(int32_t max_bodies, int32_t max_particles, int64_t max_pairs,
                                   gsap_prb** out) {
    if (!out || max_bodies <= 0 || max_particles <= 0 || max_pairs < 0)
        return pfail(GSAP_ERR_INVALID_ARG, "gsap_prb_create: bad sizes");
    *out = nullptr;
    if (gsap_device_count() <= 0) return pfail(GSAP_ERR_HIP, "gsap_prb_create: no HIP device");
    gsap_prb* h = new gsap_prb();
    h->max_bodies = max_bodies;
    h->max_particles = max_particles;
    h->max_pairs = max_pairs;
    const size_t B = (size_t)max_bodies, P = (size_t)max_particles;
    hipError_t e = hipSuccess;
    auto A = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 16);
        if (e == hipSuccess) e = hipMemset(*p, 0, bytes ? bytes : 16);
    };
    A((void**)&h->d.pos, B * 16); A((void**)&h->d.quat, B * 16);
    A((void**)&h->d.linvel, B * 16); A((void**)&h->d.angvel, B * 16);
    A((void**)&h->d.shape, B * 4); A((void**)&h->d.pstart, B * 4); A((void**)&h->d.npart, B * 4);
    A((void**)&h->d.inv_mass, B * 4); A((void**)&h->d.inv_inertia, B * 48);
    A((void**)&h->d.ppos, P * 16); A((void**)&h->d.pvel, P * 16); A((void**)&h->d.pforce, P * 16);
    A((void**)&h->d.pmass, P * 4); A((void**)&h->d.pbody, P * 4); A((void**)&h->d.pip, P * 4);
    A((void**)&h->pairs, (size_t)max_pairs * 8); A((void**)&h->entries, P * 4);
    A((void**)&h->counts, 8 * 8); A((void**)&h->aabbs, P * 24);
    if (e != hipSuccess) {
        gsap_prb_destroy(h);
        return pfail(e == hipErrorOutOfMemory ? GSAP_ERR_OOM : GSAP_ERR_HIP,
                     std::string("gsap_prb_create: ") + hipGetErrorString(e));
    }
    // ParticleRigidBody.cpp:46-47 defaults; gravity of integrateK :55.
    h->p.gravity[0] = 0.0f; h->p.gravity[1] = -5.0f; h->p.gravity[2] = 0.0f;
    h->p.sp = 0.5f; h->p.dp = 0.25f;
    h->p.k_abs = 1e4f; h->p.c_abs = 50.0f;
    h->p.floor_scale = 0.1f;
    h->p.wall_lo[0] = h->p.wall_lo[1] = 0.0f;
    h->p.wall_hi[0] = h->p.wall_hi[1] = 256.0f;
    h->p.max_lin_vel = 1e10f;
    h->p.flags = 0;
    *out = h;
    return GSAP_OK;
}

extern "C" int32_t gsap_prb_destroy(gsap_prb* h) {
    if (!h) return GSAP_OK;
    (void)hipDeviceSynchronize();
    void* ptrs[] = {h->d.pos, h->d.quat, h->d.linvel, h->d.angvel, h->d.shape, h->d.pstart,
                    h->d.npart, h->d.inv_mass, h->d.inv_inertia, h->d.ref_pos, h->d.ref_mass,
                    h->d.shape_inv_inertia, h->d.ppos, h->d.pvel, h->d.pforce, h->d.pmass,
                    h->d.pbody, h->d.pip, h->pairs, h->entries, h->counts, h->aabbs};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete h;
    return GSAP_OK;
}

extern "C" int32_t gsap_prb_set_rb_data(gsap_prb* h, int32_t nb, const float* pos, const float* quat,
                                        const float* linvel, const float* angvel,
                                        const int32_t* shape, const int32_t* pstart,
                                        const int32_t* npart, const float* inv_mass,
                                        const float* inv_inertia, int32_t total_particles) {
    if (!h || nb < 0 || nb > h->max_bodies || total_particles < 0 ||
        total_particles > h->max_particles || !pos || !quat || !linvel || !angvel || !shape ||
        !pstart || !npart || !inv_mass)
        return pfail(GSAP_ERR_INVALID_ARG, "gsap_prb_set_rb_data: bad argument");
    // setRbData normalises the quaternions on the host (ParticleRigidBody.cpp:82-86).
    std::vector<float> q(quat, quat + 4 * (size_t)nb);
    for (int i = 0; i < nb; ++i) {
        float* v = &q[4 * (size_t)i];
        float l = std::sqrt(((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3]);
        float inv = 1.0f / l;
        for (int k = 0; k < 4; ++k) v[k] = v[k] * inv;
    }
    std::vector<int32_t> pb((size_t)total_particles, 0), pip((size_t)total_particles, 0);
    for (int b = 0; b < nb; ++b) {
        if (npart[b] < 0 || pstart[b] < 0 || pstart[b] + npart[b] > total_particles)
            return pfail(GSAP_ERR_INVALID_ARG, "gsap_prb_set_rb_data: particle range out of bounds");
        for (int k = 0; k < npart[b]; ++k) {
            pb[(size_t)pstart[b] + k] = b;
            pip[(size_t)pstart[b] + k] = k;
        }
    }
    const size_t B = (size_t)nb;
    PCHK(hipMemcpy(h->d.pos, pos, B * 16, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.quat, q.data(), B * 16, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.linvel, linvel, B * 16, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.angvel, angvel, B * 16, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.shape, shape, B * 4, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.pstart, pstart, B * 4, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.npart, npart, B * 4, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.inv_mass, inv_mass, B * 4, hipMemcpyHostToDevice));
    if (inv_inertia) PCHK(hipMemcpy(h->d.inv_inertia, inv_inertia, B * 48, hipMemcpyHostToDevice));
    if (total_particles) {
        PCHK(hipMemcpy(h->d.pbody, pb.data(), pb.size() * 4, hipMemcpyHostToDevice));
        PCHK(hipMemcpy(h->d.pip, pip.data(), pip.size() * 4, hipMemcpyHostToDevice));
    }
    int32_t ms = -1, mp = 0;
    for (int b = 0; b < nb; ++b) {
        if (shape[b] < 0) return pfail(GSAP_ERR_INVALID_ARG, "gsap_prb_set_rb_data: negative shape");
        ms = shape[b] > ms ? shape[b] : ms;
        mp = npart[b] > mp ? npart[b] : mp;
    }
    h->max_shape = ms;
    h->max_npart = mp;
    h->d.nb = nb;
    h->d.np = total_particles;
    return GSAP_OK;
}

// Every body's particles must index inside the template table.
static int32_t check_templates(gsap_prb* h) {
    if (!h) return pfail(GSAP_ERR_INVALID_ARG, "null handle");
    if (h->d.nb == 0) return GSAP_OK;
    if (!h->d.ref_pos) return pfail(GSAP_ERR_INVALID_ARG, "particle templates not set");
    if (h->max_npart > h->d.max_per_shape ||
        (int64_t)h->max_shape * h->d.max_per_shape + h->max_npart > h->ref_count)
        return pfail(GSAP_ERR_INVALID_ARG, "body shapes index past the particle templates");
    return GSAP_OK;
}

extern "C" int32_t gsap_prb_set_particle_data(gsap_prb* h, const float* ref_pos, const float* ref_mass,
                                              int32_t n, int32_t max_per_shape) {
    if (!h || !ref_pos || !ref_mass || n <= 0 || max_per_shape <= 0)
        return pfail(GSAP_ERR_INVALID_ARG, "gsap_prb_set_particle_data: bad argument");
    if (n > h->ref_capacity) {
        if (h->d.ref_pos) (void)hipFree(h->d.ref_pos);
        if (h->d.ref_mass) (void)hipFree(h->d.ref_mass);
        h->d.ref_pos = nullptr;
        h->d.ref_mass = nullptr;
        PCHK(hipMalloc((void**)&h->d.ref_pos, (size_t)n * 16));
        PCHK(hipMalloc((void**)&h->d.ref_mass, (size_t)n * 4));
        h->ref_capacity = n;
    }
    PCHK(hipMemcpy(h->d.ref_pos, ref_pos, (size_t)n * 16, hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->d.ref_mass, ref_mass, (size_t)n * 4, hipMemcpyHostToDevice));
    h->d.max_per_shape = max_per_shape;
    h->ref_count = n;
    return GSAP_OK;
}

extern "C" int32_t gsap_prb_set_shape_inv_inertia(gsap_prb* h, const float* inv_inertia, int32_t ns) {
    if (!h || !inv_inertia || ns <= 0) return pfail(GSAP_ERR_INVALID_ARG, "bad argument");
    if (ns > h->n_shapes_cap) {
        if (h->d.shape_inv_inertia) (void)hipFree(h->d.shape_inv_inertia);
        h->d.shape_inv_inertia = nullptr;
        PCHK(hipMalloc((void**)&h->d.shape_inv_inertia, (size_t)ns * 48));
        h->n_shapes_cap = ns;
    }
    PCHK(hipMemcpy(h->d.shape_inv_inertia, inv_inertia, (size_t)ns * 48, hipMemcpyHostToDevice));
    h->n_shapes = ns;
    return GSAP_OK;
}

extern "C" int32_t gsap_prb_set_params(gsap_prb* h, float sp, float dp) {
    if (!h) return pfail(GSAP_ERR_INVALID_ARG, "null handle");
    h->p.sp = sp;
    h->p.dp = dp;
    return GSAP_OK;
}

// setMaxVelocity (ParticleRigidBody.h:88): maxLinVel = minRadius / dt, cap on.
extern "C" int32_t gsap_prb_set_max_velocity(gsap_prb* h, float min_radius, float dt) {
    if (!h || !(dt > 0.0f)) return pfail(GSAP_ERR_INVALID_ARG, "bad argument");
    h->p.max_lin_vel = min_radius / dt;
    h->p.flags |= GSAP_PRB_CAP_VEL;
    return GSAP_OK;
}

extern "C" int32_t gsap_prb_set_physics(gsap_prb* h, const gsap_prb_physics* ph) {
    if (!h || !ph) return pfail(GSAP_ERR_INVALID_ARG, "null argument");
    if (ph->broadphase_kind != GSAP_INPUT_SPHERE && ph->broadphase_kind != GSAP_INPUT_AABB)
        return pfail(GSAP_ERR_INVALID_ARG, "bad broadphase_kind");
    for (int k = 0; k < 3; ++k) h->p.gravity[k] = ph->gravity[k];
    h->p.k_abs = ph->k_abs;
    h->p.c_abs = ph->c_abs;
    h->p.floor_scale = ph->floor_scale;
    for (int k = 0; k < 2; ++k) {
        h->p.wall_lo[k] = ph->wall_lo[k];
        h->p.wall_hi[k] = ph->wall_hi[k];
    }
    h->p.flags = ph->flags | (h->p.flags & GSAP_PRB_CAP_VEL);
    h->bp_kind = ph->broadphase_kind;
    return GSAP_OK;
}

#define PRBLAUNCH(h, expr)                                                 \
    do {                                                                   \
        if (!(h)) return pfail(GSAP_ERR_INVALID_ARG, "null handle");       \
        PCHK(expr);                                                        \
        return GSAP_OK;                                                    \
    } while (0)

extern "C" int32_t gsap_prb_update_particle_data(gsap_prb* h, void* st) {
    if (int32_t rc = check_templates(h)) return rc;
    PRBLAUNCH(h, prb_update_particles(h->d, (hipStream_t)st));
}
extern "C" int32_t gsap_prb_prepare(gsap_prb* h, void* st) {
    if (h && h->d.nb && (!h->d.shape_inv_inertia || h->max_shape >= h->n_shapes))
        return pfail(GSAP_ERR_INVALID_ARG, "shape inertia not set for every shape");
    PRBLAUNCH(h, prb_prepare(h->d, (hipStream_t)st));
}
extern "C" int32_t gsap_prb_compute_force(gsap_prb* h, void* st, float dt) {
    PRBLAUNCH(h, prb_forces(h->d, h->p, h->pairs, h->entries, h->counts,
                            (unsigned long long)h->max_pairs, dt, (hipStream_t)st));
}
extern "C" int32_t gsap_prb_update_velocity(gsap_prb* h, void* st, float dt) {
    if (int32_t rc = check_templates(h)) return rc;
    PRBLAUNCH(h, prb_update_velocity(h->d, h->p, dt, (hipStream_t)st));
}
extern "C" int32_t gsap_prb_integrate(gsap_prb* h, void* st, float dt) {
    PRBLAUNCH(h, prb_integrate(h->d, h->p, dt, (hipStream_t)st));
}

// The broadphase on the particles; boxes_ready: the particle AABBs were
// written with the particle data (the whole step)
static int32_t prb_collide(gsap_prb* h, void* st, gsap_sap* sap, bool boxes_ready) {
    if (!h || !sap) return pfail(GSAP_ERR_INVALID_ARG, "null argument");
    const void* in = h->d.ppos;
    if (h->bp_kind == GSAP_INPUT_AABB) {
        if (!boxes_ready) PCHK(prb_particle_aabbs(h->d, h->aabbs, (hipStream_t)st));
        in = h->aabbs;
    }
    // The body index is the group id; a SaP handle created with
    // GSAP_FLAG_SKIP_SAME_GROUP drops intra-body pairs, others ignore it.
    const uint32_t* grp = (const uint32_t*)h->d.pbody;
    return gsap_sap_collide_async(sap, st, in, (uint32_t)h->d.np, grp, h->pairs,
                                  (uint64_t)h->max_pairs, h->entries, (uint64_t*)h->counts);
}

extern "C" int32_t gsap_prb_collide(gsap_prb* h, void* st, gsap_sap* sap) {
    return prb_collide(h, st, sap, false);
}

extern "C" int32_t gsap_prb_set_num_pairs(gsap_prb* h, int64_t num_pairs, int64_t num_entries) {
    if (!h || num_pairs < 0 || num_entries < 0 || num_pairs > h->max_pairs ||
        num_entries > h->max_particles)
        return pfail(GSAP_ERR_INVALID_ARG, "gsap_prb_set_num_pairs: out of range");
    unsigned long long c[5] = {(unsigned long long)num_pairs, (unsigned long long)num_entries, 0, 0, 0};
    PCHK(hipMemcpy(h->counts, c, sizeof c, hipMemcpyHostToDevice));
    h->num_pairs = num_pairs;
    h->num_entries = num_entries;
    return GSAP_OK;
}

extern "C" int32_t gsap_prb_step(gsap_prb* h, void* st, gsap_sap* sap, float dt) {
    int32_t rc;
    if ((rc = gsap_prb_integrate(h, st, dt))) return rc;
    // the particle data and, for an AABB broadphase, the particle boxes in
    // one pass
    if ((rc = check_templates(h))) return rc;
    const bool boxes = h->bp_kind == GSAP_INPUT_AABB;
    PCHK(prb_update_particles(h->d, (hipStream_t)st, boxes ? h->aabbs : nullptr));
    if ((rc = prb_collide(h, st, sap, boxes))) return rc;
    if (h->p.flags & GSAP_PRB_INERTIA)
        if ((rc = gsap_prb_prepare(h, st))) return rc;
    if ((rc = gsap_prb_compute_force(h, st, dt))) return rc;
    return gsap_prb_update_velocity(h, st, dt);
}

extern "C" int32_t gsap_prb_get_ptrs(gsap_prb* h, gsap_prb_ptrs* o) {
    if (!h || !o) return pfail(GSAP_ERR_INVALID_ARG, "null argument");
    unsigned long long c[5];
    PCHK(hipDeviceSynchronize());
    PCHK(hipMemcpy(c, h->counts, sizeof c, hipMemcpyDeviceToHost));
    o->rb_pos = (float*)h->d.pos;
    o->rb_quat = (float*)h->d.quat;
    o->rb_linvel = (float*)h->d.linvel;
    o->rb_angvel = (float*)h->d.angvel;
    o->rb_inv_inertia = (float*)h->d.inv_inertia;
    o->p_pos = (float*)h->d.ppos;
    o->p_vel = (float*)h->d.pvel;
    o->p_force = (float*)h->d.pforce;
    o->p_mass = h->d.pmass;
    o->p_body = h->d.pbody;
    o->pairs = (int32_t*)h->pairs;
    o->pair_entries = h->entries;
    o->num_bodies = h->d.nb;
    o->num_particles = h->d.np;
    o->num_pairs = (int64_t)c[0];
    o->num_pair_entries = (int64_t)c[1];
    o->rb_shape_idx = h->d.shape;
    o->rb_num_particles = h->d.npart;
    if (c[3] & 1ull) return pfail(GSAP_ERR_CAPACITY, "pair buffer overflow in the last broadphase");
    return GSAP_OK;
}
