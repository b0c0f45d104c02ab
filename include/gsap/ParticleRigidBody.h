// gsap/ParticleRigidBody.h -- drop-in for the reference class ParticleRigidBody
// (physicboxSap2.16_Fermi/ParticleRigidBody/ParticleRigidBody.h:46-165),
// header-only over gsap_prb_* in gsap.h.  Method names are the reference's,
// including comupteFoceOnParticles.  RigidBodyData holds HOST arrays (the
// reference's setRbData copies them host->device, ParticleRigidBody.cpp:101-109);
// quaternions are Quaternion (a float4 x,y,z,w) and inverse inertias
// Matrix3x3 (3 float4 rows), the reference's types (gsap/quaternion.h,
// gsap/matrix3x3.h).  No GL buffers are created.
#pragma once
#include <hip/hip_runtime.h>

#include <cassert>
#include <cstdio>

#include "../gsap.h"
#include "Sap.h"
#include "matrix3x3.h"
#include "quaternion.h"

// Template slots per shape and shape count of the reference's fixed template
// table (ParticleRigidBody.h:39-40); here only the default layout of
// setParticleData, the table itself is sized at run time.
#ifndef MAX_NUM_PARTICLES
#define MAX_NUM_PARTICLES 10
#endif
#ifndef MAX_PARTICLE_SHAPE_TYPE
#define MAX_PARTICLE_SHAPE_TYPE 4
#endif

class ParticleRigidBody {
public:
    struct RigidBodyData {
        int m_nBodies;
        float4* m_pos;
        Quaternion* m_quat;
        float4* m_linVel;
        float4* m_angVel;
        int* m_shapeIdx;
        int* m_pBufferStartOffset;
        int* m_numParticles;
        float* m_invMass;
        Matrix3x3* m_invInertia;
        int m_numTotalParticles;
    };

    ParticleRigidBody(int maxNumBodies, int maxNumParticles, int maxPairs)
        : m_maxPairs(maxPairs) {
        check(gsap_prb_create(maxNumBodies, maxNumParticles, maxPairs, &m_h));
    }
    ~ParticleRigidBody() { gsap_prb_destroy(m_h); }
    ParticleRigidBody(const ParticleRigidBody&) = delete;
    ParticleRigidBody& operator=(const ParticleRigidBody&) = delete;

    void setRbData(RigidBodyData& d) {
        check(gsap_prb_set_rb_data(m_h, d.m_nBodies, (const float*)d.m_pos, (const float*)d.m_quat,
                                   (const float*)d.m_linVel, (const float*)d.m_angVel, d.m_shapeIdx,
                                   d.m_pBufferStartOffset, d.m_numParticles, d.m_invMass,
                                   (const float*)d.m_invInertia, d.m_numTotalParticles));
    }
    // setParticleData(float4* pos, float* mass, int numElements); templates are
    // laid out maxPerShape slots per shape (reference: MAX_NUM_PARTICLES = 10).
    void setParticleData(float4* pos, float* mass, int numElements,
                         int maxPerShape = MAX_NUM_PARTICLES) {
        check(gsap_prb_set_particle_data(m_h, (const float*)pos, mass, numElements, maxPerShape));
    }
    void setParticleShapeInvInertia(Matrix3x3* invInertia, int numShapes) {
        check(gsap_prb_set_shape_inv_inertia(m_h, (const float*)invInertia, numShapes));
    }
    void setParams(float spMultip, float dpMultip) { check(gsap_prb_set_params(m_h, spMultip, dpMultip)); }
    void setMaxVelocity(float minRadius, float dt) { check(gsap_prb_set_max_velocity(m_h, minRadius, dt)); }
    void setPhysics(const gsap_prb_physics& p) { check(gsap_prb_set_physics(m_h, &p)); }

    void updateParticleData() { check(gsap_prb_update_particle_data(m_h, m_stream)); }
    void prepare() { check(gsap_prb_prepare(m_h, m_stream)); }
    void comupteFoceOnParticles(float dt) { check(gsap_prb_compute_force(m_h, m_stream, dt)); }
    void updateVelocity(float dt) { check(gsap_prb_update_velocity(m_h, m_stream, dt)); }
    void integrate(float dt) { check(gsap_prb_integrate(m_h, m_stream, dt)); }

    // Getters return device pointers, as in the reference.
    int getNumPairs() { refresh(); return (int)m_p.num_pairs; }
    void setNumPairs(int numPairs) {
        assert(numPairs < m_maxPairs);
        m_pendingPairs = numPairs;
    }
    int2* getPairs() { refresh(); return (int2*)m_p.pairs; }
    int* getPairEntries() { refresh(); return m_p.pair_entries; }
    int getNumPairEntries() { refresh(); return (int)m_p.num_pair_entries; }
    void setNumPairEntries(int numPairEntries) {
        check(gsap_prb_set_num_pairs(m_h, m_pendingPairs, numPairEntries));
    }
    float4* getRbPos() { refresh(); return (float4*)m_p.rb_pos; }
    Quaternion* getRbQuat() { refresh(); return (Quaternion*)m_p.rb_quat; }
    int* getRbShapeIdx() { refresh(); return m_p.rb_shape_idx; }
    int* getRbNumParticles() { refresh(); return m_p.rb_num_particles; }
    int getNumParticles() { refresh(); return m_p.num_particles; }
    float4* getParticlePos() { refresh(); return (float4*)m_p.p_pos; }
    float4* getParticleVel() { refresh(); return (float4*)m_p.p_vel; }
    float4* getParticleForce() { refresh(); return (float4*)m_p.p_force; }
    int getNumBodies() { refresh(); return m_p.num_bodies; }

    // Template index of particle idx of a body of shape shapeIdx in the
    // default layout (ParticleRigidBody.h:143-144).
    __host__ __device__ static int getParticleRefIdx(int shapeIdx, int idx) {
        return MAX_NUM_PARTICLES * shapeIdx + idx;
    }

    // Whole DemoParticles::step sequence (demoParticles.h:856-928) with the
    // broadphase on this object's particles through `sap`.
    void step(Sap& sap, float dt) { check(gsap_prb_step(m_h, m_stream, sap.handle(), dt)); }
    void setStream(hipStream_t s) { m_stream = s; }
    gsap_prb* handle() { return m_h; }

private:
    void refresh() { check(gsap_prb_get_ptrs(m_h, &m_p)); }
    static void check(int rc) {
        if (rc != GSAP_OK) std::fprintf(stderr, "ParticleRigidBody: %s\n", gsap_last_error());
    }
    gsap_prb* m_h = nullptr;
    gsap_prb_ptrs m_p = {};
    hipStream_t m_stream = nullptr;
    int m_maxPairs;
    int m_pendingPairs = 0;
};
