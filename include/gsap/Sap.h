// gsap/Sap.h -- drop-in for the reference class Sap
// (physicboxSap2.16_Fermi/sapDetection/Sap.h:12-84), header-only over the C ABI
// in gsap.h.  Same constructor and collision() signature; spheres in (float4
// x,y,z,r), the reference output format out (int2 pairs in both orientations
// sorted by (i,j), plus pairEntries row starts).
//
// Differences a caller can observe:
//  - n need not be a multiple of 512 (the reference drops the tail,
//    SapKernel.cu:2583-2585).
//  - collision() has no capacity argument, like the reference; the capacity
//    assumed for `pairs` is setPairCapacity() (default 40 n directed pairs, the
//    reference's own d_pairs scratch size, Sap.cpp:83).  On overflow NPairs and
//    NPEntries are set to 0, as the reference does when its thread budget
//    overflows (SapKernel.cu:2496-2497), and lastStatus() is GSAP_ERR_CAPACITY.
//  - `spacing` seeded the reference's cell width on the first call
//    (Sap.cpp:226-232); the single global sweep has no cells and ignores it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../gsap.h"

class Sap {
public:
    explicit Sap(unsigned int numSpheres, int axisMode = GSAP_AXIS_COORD_MAXVAR)
        : m_n(numSpheres), m_capacity(40ull * numSpheres) {
        gsap_sap_desc d = {};
        d.max_n = numSpheres;
        d.input_kind = GSAP_INPUT_SPHERE;
        d.axis_mode = axisMode;
        d.output_kind = GSAP_OUT_SYM_CSR;
        m_status = gsap_sap_create(&d, &m_h);
        if (m_status != GSAP_OK) std::fprintf(stderr, "Sap: %s\n", gsap_last_error());
    }
    ~Sap() { gsap_sap_destroy(m_h); }
    Sap(const Sap&) = delete;
    Sap& operator=(const Sap&) = delete;

    void setPairCapacity(unsigned long long directedPairs) { m_capacity = directedPairs; }
    // Scratch of one collision() at the current pair capacity, and carving it
    // from caller memory instead of allocating (the reference's constructor
    // allocates its scratch once, Sap.cpp:12-114).
    unsigned long long workspaceSize() {
        uint64_t b = 0;
        m_status = gsap_sap_query_workspace(m_h, m_n, m_capacity, &b);
        return b;
    }
    void setWorkspace(void* d_ws, unsigned long long bytes) {
        m_status = gsap_sap_set_workspace(m_h, d_ws, bytes);
        if (m_status != GSAP_OK) std::fprintf(stderr, "Sap::setWorkspace: %s\n", gsap_last_error());
    }
    void setStream(hipStream_t s) { m_stream = s; }
    int lastStatus() const { return m_status; }
    gsap_sap* handle() { return m_h; }

    // Sap::collision (Sap.cpp:211-335)
    void collision(float spacing, float4* d_particleData, int2* pairs, int* pairsEntries,
                   int* NPairs, int* NPEntries) {
        (void)spacing;
        uint64_t np = 0;
        uint32_t ne = 0;
        m_status = gsap_sap_collide(m_h, m_stream, d_particleData, m_n, nullptr, pairs, m_capacity,
                                    pairsEntries, &np, &ne, nullptr);
        if (m_status == GSAP_OK) {
            *NPairs = (int)np;
            *NPEntries = (int)ne;
        } else {
            std::fprintf(stderr, "Sap::collision: %s\n", gsap_last_error());
            *NPairs = 0;
            *NPEntries = 0;
        }
    }
    // collisionWithSub (Sap.h:23): its body is commented out in the reference
    // (Sap.cpp:337-427); it forwards to collision().
    void collisionWithSub(float spacing, float4* d, int2* pairs, int* entries, int* NPairs,
                          int* NPEntries) {
        collision(spacing, d, pairs, entries, NPairs, NPEntries);
    }

private:
    gsap_sap* m_h = nullptr;
    unsigned int m_n;
    unsigned long long m_capacity;
    hipStream_t m_stream = nullptr;
    int m_status = GSAP_OK;
};
