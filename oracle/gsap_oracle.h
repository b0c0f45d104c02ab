/*
 * gsap_oracle.h -- CPU oracle for the sweep-and-prune broadphase and the
 * particle rigid-body step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker or the timed CPU baseline.
 *
 * Every function restates the reference algorithm (liufububai/gSap,
 * physicboxSap2.16_Fermi = "PB/") in plain C99 with a fixed evaluation
 * order and no FMA contraction (built with -ffp-contract=off).
 *
 * Pinning (how each function is checked independently of its own formula):
 *   or_flip / or_unflip      exhaustive monotonicity + bijection over all 2^32
 *                            patterns (tests/test_oracle.py)
 *   or_radix_sort_*          equals numpy stable argsort (library routine)
 *   or_moments / or_pca_axis numpy.linalg.eigh on the same fp64 covariance,
 *                            closed-form anisotropic Gaussian
 *   or_axis_coord_maxvar     closed form on hand-built covariances
 *   or_project               containment of every box corner projection
 *   or_sap                   equals or_brute (O(n^2)) on every tested input
 *   or_brute                 hand-built touching / separated cases,
 *                            PB/cudaPhysics/2_1024.txt fixture (sphere mode)
 *   or_sym_csr               PB/cudaPhysics/2_1024.txt pairsEntries lists
 *   or_prb_*                 closed forms: free fall, floor equilibrium,
 *                            2-body momentum, |q| = 1, torque = r x F
 */
#ifndef GSAP_ORACLE_H
#define GSAP_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_KIND_AABB = 0, OR_KIND_SPHERE = 1 };
enum { OR_AXIS_PCA = 0, OR_AXIS_COORD_MAXVAR = 1, OR_AXIS_FIXED = 2, OR_AXIS_AUTO = 4 };

uint32_t or_flip(float f);
float or_unflip(uint32_t u);
void or_flip_array(const float* in, uint32_t* out, size_t n);

/* out[0..2] = sum of centres, out[3..8] = sum c c^T (xx,xy,xz,yy,yz,zz). fp64. */
void or_moments(const float* boxes, int kind, size_t n, double out[9]);
/* Major eigenvector of the covariance (cyclic Jacobi, fp64), sign fixed so the
 * largest-magnitude component is positive, cast to fp32 and renormalised. */
void or_pca_axis(const double mom[9], size_t n, float axis[3], double evals[3]);
/* Reference rule: coordinate axis of largest diagonal variance. */
int or_axis_coord_maxvar(const double mom[9], size_t n, float axis[3], int lamda[3]);

/* Per-coordinate sums and maxima of the box extents (hi - lo; 2r for spheres). */
void or_extents(const float* boxes, int kind, size_t n, double sum[3], double mx[3]);
/* Top eigenvector of P C P, P = I - u u^T: the slab axis of the slab sweep. */
void or_slab_axis(const double mom[9], size_t n, const float u[3], float v[3]);
/* Sweep-test model cost of axis u (v == NULL) or of the slab pair (u, v). */
double or_axis_cost(const double mom[9], const double xsum[3], const double xmax[3], size_t n,
                    const float u[3], const float v[3]);
/* PCA axis (pair) unless a coordinate axis (pair) is strictly cheaper by the
 * model; returns 0 (PCA kept) or 1 (coordinate).  v is written when slabs. */
int or_axis_auto(const double mom[9], const double xsum[3], const double xmax[3], size_t n,
                 int slabs, float u[3], float v[3]);

/* Conservative projection of every box onto axis: keys = flip(lower end),
 * hikeys = flip(upper end). */
void or_project(const float* boxes, int kind, size_t n, const float axis[3],
                uint32_t* keys, uint32_t* hikeys);

/* Stable LSD radix sort of (key, value) pairs, digit_bits in [1,16]. */
void or_radix_sort_kv(uint32_t* keys, uint32_t* vals, size_t n, int digit_bits);
/* Stable LSD radix sort of packed pairs {a,b} (uint32x2, a in the low word)
 * by the logical key (a << bbits) | b over key_bits bits. */
void or_radix_sort_pairs(uint32_t* pairs, size_t n, int bbits, int key_bits, int digit_bits);

/* Exact pair predicates. */
int or_overlap_aabb(const float* a, const float* b);
int or_overlap_sphere(const float* a, const float* b);

/* Full SaP update.  Writes canonical pairs (a<b, lexicographic) into pairs
 * (capacity cap pairs; uint32x2).  Returns the pair count (may exceed cap, in
 * which case only cap pairs are written and the result is unsorted-truncated).
 * tests_out: sum over sorted objects of the sweep interval length.
 * sample_stride > 1 sweeps only sorted objects s with s % stride == 0 (CPU
 * baseline timing of a bounded sample); the pairs found are still sorted. */
int64_t or_sap(const float* boxes, int kind, size_t n, int axis_mode, const float* fixed_axis,
               const uint32_t* group, uint32_t* pairs, uint64_t cap, uint64_t* tests_out,
               float axis_out[3], uint32_t sample_stride);

/* Wall times (s) of the last or_sap phases: setup (axis, projection, sort),
 * sweep, canonical sort. */
void or_last_times(double out[3]);

/* Sweep-test count for a given axis (label source for the learned axis). */
uint64_t or_sweep_tests(const float* boxes, int kind, size_t n, const float axis[3]);

/* The reference's 4-cell sweep plus cross pass (Sap.cpp:211-295), for its work
 * counts: stats = {cell-pass tests, cross-pass tests, relevant boxes, pairs
 * before deduplication}.  Returns the unique pair count. */
int64_t or_sap_ref_cells(const float* boxes, int kind, size_t n, int count_only, uint32_t* pairs,
                         uint64_t cap, uint64_t stats[4]);

/* Brute force O(n^2), canonical output. */
int64_t or_brute(const float* boxes, int kind, size_t n, const uint32_t* group,
                 uint32_t* pairs, uint64_t cap);

/* Lexicographic sort of canonical pairs. */
void or_canon_sort(uint32_t* pairs, uint64_t P);

/* Reference output format: both orientations sorted by (i,j) as int2, plus
 * pairEntries (start offset of each distinct i).  Returns 2P. */
uint64_t or_sym_csr(const uint32_t* canon, uint64_t P, int32_t* pairs_out,
                    int32_t* entries_out, uint32_t* n_entries);

/* ---------------- particle rigid body ---------------- */
typedef struct {
    float gravity[3];
    float sp, dp;          /* reference spring / damping multipliers           */
    float k_abs, c_abs;    /* absolute spring constant / damping (mode abs)    */
    float floor_scale;     /* 0.1 in the reference floor force                 */
    float wall_lo[2], wall_hi[2]; /* x and z walls (active when flags & WALLS) */
    float max_lin_vel;     /* velocity cap (flags & CAP_VEL)                   */
    uint32_t flags;
} or_prb_params;

enum {
    OR_PRB_ABS_SPRING = 1u,   /* k,c absolute instead of mass/dt-derived       */
    OR_PRB_WALLS = 2u,        /* walls at x,z planes                            */
    OR_PRB_FLOOR_CLAMP = 4u,  /* reference y clamp in integrate                 */
    OR_PRB_CAP_VEL = 8u,      /* velocity cap in updateVelocity                 */
    OR_PRB_SKIP_SAME_BODY = 16u, /* ignore pairs within one body                */
    OR_PRB_INERTIA = 32u      /* omega += Iw^-1 tau dt instead of tau dt        */
};

/* Body arrays are float4 (x,y,z,w).  inv_inertia: 12 floats per body (3 rows
 * of float4).  Particle templates: ref_pos float4 (w = radius), ref_mass. */
void or_prb_integrate(float* pos, float* quat, float* linvel, const float* angvel,
                      int nb, float dt, const or_prb_params* p);
void or_prb_update_particles(const float* pos, const float* quat, const float* linvel,
                             const float* angvel, const int32_t* shape, const int32_t* pstart,
                             const int32_t* npart, const float* ref_pos, const float* ref_mass,
                             int max_per_shape, float* ppos, float* pvel, float* pmass, int nb);
void or_prb_prepare(const float* quat, const int32_t* shape, const float* shape_inv_inertia,
                    float* inv_inertia_out, int nb);
void or_prb_forces(const float* ppos, const float* pvel, const float* pmass,
                   const int32_t* sym_pairs, uint64_t n_sym, const int32_t* pbody,
                   float* pforce, int np, float dt, const or_prb_params* p);
void or_prb_update_velocity(float* linvel, float* angvel, const float* quat,
                            const float* inv_mass, const float* inv_inertia,
                            const int32_t* shape, const int32_t* pstart, const int32_t* npart,
                            const float* ref_pos, int max_per_shape, const float* pforce,
                            int nb, float dt, const or_prb_params* p);

#ifdef __cplusplus
}
#endif
#endif
