/*
 * gsap.h -- C ABI of the MI355X (gfx950) sweep-and-prune broadphase and
 * particle rigid-body step.
 *
 * Conventions
 *  - Every entry point returns an int32 status (GSAP_OK = 0).  The message of
 *    the last failure on the calling thread is returned by gsap_last_error().
 *  - All array arguments named d_* are device pointers owned by the caller;
 *    h_* are host pointers.  `stream` is a hipStream_t passed as void*
 *    (NULL = the null stream).
 *  - A handle is not thread-safe: use one handle per stream.
 *  - No torch or C++ types cross this boundary.
 *
 * Reference interfaces replaced (paths relative to physicboxSap2.16_Fermi/):
 *  gsap_sap_*         class Sap (sapDetection/Sap.h:12-84; collision
 *                     Sap.cpp:211-335) and its extern "C" stage wrappers
 *                     (sapDetection/SapKernel.cuh:4-56)
 *  gsap_radix_sort_*  nvRadixSort::RadixSort (sapDetection/radixsort.h:110-278)
 *  gsap_prb_*         class ParticleRigidBody (ParticleRigidBody/
 *                     ParticleRigidBody.h:46-165) and PRBWrap::*
 *                     (ParticleRigidBody/ParticleRigidBodyWrapper.h:11-54)
 */
#ifndef GSAP_H
#define GSAP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSAP_ABI_VERSION 2

/* ------------------------------------------------------------------ status */
enum {
    GSAP_OK = 0,
    GSAP_ERR_INVALID_ARG = 1, /* bad pointer / size / mode                    */
    GSAP_ERR_CAPACITY = 2,    /* output too small; required size returned     */
    GSAP_ERR_NONFINITE = 3,   /* NaN or Inf in the input boxes                */
    GSAP_ERR_HIP = 4,         /* a HIP runtime call failed                    */
    GSAP_ERR_UNSUPPORTED = 5, /* mode not available (e.g. no trained model)   */
    GSAP_ERR_OOM = 6          /* device allocation failed                      */
};

/* Message of the last failure of any gsap_* call on the calling thread. */
const char* gsap_last_error(void);
/* Library version string (build id). */
const char* gsap_version(void);
/* Number of visible devices (0 when no GPU); never fails. */
int32_t gsap_device_count(void);

/* ------------------------------------------------------------------ SaP */
enum { GSAP_INPUT_AABB = 0,   /* 6 floats per object: lo.x lo.y lo.z hi.x hi.y hi.z */
       GSAP_INPUT_SPHERE = 1  /* float4 per object: x y z r (reference layout)      */ };

enum { GSAP_AXIS_PCA = 0,          /* major eigenvector of the fp64 covariance     */
       GSAP_AXIS_COORD_MAXVAR = 1, /* reference rule: coordinate axis of max var.  */
       GSAP_AXIS_FIXED = 2,        /* desc.fixed_axis                              */
       GSAP_AXIS_LEARNED = 3,      /* learned selector, PCA fallback               */
       /* Cheapest of the PCA axis and the three coordinate axes (with
        * GSAP_FLAG_SLABS: of the PCA pair and the six coordinate pairs) by
        * the sweep-test model E[extent on u] / sqrt(var u), or E[extent on u] x
        * (largest extent on v) / sqrt(det cov(u, v)) for slabs, from the same
        * fp64 moments plus per-coordinate mean and largest extents.  No
        * reference counterpart; axis_source reports PCA or COORD_MAXVAR. */
       GSAP_AXIS_AUTO = 4 };

enum { GSAP_OUT_CANON = 0,  /* uint32x2 (i<j), lexicographic, each pair once    */
       GSAP_OUT_SYM_CSR = 1 /* reference format: int32x2 both orientations,
                               sorted by (i,j), plus pairEntries row starts     */ };

enum {
    GSAP_FLAG_SKIP_SAME_GROUP = 1u, /* drop pairs whose d_group ids are equal              */
    /* Sweep inside slabs of the second principal axis (the reference's cell split,
     * SapKernel.cu:150-213 and 2634-2678, generalised to up to 255 slabs with no
     * cross pass): each box is keyed into the one or two slabs its projection
     * touches and a pair is reported only in the slab of the larger lower bound.
     * Same pair set, far fewer sweep tests on spread-out scenes.  Batched
     * updates split each scene with its own slab axis and parameters; n <= 2^30. */
    GSAP_FLAG_SLABS = 2u,
    /* Temporal coherence for successive updates of one object set (the
     * reference carries only its cell interval between frames, Sap.cpp:226-243).
     * The key function (axes, slab and key parameters) of a refresh is kept
     * for the next 63 single-scene updates of the same n, and the object
     * sort starts from the previous update's order: blocks of it keep the
     * entries that stay in their previous key range and fix their local
     * order in LDS; the entries that moved farther are sorted on their own
     * and merged in.  The pair list is identical to a from-scratch update's;
     * too many far movers fall back to the regular sort on the device.
     * Batched updates ignore the flag. */
    GSAP_FLAG_COHERENT = 4u
};

typedef struct gsap_sap gsap_sap;

typedef struct gsap_sap_desc {
    uint32_t max_n;      /* objects per call the workspace is sized for (grows) */
    int32_t input_kind;  /* GSAP_INPUT_*                                          */
    int32_t axis_mode;   /* GSAP_AXIS_*                                           */
    int32_t output_kind; /* GSAP_OUT_*                                            */
    float fixed_axis[3]; /* GSAP_AXIS_FIXED (need not be normalised)              */
    uint32_t flags;      /* GSAP_FLAG_*                                           */
    int32_t digit_bits;  /* radix digit width of both sorts (0 = default 8; 6..8) */
    int32_t sort_items;  /* keys per thread in a sort tile (0 = default 8)        */
} gsap_sap_desc;

typedef struct gsap_stats {
    uint64_t num_pairs;      /* unordered pairs found                             */
    uint64_t sweep_tests;    /* sum over sorted objects of candidates inside the
                                projected interval (the sweep-test count)         */
    float axis[3];           /* sweep axis used                                   */
    int32_t axis_source;     /* GSAP_AXIS_* actually used (LEARNED may fall back) */
    float ms_total;          /* hipEvent time of the whole update (stream-ordered)*/
    uint32_t n;              /* objects in the call                               */
} gsap_stats;

/* Sap(uint n) (Sap.h:15): creates a handle and its workspace. */
int32_t gsap_sap_create(const gsap_sap_desc* desc, gsap_sap** out);
/* ~Sap() */
int32_t gsap_sap_destroy(gsap_sap* h);

/* Sap::collision (Sap.cpp:211-335).  One complete update from device boxes to
 * a device pair list.
 *   d_in           n objects (layout by desc.input_kind)
 *   d_group        optional n uint32 group ids (GSAP_FLAG_SKIP_SAME_GROUP)
 *   d_pairs        CANON: uint32x2 pairs; SYM_CSR: int32x2 directed pairs
 *   pair_capacity  capacity of d_pairs in pairs (CANON) or directed pairs
 *                  (SYM_CSR); 0 or d_pairs == NULL => count only (pre-count)
 *   d_entries      SYM_CSR only: int32 row starts, capacity n
 *   h_num_pairs    out: P (CANON) or 2P (SYM_CSR); on GSAP_ERR_CAPACITY the
 *                  required capacity
 *   h_num_entries  out (SYM_CSR): number of distinct first indices
 *   stats          optional
 * Blocks until the counts are on the host (one device->host copy). */
int32_t gsap_sap_collide(gsap_sap* h, void* stream, const void* d_in, uint32_t n,
                         const uint32_t* d_group, void* d_pairs, uint64_t pair_capacity,
                         int32_t* d_entries, uint64_t* h_num_pairs, uint32_t* h_num_entries,
                         gsap_stats* stats);

/* Stream-ordered variant: nothing is copied to the host and the call does not
 * block.  d_counts (device, 5 x uint64) receives [0] = P or 2P, [1] = entries,
 * [2] = sweep tests, [3] = status bits (1 = capacity overflow, 2 = non-finite
 * input), [4] = candidate tests executed.  Suitable for hipGraph capture. */
int32_t gsap_sap_collide_async(gsap_sap* h, void* stream, const void* d_in, uint32_t n,
                               const uint32_t* d_group, void* d_pairs, uint64_t pair_capacity,
                               int32_t* d_entries, uint64_t* d_counts);

/* Allocation-free variant of gsap_sap_collide_async.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and status bits; the update's scratch comes from a static
 * device array of GSAP_NOALLOC_SCRATCH_BYTES per GPU (gsap_sap_query_workspace
 * gives an update's need), so the call makes no hipMalloc/hipFree and leaves
 * the handle's own buffers untouched.  An update that needs more returns
 * GSAP_ERR_CAPACITY before any device work.  Every call on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (other streams, other
 * handles, other threads) race and must be serialised by the caller.  The
 * handle itself is created beforehand (gsap_sap_create allocates). */
#define GSAP_NOALLOC_SCRATCH_BYTES (64ull << 20)
int32_t gsap_sap_collide_noalloc(gsap_sap* h, void* stream, const void* d_in, uint32_t n,
                                 const uint32_t* d_group, void* d_pairs, uint64_t pair_capacity,
                                 int32_t* d_entries, uint64_t* d_counts);

/* Batched update over many independent scenes packed back to back.
 *   h_scene_offsets  n_scenes+1 non-decreasing offsets into d_in (host
 *                    array); a scene may be empty or hold a single box
 *   d_pairs          CANON pairs with GLOBAL indices (into d_in), sorted,
 *                    so scene s occupies d_scene_pair_offsets[s..s+1)
 *   d_scene_pair_offsets  device, n_scenes+1 uint64
 * One launch per stage covers every scene; the stage list is recorded into a
 * hipGraph on first use for a given scene table and replayed afterwards. */
int32_t gsap_sap_collide_batched(gsap_sap* h, void* stream, const void* d_in,
                                 const uint32_t* h_scene_offsets, uint32_t n_scenes,
                                 void* d_pairs, uint64_t pair_capacity,
                                 uint64_t* d_scene_pair_offsets, uint64_t* h_num_pairs,
                                 gsap_stats* stats);

/* Workspace.  Sap::Sap (Sap.cpp:12-114) allocates its scratch once, sized by
 * the object count.  Here the scratch of an update depends on n, the pair
 * capacity and the launch configuration (gsap_tune), and is allocated and
 * grown by the handle unless the caller provides it:
 *   gsap_sap_query_workspace  bytes an update of n objects with room for
 *                             pair_capacity pairs (0: count only) needs under
 *                             the handle's current configuration
 *   gsap_sap_set_workspace    use d_ws (device, 256-byte aligned, `bytes`
 *                             long) for every later update's scratch instead
 *                             of allocating; d_ws == NULL returns to internal
 *                             allocation.  Internal buffers are freed.  The
 *                             memory must stay valid until the handle is
 *                             destroyed or another workspace is set.  An update
 *                             that needs more returns GSAP_ERR_CAPACITY with
 *                             the required size in the message.  Setting it
 *                             synchronises the device.
 * Batched updates need the scene tables of every scene in it as well; the
 * learned model and the judge's axis table stay internal allocations. */
int32_t gsap_sap_query_workspace(gsap_sap* h, uint32_t n, uint64_t pair_capacity, uint64_t* bytes);
int32_t gsap_sap_set_workspace(gsap_sap* h, void* d_ws, uint64_t bytes);

/* Per-stage timing of the last gsap_sap_collide with timing enabled (ms,
 * hipEvents on the launch stream).  Stages, in order: moments+project,
 * object_sort, gather, sweep, then (unless count-only) pair_hist, one entry
 * per pair-sort onesweep pass, pair_local (only with gsap_tune.pair_local 1),
 * pair_tail (row starts / scene offsets); last: finish.  With pair_local 2 the
 * row sort takes the place of pair_hist and there are no passes. */
int32_t gsap_sap_stage_times(gsap_sap* h, float* h_ms, int32_t max_stages, int32_t* n_stages);
int32_t gsap_sap_set_timing(gsap_sap* h, int32_t enable);

/* ------------------------------------------------------------------ autotuner */
/* Launch configuration of a SaP handle.  The reference's analogue is
 * computeNumCTAs (sapDetection/radixsort.cpp:49-117), which sized the
 * persistent grids of its pre-Fermi sort from the device; here every knob is
 * measured on the actual update. */
typedef struct gsap_tune {
    int32_t sort_items;  /* object sort: keys per thread, 4/8/12/16 (tile = 256 x items) */
    int32_t sort_bits;   /* object sort: digit width, 6/7/8                               */
    int32_t pair_items;  /* pair sort: keys per thread, 8/16                              */
    int32_t pair_bits;   /* pair sort: digit width, 6/7/8                                 */
    int32_t pair_grid;   /* pair sort: persistent workgroups per pass, 256..16384         */
    int32_t slabs;       /* 1: slab sweep (GSAP_FLAG_SLABS), per scene when batched        */
    int32_t sweep_chunk; /* sweep: sub-tiles per XCD chunk, 0 (round-robin)/4/16/64       */
    int32_t pair_regions;/* sweep output regions with one allocation counter each, 1 or 8;
                          * 8 holds 8x the pair capacity in device memory.  A handle
                          * whose configuration was never set or tuned switches to 8
                          * at its first batched update with pairs                      */
    int32_t pair_local;  /* pair sort: 0 = onesweep passes over the whole key; 1 = onesweep
                          * passes over the high key bits only (8-bit digits, pair_bits
                          * unused), then each run of equal high bits sorted in LDS; 2 =
                          * by rows: a counting sort on the first index, then each row by
                          * the second (for lists of about one pair per object; objects
                          * 2^8 .. 2^24, else the whole-key passes run)                    */
    float ms;            /* out: measured update time of this configuration (autotune)    */
    int32_t from_cache;  /* out: 1 if the configuration came from the cache file          */
} gsap_tune;
int32_t gsap_sap_get_tune(gsap_sap* h, gsap_tune* out);
int32_t gsap_sap_set_tune(gsap_sap* h, const gsap_tune* t);
/* Time the update of this input under candidate configurations (a staged
 * search: sweep kind, then object sort items x digit width, then pair sort
 * items x digit width x grid, with and without the local stage), keep the
 * fastest in the handle and return it.
 * h_scene_offsets == NULL tunes the single-scene update of d_in[0..n)
 * (d_group as in gsap_sap_collide); otherwise the batched update of n_scenes
 * scenes (n and d_group are ignored).
 * cache_path (optional) is a text file keyed by device, input and output
 * kind, axis mode, flags and size bucket: a hit sets the stored configuration
 * without measuring, a miss appends the measured one.  Allocates temporary
 * device memory for the pairs it produces while measuring. */
int32_t gsap_sap_autotune(gsap_sap* h, void* stream, const void* d_in, uint32_t n,
                          const uint32_t* d_group, const uint32_t* h_scene_offsets,
                          uint32_t n_scenes, const char* cache_path, gsap_tune* out);

/* ------------------------------------------------------------------ learned axis */
/* No reference counterpart.  A 64-128-128-4 MLP (bf16 weights/activations,
 * fp32 accumulation, hand-written HIP forward/backward/Adam) maps 64 per-scene
 * features to a sweep axis (snapped to the nearest of 64 fixed candidates) and
 * a log pair count.  GSAP_AXIS_LEARNED uses it per scene and keeps the PCA axis
 * where the prediction is not confident.  The axis never changes the pair set. */
typedef struct gsap_mlp gsap_mlp;
int32_t gsap_mlp_create(uint64_t seed, gsap_mlp** out);
int32_t gsap_mlp_destroy(gsap_mlp* m);
int32_t gsap_mlp_num_params(void);
/* fp32 parameters: W1[64][128] b1[128] W2[128][128] b2[128] W3[128][4] b3[4] */
int32_t gsap_mlp_set_params(gsap_mlp* m, const float* h_params, uint32_t n);
int32_t gsap_mlp_get_params(gsap_mlp* m, float* h_params, uint32_t n);
/* d_x: n x 64 fp32 features (rounded to bf16 on input); d_y: n x 4 */
int32_t gsap_mlp_forward(gsap_mlp* m, void* stream, const float* d_x, uint32_t n, float* d_y);
/* d_t: n x 65 targets per scene -- the regret tests_c / min_c tests - 1 of each
 * of the 64 candidate axes, then log(1 + pairs).  Loss = mean over rows of
 * sum_c softmax_c(30 (u.a_c)^2) regret_c + lambda (y3 - log(1+pairs))^2 with
 * u = y[0:3]/|y[0:3]| (expected regret of the axes the prediction points at).
 * Gradient only. */
int32_t gsap_mlp_loss_grad(gsap_mlp* m, void* stream, const float* d_x, const float* d_t,
                           uint32_t n, float lambda_count, float* h_loss, float* h_grads);
/* Adam (0.9, 0.999, 1e-8) over contiguous minibatches, `epochs` passes. */
int32_t gsap_mlp_train(gsap_mlp* m, void* stream, const float* d_x, const float* d_t, uint32_t n,
                       uint32_t epochs, uint32_t batch, float lr, float lambda_count,
                       float* h_epoch_loss);
/* The 64 fixed candidate axes (64 x 3 floats): coordinate axes first. */
int32_t gsap_axis_candidates(float* h_axes);
/* Attach a trained model to a handle (weights are copied).  Plain sweep: the
 * model names the sweep axis, and conf_min in (0,1] is the |cos| between
 * prediction and snapped candidate required to use it (else PCA).  Slab sweep
 * (GSAP_FLAG_SLABS): the model names the normal of the sweep plane, snapped to
 * the nearest coordinate axis; the sweep and slab axes are the two other
 * coordinate axes (sweep along the one of larger variance); conf_min is not
 * used, only a degenerate prediction keeps the PCA pair. */
int32_t gsap_sap_set_learned_model(gsap_sap* h, const gsap_mlp* m, float conf_min);
/* Per-scene features (n_scenes x 64 fp32, device) of packed AABB scenes. */
int32_t gsap_sap_features_batched(gsap_sap* h, void* stream, const void* d_in,
                                  const uint32_t* h_scene_offsets, uint32_t n_scenes,
                                  float* d_features);
/* The judge: exact sweep-test count of every scene on every axis
 * (d_tests: n_scenes x n_axes) and each scene's pair count (d_pairs, may be
 * NULL).  Plain sweep with scenes of at most 2048 objects: in-LDS sort per
 * (scene, axis).  Otherwise (AABB, canonical output, no group filter) the
 * counts are the sweep's own, from one count-only batched update per axis,
 * scenes of any size; on a slab handle (GSAP_FLAG_SLABS) the axes are
 * sweep-plane normals: a coordinate normal sweeps the two other coordinate
 * axes (the larger variance first), any other normal the two directions of
 * largest spread within its plane. */
int32_t gsap_sap_axis_tests_batched(gsap_sap* h, void* stream, const void* d_in,
                                    const uint32_t* h_scene_offsets, uint32_t n_scenes,
                                    const float* h_axes, uint32_t n_axes, uint64_t* d_tests,
                                    uint64_t* d_pairs);

/* ------------------------------------------------------------------ radix sort */
typedef struct gsap_radix_sort gsap_radix_sort;

/* RadixSort(maxElements, keysOnly) (radixsort.h:121-145) */
int32_t gsap_radix_sort_create(uint32_t max_elements, int32_t keys_only, gsap_radix_sort** out);
int32_t gsap_radix_sort_destroy(gsap_radix_sort* s);
/* RadixSort::sort(uint* keys, uint* values, uint n, uint keyBits)
 * (radixsort.h:151-178): stable, in place; d_values may be NULL when keys_only. */
int32_t gsap_radix_sort_u32(gsap_radix_sort* s, void* stream, uint32_t* d_keys,
                            uint32_t* d_values, uint32_t n, uint32_t key_bits);
/* RadixSort::sort(float* keys, uint* values, uint n, uint keyBits, bool
 * negativeKeys) (radixsort.h:180-198): float keys ordered numerically. */
int32_t gsap_radix_sort_f32(gsap_radix_sort* s, void* stream, float* d_keys, uint32_t* d_values,
                            uint32_t n, uint32_t key_bits, int32_t negative_keys);

/* ------------------------------------------------------------------ rigid bodies */
typedef struct gsap_prb gsap_prb;

enum {
    GSAP_PRB_ABS_SPRING = 1u,     /* k, c absolute (walled cube scene), not mass/dt derived */
    GSAP_PRB_WALLS = 2u,          /* x and z walls                                       */
    GSAP_PRB_FLOOR_CLAMP = 4u,    /* reference y clamp in integrate                      */
    GSAP_PRB_CAP_VEL = 8u,        /* velocity cap (setMaxVelocity)                       */
    GSAP_PRB_SKIP_SAME_BODY = 16u,/* broadphase drops pairs inside one body             */
    GSAP_PRB_INERTIA = 32u        /* omega += Iw^-1 tau dt (reference: tau dt)          */
};

typedef struct gsap_prb_physics {
    float gravity[3];   /* reference: (0,-5,0)                                     */
    float k_abs, c_abs; /* GSAP_PRB_ABS_SPRING constants                           */
    float floor_scale;  /* reference 0.1                                           */
    float wall_lo[2];   /* x, z                                                    */
    float wall_hi[2];
    uint32_t flags;     /* GSAP_PRB_*                                              */
    int32_t broadphase_kind; /* GSAP_INPUT_SPHERE (reference) or GSAP_INPUT_AABB  */
} gsap_prb_physics;

typedef struct gsap_prb_ptrs {
    float* rb_pos;       /* float4 per body                                 */
    float* rb_quat;      /* float4 per body (x,y,z,w)                       */
    float* rb_linvel;
    float* rb_angvel;
    float* rb_inv_inertia; /* 12 floats per body (3 rows of float4)       */
    float* p_pos;        /* float4 per particle, w = radius                 */
    float* p_vel;
    float* p_force;
    float* p_mass;
    int32_t* p_body;     /* body of each particle                           */
    int32_t* pairs;      /* int32x2 symmetric sorted pair list              */
    int32_t* pair_entries;
    int32_t num_bodies, num_particles;
    int64_t num_pairs, num_pair_entries;
    int32_t* rb_shape_idx;     /* shape of each body (getRbShapeIdx)        */
    int32_t* rb_num_particles; /* particles of each body (getRbNumParticles) */
} gsap_prb_ptrs;

/* ParticleRigidBody(maxNumBodies, maxNumParticles, maxPairs) */
int32_t gsap_prb_create(int32_t max_bodies, int32_t max_particles, int64_t max_pairs,
                        gsap_prb** out);
int32_t gsap_prb_destroy(gsap_prb* h);
/* setRbData (ParticleRigidBody.cpp:78-110): host arrays, normalises quats. */
int32_t gsap_prb_set_rb_data(gsap_prb* h, int32_t num_bodies, const float* h_pos,
                             const float* h_quat, const float* h_linvel, const float* h_angvel,
                             const int32_t* h_shape, const int32_t* h_pstart,
                             const int32_t* h_npart, const float* h_inv_mass,
                             const float* h_inv_inertia, int32_t total_particles);
/* setParticleData: per-shape particle templates, max_per_shape slots/shape. */
int32_t gsap_prb_set_particle_data(gsap_prb* h, const float* h_ref_pos, const float* h_ref_mass,
                                   int32_t num_elements, int32_t max_per_shape);
int32_t gsap_prb_set_shape_inv_inertia(gsap_prb* h, const float* h_inv_inertia, int32_t n_shapes);
int32_t gsap_prb_set_params(gsap_prb* h, float sp_multip, float dp_multip);
int32_t gsap_prb_set_max_velocity(gsap_prb* h, float min_radius, float dt);
int32_t gsap_prb_set_physics(gsap_prb* h, const gsap_prb_physics* phys);
int32_t gsap_prb_update_particle_data(gsap_prb* h, void* stream);
int32_t gsap_prb_prepare(gsap_prb* h, void* stream);
/* comupteFoceOnParticles (sic, ParticleRigidBody.cpp:148-167) */
int32_t gsap_prb_compute_force(gsap_prb* h, void* stream, float dt);
int32_t gsap_prb_update_velocity(gsap_prb* h, void* stream, float dt);
int32_t gsap_prb_integrate(gsap_prb* h, void* stream, float dt);
/* Broadphase on the particles through a SaP handle, result kept in the
 * handle's pair buffers (the reference calls Sap::collision + setNumPairs). */
int32_t gsap_prb_collide(gsap_prb* h, void* stream, gsap_sap* sap);
/* setNumPairs / setNumPairEntries after an external broadphase. */
int32_t gsap_prb_set_num_pairs(gsap_prb* h, int64_t num_pairs, int64_t num_pair_entries);
/* One DemoParticles::step (demoParticles.h:856-928): integrate, update
 * particles, broadphase, prepare, forces, update velocity. */
int32_t gsap_prb_step(gsap_prb* h, void* stream, gsap_sap* sap, float dt);
int32_t gsap_prb_get_ptrs(gsap_prb* h, gsap_prb_ptrs* out);

#ifdef __cplusplus
}
#endif
#endif /* GSAP_H */
