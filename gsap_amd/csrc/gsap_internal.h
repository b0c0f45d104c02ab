// gsap_internal.h -- device-side data layout and host launch wrappers shared by
// the HIP translation units of libgsap.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsap {

constexpr int kWG = 256;          // threads per workgroup for the object stages
constexpr int kWaves = kWG / 64;  // wave64
constexpr int kSweepTile = 256;   // sorted objects per sweep workgroup
constexpr int kMaxPasses = 8;
// Per-tile partials of k_moments: [0..8] sum c, sum c c^T; [9..11] sum of the
// extents per coordinate; [12..14] largest extent per coordinate.
constexpr int kMomStride = 16;

// Scene (segment) table for the batched path.  A "tile" is a run of at most
// `tile` consecutive objects of one scene; every object stage maps one tile
// (or a 256-object sub-tile) to a workgroup.  tile_seg == nullptr means a
// single scene [0, n).
struct SegTable {
    const uint32_t* seg_start;  // [nseg] first object
    const uint32_t* seg_n;      // [nseg] objects
    const uint32_t* seg_tile0;  // [nseg+1] first tile of each scene
    const uint32_t* tile_seg;   // [ntiles] scene of each tile
    uint32_t nseg;
    uint32_t ntiles;
};

// Sorted-object record (32 bytes).  Two copies are written after the sort:
// the candidate stream in sorted order (aux unused) and, per 64 positions
// (one sweep wave), the objects re-ordered by sweep end (aux = sorted position), read
// by the sweep through the scalar cache.
struct alignas(16) RecAabb {
    float lo[3];
    float hi[3];
    uint32_t orig;  // index into the caller's input (global in batched mode)
    uint32_t aux;
};
struct alignas(16) RecSphere {
    float c[4];     // x y z r
    uint32_t orig;
    uint32_t aux;
    uint32_t pad[2];
};

// Device control block of one SaP handle.  Zeroed once at allocation; the
// first kernel of every update resets what the update accumulates.
constexpr int kMaxRegions = 8;  // pair regions (one allocation counter each)
struct Ctrl {
    unsigned long long num_pairs;    // pairs (or directed pairs) found: the sum of the region
                                     // fills (set by k_pair_regions), or of part[0] (count only)
    unsigned long long tests;        // sweep-test count
    unsigned long long num_entries;  // SYM_CSR row starts
    unsigned long long exec_unused;  // (layout)
    // sweep workgroup totals spread over kMaxRegions slots (workgroup w adds
    // to slot w mod kMaxRegions: one contended address per XCD, not one for
    // the grid): [0] pairs of a count-only sweep, [1] candidate tests executed
    unsigned long long part[2][kMaxRegions];
    unsigned long long region_fill[kMaxRegions];       // pairs written to each region
    unsigned long long region_start[kMaxRegions + 1];  // logical start of each region
    unsigned long long region_valid[kMaxRegions];      // logical end of each region's pairs
    unsigned long long sort_n;       // pairs (+ tile padding) the pair sort reads
    unsigned long long sort_real;    // pairs the pair sort holds (padding excluded)
    uint32_t flags;                  // 1 = capacity overflow, 2 = non-finite input
    uint32_t epoch;                  // bumped per update; lookback tags
    uint32_t tile_ctr[2 * kMaxPasses];
    uint32_t local_big;              // windows the pair sort's LDS stage handed on
    uint32_t coh_fallback;           // coherent object sort: 1 = the regular projection and
                                     // LSD passes run instead
    uint32_t coh_far;                // coherent object sort: side-list entries claimed
    uint32_t pad;
};

enum : uint32_t { kFlagCapacity = 1u, kFlagNonFinite = 2u, kFlagTimeout = 4u, kFlagSlab = 8u };

// ---- launch wrappers (k_sap.hip) ----
struct ObjPipeArgs {
    const void* in;         // caller input
    int kind;               // 0 aabb, 1 sphere
    uint32_t n;             // total objects
    SegTable seg;           // scene table (single scene when tile_seg == nullptr)
    uint32_t tile;          // objects per tile (= sort tile)
    uint32_t ntiles;
    int axis_mode;          // GSAP_AXIS_*
    float fixed_axis[3];
    float* axis;            // [nseg][4] device: axis used (w = source)
    double* partials;       // [ntiles][kMomStride] moment and extent partials
    Ctrl* ctrl;
    uint32_t* hist_obj;     // [nseg][npass][256]
    uint32_t npass_obj;
    uint32_t bits_obj;      // object-sort digit width (6..8)
    uint32_t* hist_pair;    // [npass_pair][256]
    uint32_t npass_pair;
    uint32_t* keys;         // [n] keys (ping)
    uint32_t* vals;         // [n] orig index (ping)
    uint32_t* keys2;
    uint32_t* vals2;
    void* rec;              // [n] sorted records (candidate stream)
    void* grp;              // [n] records grouped by sweep end per 64 positions (waves
                            // with a window longer than kShortWindow only)
    uint32_t* grp_end;      // [n] sweep end - 1 per grouped slot; top bit: grp written
    const float* learned_axis;  // [nseg][4] learned axis proposal (w = confidence) or null
    double* seg_mom;        // [nseg][9] per-scene fp64 moments (written by k_axis)
    // if set: the sweep tests of scene s are added to seg_tests[s * seg_tests_stride]
    // (k_gather; the labels of the learned selector in slab mode)
    unsigned long long* seg_tests;
    uint32_t seg_tests_stride;
    // per-scene histograms of the box centres on the three coordinate planes
    // and axes, [nseg][kDensWords] (k_features; density features)
    uint32_t* plane_hist;
    // slab mode (GSAP_FLAG_SLABS, one scene): objects i -> entries 2i, 2i+1
    int slabs;
    float* axis2;           // [4] slab axis v (largest spread orthogonal to the sweep axis)
    float* vr_partials;     // [ntiles][kVr] per-tile ranges for the slab parameters
    float* slab;            // [kSlabParams] slab and key-quantisation parameters
    uint32_t* slab_cnt;     // [256] entries per slab (zeroed by k_moments, counted by k_project)
    // [3][kMaxRegions][256] or null: the pair sort's first-pass digit counts
    // per pair region (counted by the sweep), that pass's slot counters, and
    // the second-pass digit counts per part (counted by the sweep); zeroed
    // by k_moments
    uint32_t* pair_claim;
    // k_project runs only if *run_if != 0 (null: always); no_slab_count: the
    // entries per slab were counted already (the coherent projection)
    const uint32_t* run_if;
    int no_slab_count;
    // keep the key function of the last refresh (coherent updates): k_axis is
    // not launched and k_slab_params keeps the slab parameters while the slab
    // width still exceeds every box's slab-axis extent
    int keep_key;
};
constexpr uint32_t kPairClaimWords = 3u * kMaxRegions * 256u;
constexpr uint32_t kPairDigitStartWords = 260u;  // digit starts of the 32-bit items (257 used)
constexpr int kVr = 5;           // max v extent, min lo_v, max hi_v, min lo_u, max hi_u
constexpr int kSlabParams = 8;   // v0, 1/W, S, u0, 1/du, W, -, -
// Slab-mode key: slab << kSlabQBits | quantised lower bound on the sweep axis
// (any non-decreasing quantisation keeps the sweep exact; 16 bits over the
// sweep-axis range only adds ties).  24-bit keys: three 8-bit sort passes.
constexpr uint32_t kSlabQBits = 16u;
constexpr uint32_t kSlabKeyBits = 8u + kSlabQBits;
constexpr float kSlabQMax = (float)((1u << kSlabQBits) - 1u);
// key of an unused second entry: sorts after every slab (slab indices <= 254)
constexpr uint32_t kSlabNull = (1u << kSlabKeyBits) - 1u;
constexpr uint32_t kFirstBit = 0x80000000u;  // entry is its box's first slab

hipError_t launch_moments(const ObjPipeArgs& a, hipStream_t st);
// slab mode: per-tile projection ranges and the slab parameters (one scene)
hipError_t launch_slab_params(const ObjPipeArgs& a, hipStream_t st);
hipError_t launch_project(const ObjPipeArgs& a, hipStream_t st);
hipError_t launch_gather(const ObjPipeArgs& a, const uint32_t* sorted_keys,
                         const uint32_t* perm, hipStream_t st);

struct SweepArgs {
    const void* rec;         // candidates, sorted order
    const void* grp;         // objects, end-grouped per 64 positions (one wave)
    const uint32_t* grp_end;
    uint32_t n;
    SegTable seg;
    uint32_t tile;
    uint32_t ntiles;
    const uint32_t* group;   // per original index, or null
    uint2* out;              // pair regions (each dense, unsorted)
    unsigned long long slot_cap;   // pairs per region
    unsigned long long region_stride;  // slots between region starts
    uint32_t nregions;       // 1 or kMaxRegions (workgroup w uses region w mod nregions)
    Ctrl* ctrl;
    int kind;
    int sym;                 // emit both orientations
    int count_only;
    int slabs;               // entries carry kFirstBit in orig and their slab's first position in aux
    uint32_t xcd_chunk;      // sub-tiles per XCD chunk (0: hardware round-robin)
    // if set: 8-bit digit histograms of the written pairs' keys (a << bbits) | b
    // at bits hist_shift0 + 8 q, q < hist_npass <= kSweepHistPasses
    uint32_t* hist;
    int hist_npass;
    int hist_shift0;
    int bbits;
    uint32_t* hist_reg;      // if set: the first digit's counts per region, [kMaxRegions][256]
    // if set (with hist, two digits): the second digit's counts per part
    // (part = first digit >> 5), [kMaxRegions][256], in place of its histogram
    uint32_t* joint;
    // if set (count-only sweep): the pairs of scene s are added to seg_pairs[s]
    unsigned long long* seg_pairs;
};
hipError_t launch_sweep(const SweepArgs& a, hipStream_t st);
// Slab mode: scenes whose normals[4 s + 3] > 0 sweep the plane normal to
// normals[4 s .. 4 s + 2]: a coordinate normal gives the two other coordinate
// axes (the sweep axis the one of larger variance), any other normal the two
// directions of largest spread within the plane.  Writes axis (source 3) and
// axis2 of those scenes; the others keep theirs.  axis2 == null (plain
// sweep): the vector itself becomes the scene's sweep axis.
hipError_t launch_plane_pairs(const double* seg_mom, const uint32_t* seg_n, uint32_t n, uint32_t nseg,
                              const float* normals, float* axis, float* axis2, hipStream_t st);

// Digit histograms of the pair keys for npass digits starting at bit shift0;
// padding slots past region_valid[r] are not counted.
hipError_t launch_pair_hist(const uint2* pairs, const unsigned long long* n_dev,
                            unsigned long long cap, uint32_t* hist, int npass, int bbits, int bits,
                            int shift0, const unsigned long long* region_start,
                            const unsigned long long* region_valid,
                            unsigned long long region_stride, uint32_t nregions, hipStream_t st);
// Lay the sweep's pair regions out for the sort: clamp the fills so that
// together they hold at most region_cap pairs, round each region up to a whole
// tile (the padding is never read), and write the logical region starts and
// ends, the padded total (ctrl->sort_n) and the pair total (ctrl->sort_real).
hipError_t launch_pair_regions(Ctrl* ctrl, unsigned long long region_cap, uint32_t nregions,
                               uint32_t tile, hipStream_t st);
hipError_t launch_entries(const uint2* sorted_pairs, const unsigned long long* n_dev,
                          unsigned long long cap, uint32_t nobj, uint32_t* rowstart,
                          uint32_t* has, uint32_t* block_sums, int32_t* entries, Ctrl* ctrl,
                          hipStream_t st);
hipError_t launch_scene_offsets(const uint2* P, Ctrl* ctrl, unsigned long long cap,
                                const uint32_t* seg_start, uint32_t nseg, unsigned long long* out,
                                hipStream_t st);
hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t st);
// ---- coherent object sort (k_coherent.hip) ----
// Successive updates of one object set reuse the previous update's sorted
// order: each block of kCohB entries of it keeps the entries whose new
// composite (key << 32 | entry position) stays in the block's previous range
// and puts them in order in LDS; the others go to a side list that is sorted
// and merged in -- the order of the stable LSD sort, so the result is
// identical.  A side list of more than kCohFar entries, or an out-of-order
// run within a block reaching more than kCohRun entries to either side of one
// of its entries, sets ctrl->coh_fallback.  Kept entries out of order with
// the one kCohDist places away go to the side list as well.
constexpr uint32_t kCohB = 2048u;      // entries per block of the previous order
constexpr uint32_t kCohMaxB = 8192u;   // blocks (entries <= kCohB * kCohMaxB)
constexpr uint32_t kCohFar = 4096u;    // side list: entries that left their block's range
constexpr uint32_t kCohDist = 64u;    // an entry out of order with the one this far away: side list
constexpr uint32_t kCohRun = 128u;
constexpr uint32_t kCohRefresh = 64u;  // updates between refreshes of the key function
// small buffer: split[kCohMaxB] u64, far[kCohFar] u64, kept[kCohMaxB] u32,
// fstart[kCohMaxB + 1] u32, pstart[kCohMaxB] u32
constexpr size_t kCohSmallBytes =
    (size_t)kCohMaxB * 8 + (size_t)kCohFar * 8 + (size_t)kCohMaxB * 4 * 3 + 4;
struct CohArgs {
    const void* in;
    int kind;
    uint32_t ne;                 // entries (slab mode 2n, else n)
    uint32_t nb;                 // blocks, ceil(ne / kCohB)
    int slabs;
    const float* axis;           // sweep axis u
    const float* axis2;          // slab axis v
    const float* slab;           // slab parameters
    unsigned long long* split;   // [nb] split[k] (k >= 1): the previous update's composite
                                 // at position k kCohB
    unsigned long long* far;     // [kCohFar] side list
    uint32_t* kept;              // [nb] entries each block keeps
    uint32_t* fstart;            // [nb + 1] each block's first side-list entry
    uint32_t* pstart;            // [nb] each block's first final position
    uint32_t* keysX;             // kept entries, in order within each block (at its positions)
    uint32_t* valsX;
    uint32_t* keysF;             // the previous sorted order in, this update's out (the LSD
    uint32_t* valsF;             // path's final buffers)
    uint32_t* slab_cnt;
    Ctrl* ctrl;                  // coh_far and coh_fallback: zeroed by k_moments
};
// the blocks (new keys in the previous order, the kept entries in order, the
// side list and the entries per slab)
hipError_t launch_coh_block(const CohArgs& a, hipStream_t st);
// the side list sorted, then every block's entries at their final positions
hipError_t launch_coh_merge(const CohArgs& a, hipStream_t st);
// the splits of the next update from the sorted keys and values (keysF, valsF)
hipError_t launch_coh_save(const CohArgs& a, hipStream_t st);
// k_scratch.hip: device address (this GPU's copy) and size of the static
// scratch of the allocation-free entry point
hipError_t scratch_region(void** ptr, size_t* bytes);
// d_counts[0..4] = pairs, entries, sweep tests, flags, executed tests; with
// axis_out: axis_out[0..1] = the float bits of axis[0..3]
hipError_t launch_finish(Ctrl* ctrl, unsigned long long cap, unsigned long long* d_counts, int sym,
                         const float* axis, unsigned long long* axis_out, hipStream_t st);

// ---- learned sweep axis (k_mlp.hip) ----
constexpr int kFeat = 64, kHid = 128, kOut = 4, kCand = 64;
constexpr int kTgt = kCand + 1;  // training target row: 64 candidate regrets, log(1 + pairs)
constexpr int kFeatSums = 39;  // per-tile partial sums feeding the features
// density features: histograms of the box centres over mean +- 2 standard
// deviations per axis -- a kPlaneGrid^2 grid on each coordinate plane, then
// kAxisBins bins on each coordinate axis; kDensWords counters per scene
constexpr int kPlaneGrid = 8, kPlaneBins = kPlaneGrid * kPlaneGrid;
constexpr int kAxisBins = 32;
constexpr int kDensWords = 3 * kPlaneBins + 3 * kAxisBins;
struct MlpWeights {          // bf16 copies used by the kernels (row-major [in][out])
    const uint16_t *w1, *w2, *w3;
    const float *b1, *b2, *b3;
};
hipError_t launch_features(const ObjPipeArgs& a, float* tile_sums, float* feats, hipStream_t st);
// normals == null: confident scenes get the chosen candidate as their sweep
// axis in axis; else every scene's normals[4 s ..] is written (w = 1 when
// confident, else 0) for launch_plane_pairs
hipError_t launch_axis_infer(const float* feats, uint32_t nseg, MlpWeights w, const float* cand,
                             float conf_min, float* axis, float* normals, hipStream_t st);
hipError_t launch_axis_tests(const void* in, int kind, const uint32_t* seg_start,
                             const uint32_t* seg_n, uint32_t nseg, const float* axes, uint32_t naxes,
                             unsigned long long* tests, unsigned long long* pairs, hipStream_t st);
hipError_t mlp_forward(const float* x, uint32_t B, MlpWeights w, uint16_t* xb, uint16_t* h1,
                       uint16_t* h2, float* y, hipStream_t st);
hipError_t mlp_backward(const float* y, const float* target, uint32_t B, MlpWeights w,
                        const uint16_t* xb, const uint16_t* h1, const uint16_t* h2, float* dh1,
                        float* dh2, float* dy, float* grads, float lambda_count, float kappa,
                        const float* cand, float* loss, hipStream_t st);
hipError_t mlp_adam(float* master, const float* grads, float* m, float* v, uint32_t nparam,
                    float lr, float b1, float b2, float eps, uint32_t step, uint16_t* w_bf16,
                    const uint32_t* bf16_map, hipStream_t st);

// ---- radix sort (k_sort.hip) ----
struct SortArgs {
    uint32_t n;                         // host element count (ignored if n_dev)
    const unsigned long long* n_dev;    // device element count (pairs) or null
    unsigned long long cap;             // clamp for n_dev
    SegTable seg;                       // segmented (tile table) or single
    uint32_t ntiles;                    // tiles of the segmented table (if seg.tile_seg)
    const uint32_t* hist;               // [nseg][npass][256]
    uint32_t npass;
    unsigned long long* lookback;       // [max_tiles][256]
    Ctrl* ctrl;
    int ctr_base;                       // 0 object sort, kMaxPasses pair sort
    int bits;                           // digit width
    int items;                          // keys per thread
    int bbits;                          // pairs: bits of the second index
    uint32_t grid;                      // persistent grid size
    const unsigned long long* last_n;   // final pass writes only positions < min(*last_n,
    unsigned long long last_cap;        //   last_cap) (pairs: the caller's capacity)
    hipEvent_t* pass_ev;                // if set: recorded after each pass (timing)
    // pair sort, first pass: the input is nregions tile-aligned regions of
    // region_stride slots, logically concatenated at region_start[]; the
    // slots from region_valid[r] to the next region's start are padding
    const unsigned long long* region_start;
    const unsigned long long* region_valid;
    unsigned long long region_stride;
    uint32_t nregions;
    const unsigned long long* n_dev_rest;  // element count of the passes after the first (or null)
    int shift0;                            // bit offset of the first digit
    // pair sort, first pass without look-back (an LSD sort's first pass need
    // not be stable): per-region digit counts [kMaxRegions][256] and slot
    // counters of the same shape (zero), or null for a onesweep first pass
    const uint32_t* hist_reg;
    uint32_t* claim;
    // pair sort of two passes after a per-region first pass (keys of 34 to
    // 40 bits): [kMaxRegions][256] counts of the second digit per part
    // (part = first digit >> 5), taken by the sweep; the second pass then
    // runs per part, on xgrid workgroups, bucket by bucket of equal (second
    // digit, first digit).  Both passes write 32-bit items: the first the
    // key without its digit, the second the key's low 32 bits (the bits
    // above are its digit's).  The second pass leaves the digit starts in
    // st_out[0..256] and the bucket starts in bst[kBucketWords] ([d2][d1],
    // then the key count); xcnt (bucket sizes) and xcur (slot counters) are
    // kRadixSq words each
    uint32_t* joint;
    uint32_t xgrid;
    uint32_t* st_out;
    uint32_t* bst;
    uint32_t* xcnt;
    uint32_t* xcur;
    // every pass returns at once unless *run_if != 0 (null: always runs)
    const uint32_t* run_if;
    // key + value sorts with a host count n: reduce-then-scan passes (count,
    // scan, scatter; the look-back buffer holds the tile tables) instead of
    // decoupled look-back
    int rts;
};
// Reduce-then-scan sort passes from this many tiles (the measured crossover
// with the look-back passes on one MI355X: below it the chain is short and
// the two extra launches per pass cost more) and, for pair sorts, up to
// kRtsMaxPairTiles tiles of the capacity (past it a pass is bound by its
// bytes, and the extra read of the keys costs more than the chain: 1.14 vs
// 1.03 ms per pass over 150 M pairs).
constexpr uint32_t kRtsMinTiles = 128;
constexpr uint64_t kRtsMaxPairTiles = 4096;
// u32 key + u32 value; ping-pong between (k0,v0) and (k1,v1); returns in *which
// the buffer index (0/1) that holds the result.
hipError_t sort_kv32(const SortArgs& a, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                     int* which, hipStream_t st);
// Pairs: passes ping-pong between p0 and p1; the final pass writes `out`
// (or, with out == nullptr, the ping-pong buffer: p1 after an odd number of
// passes, p0 after an even number).
hipError_t sort_pairs(const SortArgs& a, uint2* p0, uint2* p1, uint2* out, hipStream_t st);

// ---- local stage of the pair sort (k_pair_local.hip) ----
// After the onesweep passes over key bits [L, 2 bbits), sort every run of
// equal key >> L ("bucket") by the low L bits (8 <= L <= 24).  x holds the
// bucket-sorted pairs (*n_dev of them), y is free scratch of the same size;
// out receives positions < min(*n_dev, *last_n, last_cap).
constexpr uint32_t kLocW = 3072;   // window: pairs whose bucket starts are one workgroup's
constexpr uint32_t kLocC = 8192;   // pairs sorted in LDS by one workgroup
uint64_t pair_local_wstart_words(uint64_t cap);  // u32 words of the window-start table
// big_ctr: a device word the launch zeroes and counts the handed-on windows in.
// st non-null: x holds 32-bit items (each key's low 32 bits; keys of at
// most 40 bits after a two-pass sort whose second pass ran per part, the
// bits above implied by the position: that pass's digit d occupies
// [st[d], st[d + 1])), and x has room for 2 cap pairs (the upper half is
// scratch for oversized windows).
// bst (with st): the second pass's bucket starts (SortArgs::bst), which
// replace the search for the window starts
hipError_t launch_pair_local(uint2* x, uint2* y, uint2* out, const unsigned long long* n_dev,
                             const unsigned long long* last_n, unsigned long long last_cap,
                             unsigned long long cap, uint32_t* wstart, uint32_t* big_ctr, int bbits,
                             int L, const uint32_t* st, const uint32_t* bst, hipStream_t stream);
// The big-window kernel alone, over a given list: window k of big[0, *big_ctr)
// is positions [wst[big[k]], wst[big[k] + 1]) of x, sorted by its low L bits
// into out (y: scratch at the same positions).
hipError_t launch_local_big_windows(uint2* x, uint2* y, uint2* out, const unsigned long long* n_dev,
                                    const unsigned long long* last_n, unsigned long long last_cap,
                                    const uint32_t* wst, const uint32_t* big, const uint32_t* big_ctr,
                                    int bbits, int L, hipStream_t st);
constexpr uint32_t kRadixSq = 256u * 256u;
constexpr uint32_t kBucketWords = kRadixSq + 1u;  // bucket starts of a two-pass pair sort
// In the window-start table: the two-pass sort's bucket starts
// (kBucketWords), then its bucket sizes and slot counters (kRadixSq each).
uint32_t* pair_local_bucket_starts(uint32_t* wstart, uint64_t cap);

// ---- pair sort by rows (k_pair_rows.hip) ----
// Counting sort of the pairs in P's regions (ctrl's region table) on the
// first index into X (dense), then each row ordered by the second index into
// out (positions < min(sort_real, num_pairs, cap)); Y: scratch of the sorted
// size.  n objects, 8 <= bbits <= 24.  cnt: n zeroed words (left zeroed),
// rs: n words, bsum: 2 ceil(n / 4096) words, table: pair_rows_table_words(cap)
// words.  entries (non-null): the symmetric output's row starts and
// ctrl->num_entries.
uint64_t pair_rows_table_words(uint64_t cap);
hipError_t launch_pair_rows(const uint2* P, uint2* X, uint2* Y, uint2* out, Ctrl* ctrl,
                            uint32_t nreg, unsigned long long stride,
                            unsigned long long sort_cap, unsigned long long cap, uint32_t n,
                            int bbits, uint32_t* cnt, uint32_t* rs, uint32_t* bsum,
                            int32_t* entries, uint32_t* table, hipStream_t st);
constexpr uint32_t kSweepStage = 512;      // pairs staged in LDS per sweep wave
constexpr uint32_t kSweepStageSlab = 560;  // the same, slab sweep (fewer flushes)
// One-word stage entries (inputs of fewer than 2^26 entries): the object's
// slot in the wave (6 bits) above the candidate's index (26 bits).  The stage
// fills the LDS of six workgroups per CU, the residency the sweep's registers
// allow.
constexpr uint32_t kStage4Shift = 26u;
constexpr uint32_t kStage4Mask = (1u << kStage4Shift) - 1u;
constexpr uint32_t kSweepStage4 = 1280;
constexpr uint32_t kShortWindow = 64;      // sweep: lane-per-object path up to this window
constexpr uint32_t kGatherWin = 320;       // gather: sorted keys each wave holds in LDS
constexpr uint32_t kSweepHistPasses = 2;   // pair-sort digit histograms the sweep can count
// Histogram of u32 keys for a plain (non-fused) sort, all passes.
hipError_t launch_key_hist(const uint32_t* keys, uint32_t n, uint32_t* hist, int npass, int bits,
                           int flip_float, hipStream_t st);
hipError_t launch_flip(uint32_t* keys, uint32_t n, int unflip, hipStream_t st);
hipError_t launch_sort_prep(Ctrl* ctrl, uint32_t* hist, uint32_t nhist, int ctr_base,
                            hipStream_t st);
uint32_t sort_tile(int items);

// ---- rigid bodies (k_prb.hip) ----
struct PrbDev {
    float4* pos; float4* quat; float4* linvel; float4* angvel;
    int32_t* shape; int32_t* pstart; int32_t* npart; float* inv_mass;
    float4* inv_inertia;    // 3 float4 per body
    float4* ref_pos; float* ref_mass; float4* shape_inv_inertia;
    float4* ppos; float4* pvel; float4* pforce; float* pmass; int32_t* pbody; int32_t* pip;
    int32_t nb, np, max_per_shape;
};
struct PrbPhys {
    float gravity[3];
    float sp, dp, k_abs, c_abs, floor_scale;
    float wall_lo[2], wall_hi[2];
    float max_lin_vel;
    uint32_t flags;
};
hipError_t prb_integrate(const PrbDev& d, const PrbPhys& p, float dt, hipStream_t st);
// aabb non-null: also each particle's AABB (6 floats), as prb_particle_aabbs
hipError_t prb_update_particles(const PrbDev& d, hipStream_t st, float* aabb = nullptr);
hipError_t prb_prepare(const PrbDev& d, hipStream_t st);
hipError_t prb_forces(const PrbDev& d, const PrbPhys& p, const int2* sym, const int32_t* entries,
                      const unsigned long long* counts, unsigned long long max_pairs, float dt,
                      hipStream_t st);
hipError_t prb_update_velocity(const PrbDev& d, const PrbPhys& p, float dt, hipStream_t st);
hipError_t prb_particle_aabbs(const PrbDev& d, float* out, hipStream_t st);

}  // namespace gsap
