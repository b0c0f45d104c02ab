// gsap_api.cpp -- host side of the C ABI declared in include/gsap.h.
//
// Owns the device workspaces, sequences the stage kernels on the caller's
// stream, validates arguments and turns device status bits into status codes.
// The reference orchestration this replaces is Sap::collision
// (physicboxSap2.16_Fermi/sapDetection/Sap.cpp:211-335: >= 8 blocking
// device->host copies and per-call CUDPP plan allocations) and
// ParticleRigidBody (ParticleRigidBody/ParticleRigidBody.cpp:15-191).  Here an
// update is one stream-ordered chain of kernels with a single device->host
// copy of the counts at the end (none in the async / batched-graph paths).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gsap.h"
#include "gsap_internal.h"

using namespace gsap;

namespace {

thread_local std::string g_err;

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = false;  // allocated here (false: carved from a caller workspace)
};

void release(Buf& b) {
    if (b.p && b.owned) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    b.owned = false;
}

// Grows b to at least `bytes`; contents are not kept.  *gen is bumped when
// the buffer moves, so anything that captured the old address (a recorded
// graph) can tell it is stale.
hipError_t grow(Buf& b, size_t bytes, uint64_t* gen) {
    if (bytes <= b.bytes && b.p) return hipSuccess;
    release(b);
    ++*gen;
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return e;
    }
    b.bytes = want;
    b.owned = true;
    return hipSuccess;
}

uint32_t bits_for(uint64_t n) {
    uint32_t b = 1;
    while (b < 32 && (1ull << b) < n) ++b;
    return b;
}

}  // namespace

struct gsap_sap {
    gsap_sap_desc desc{};
    std::string err;
    // launch configuration (gsap_tune)
    int items = 8;        // object sort: keys per thread (tile = 256 * items)
    int obj_bits = 8;     // object sort digit width
    int pair_items = 16;  // pair sort
    int pair_bits = 8;    // pair sort digit width
    int pair_grid = 2048; // pair sort persistent workgroups
    int slabs = 0;        // slab sweep for single-scene updates
    int sweep_chunk = 0;  // sweep sub-tiles per XCD chunk
    int pair_regions = 1; // sweep output regions (1 or kMaxRegions)
    bool tune_set = false;  // a configuration was set or measured (gsap_sap_set_tune/autotune)
    int pair_local = 0;   // pair sort: onesweep over the high key bits + per-bucket LDS sort
    float tune_ms = 0.0f;
    int tune_cached = 0;
    uint32_t cap_em = 1;  // entries per object the sort buffers are sized for
    Buf keys[2], vals[2], rec, partials, axis, hist_obj, hist_pair, lb_obj, lb_pair;
    Buf ctrl, counts, ptmp[2], rowstart, has, block_sums, learned_axis, grp, grp_end;
    Buf seg_start, seg_n, seg_tile0, tile_seg;
    Buf seg_start_e, seg_n_e, seg_tile0_e, tile_seg_e;  // slab mode: the scene table of entries
    uint32_t ntiles_e = 0;                               // entry tiles of the uploaded table
    Buf feats, tile_sums, seg_mom, cand, mlp_wb, mlp_master, ax_tests_axes;
    Buf plane_hist;  // per-scene centre histograms on the coordinate planes (features)
    Buf normals;     // slab mode: per-scene sweep-plane normals (learned or forced)
    // Labels of the learned selector in slab mode (gsap_sap_axis_tests_batched):
    // every scene sweeps the plane normal to normals[s] (force_normal), and
    // the update adds each scene's tests to lab_tests[s * lab_stride] and, when
    // counting, its pairs to lab_pairs[s].
    bool force_normal = false;
    unsigned long long* lab_tests = nullptr;
    uint32_t lab_stride = 0;
    unsigned long long* lab_pairs = nullptr;
    Buf axis2, vr_partials, slabp, slab_cnt;  // slab mode
    Buf wstart;                               // window starts of the pair sort's local stage
    Buf coh;                                  // coherent object sort: splitters and buckets
    bool has_model = false;
    float conf_min = 0.95f;
    unsigned long long* h_counts = nullptr;  // pinned
    uint32_t cap_n = 0, cap_tiles = 0, cap_seg = 0;
    uint64_t cap_pair_tiles = 0;
    uint64_t slot_cap = 0;       // pairs per sweep region (the caller's capacity)
    uint64_t region_stride = 0;  // slots between region starts
    uint64_t sort_cap = 0;       // capacity of the second pair buffer (padded total)
    bool timing = false;
    // coherent updates (GSAP_FLAG_COHERENT): the sorted order and splitters
    // in the buffers belong to an update of coh_n objects at buffer
    // generation coh_gen; coh_age counts updates since the key function was
    // refreshed
    bool coh_valid = false;
    uint32_t coh_n = 0, coh_age = 0;
    int coh_slabs = 0;
    uint64_t coh_gen = 0;
    uint32_t cus = 256;  // compute units of the device the handle was created on
    static constexpr int kEv = 20;
    hipEvent_t ev[kEv] = {};
    int nev = 0;
    float stage_ms[kEv] = {};
    // Buffer generation: bumped whenever a buffer an update uses moves (grown,
    // re-carved) or a scene table is rewritten.  The captured batched graph
    // holds raw buffer addresses and is replayed only at the generation it was
    // recorded at.
    uint64_t buf_gen = 1;
    // caller-provided workspace (gsap_sap_set_workspace): the update buffers
    // are carved from it instead of allocated; ws_size[] is the current carve
    void* ws_base = nullptr;
    uint64_t ws_bytes = 0;
    size_t ws_size[40] = {};
    bool ws_carved = false;
    // batched graph cache
    uint64_t g_gen = 0;
    hipGraphExec_t gexec = nullptr;
    hipStream_t cap_stream = nullptr;
    std::vector<uint32_t> g_offsets;
    const void* g_in = nullptr;
    void* g_pairs = nullptr;
    uint64_t g_cap = 0;
    unsigned long long* g_scene_out = nullptr;
};

static int32_t fail(gsap_sap* h, int32_t code, const std::string& msg) {
    if (h) h->err = msg;
    g_err = msg;
    return code;
}

#define HIPCHK(h, expr)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(h, e_ == hipErrorOutOfMemory ? GSAP_ERR_OOM : GSAP_ERR_HIP,          \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

extern "C" const char* gsap_last_error(void) { return g_err.c_str(); }

namespace gsap_err_bridge {
void set(const std::string& m) { g_err = m; }
}

extern "C" const char* gsap_version(void) { return "gsap-mi355x 0.1 (gfx950)"; }

extern "C" int32_t gsap_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int32_t gsap_sap_create(const gsap_sap_desc* desc, gsap_sap** out) {
    if (!desc || !out) return fail(nullptr, GSAP_ERR_INVALID_ARG, "gsap_sap_create: null argument");
    *out = nullptr;
    if (desc->input_kind != GSAP_INPUT_AABB && desc->input_kind != GSAP_INPUT_SPHERE)
        return fail(nullptr, GSAP_ERR_INVALID_ARG, "gsap_sap_create: bad input_kind");
    if (desc->axis_mode < GSAP_AXIS_PCA || desc->axis_mode > GSAP_AXIS_AUTO)
        return fail(nullptr, GSAP_ERR_INVALID_ARG, "gsap_sap_create: bad axis_mode");
    if (desc->output_kind != GSAP_OUT_CANON && desc->output_kind != GSAP_OUT_SYM_CSR)
        return fail(nullptr, GSAP_ERR_INVALID_ARG, "gsap_sap_create: bad output_kind");
    if (desc->digit_bits != 0 && (desc->digit_bits < 6 || desc->digit_bits > 8))
        return fail(nullptr, GSAP_ERR_UNSUPPORTED, "gsap_sap_create: digit_bits must be 0 or 6..8");
    if (desc->sort_items != 0 && desc->sort_items != 4 && desc->sort_items != 8 &&
        desc->sort_items != 12 && desc->sort_items != 16)
        return fail(nullptr, GSAP_ERR_INVALID_ARG, "gsap_sap_create: sort_items in {4,8,12,16}");
    if (desc->axis_mode == GSAP_AXIS_FIXED &&
        !(std::isfinite(desc->fixed_axis[0]) && std::isfinite(desc->fixed_axis[1]) &&
          std::isfinite(desc->fixed_axis[2])) )
        return fail(nullptr, GSAP_ERR_INVALID_ARG, "gsap_sap_create: non-finite fixed axis");
    if (gsap_device_count() <= 0) return fail(nullptr, GSAP_ERR_HIP, "gsap_sap_create: no HIP device");
    gsap_sap* h = new gsap_sap();
    h->desc = *desc;
    if (desc->sort_items) h->items = desc->sort_items;
    if (desc->digit_bits) h->obj_bits = h->pair_bits = desc->digit_bits;
    h->slabs = (desc->flags & GSAP_FLAG_SLABS) ? 1 : 0;
    hipError_t e = grow(h->ctrl, sizeof(Ctrl), &h->buf_gen);
    if (e == hipSuccess) e = hipMemset(h->ctrl.p, 0, sizeof(Ctrl));
    if (e == hipSuccess) e = grow(h->counts, 8 * sizeof(unsigned long long), &h->buf_gen);
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->h_counts, 8 * sizeof(unsigned long long), 0);
    // stage events only time the work (read after the stream is synchronised),
    // so they skip the system-scope fence (an L2 write-back on every record)
    for (int i = 0; i < gsap_sap::kEv && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&h->ev[i], hipEventDisableSystemFence);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        int dev = 0, cus = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess && cus > 0) h->cus = (uint32_t)cus;
    }
    if (e != hipSuccess) {
        std::string msg = std::string("gsap_sap_create: ") + hipGetErrorString(e);
        gsap_sap_destroy(h);
        return fail(nullptr, GSAP_ERR_HIP, msg);
    }
    *out = h;
    return GSAP_OK;
}

extern "C" int32_t gsap_sap_destroy(gsap_sap* h) {
    if (!h) return GSAP_OK;
    (void)hipDeviceSynchronize();
    Buf* bufs[] = {&h->keys[0], &h->keys[1], &h->vals[0], &h->vals[1], &h->rec, &h->partials,
                   &h->axis, &h->hist_obj, &h->hist_pair, &h->lb_obj, &h->lb_pair,
                   &h->ctrl, &h->counts, &h->ptmp[0], &h->ptmp[1], &h->rowstart, &h->has,
                   &h->block_sums, &h->learned_axis, &h->seg_start, &h->seg_n, &h->seg_tile0,
                   &h->tile_seg, &h->grp, &h->grp_end, &h->feats, &h->tile_sums, &h->seg_mom,
                   &h->cand, &h->mlp_wb, &h->mlp_master, &h->ax_tests_axes, &h->axis2,
                   &h->vr_partials, &h->slabp, &h->slab_cnt, &h->wstart, &h->plane_hist,
                   &h->normals};
    for (Buf* b : bufs) release(*b);
    if (h->h_counts) (void)hipHostFree(h->h_counts);
    for (int i = 0; i < gsap_sap::kEv; ++i)
        if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    delete h;
    return GSAP_OK;
}

static bool slab_mode(const gsap_sap* h) { return h->slabs != 0; }

// ---------------------------------------------------------------- workspace
// The device buffers of one update (besides the caller's input and output,
// the control block and the counts), in carve order.
enum WsId : int {
    W_KEYS0, W_KEYS1, W_VALS0, W_VALS1, W_REC, W_GRP, W_GRP_END, W_ROWSTART, W_HAS,
    W_BLOCK_SUMS, W_PARTIALS, W_LB_OBJ, W_VR_PARTIALS, W_TILE_SEG, W_TILE_SUMS, W_AXIS,
    W_HIST_OBJ, W_SEG_START, W_SEG_N, W_SEG_TILE0, W_SEG_MOM, W_FEATS, W_HIST_PAIR, W_AXIS2,
    W_SLABP, W_SLAB_CNT, W_PTMP0, W_PTMP1, W_WSTART, W_LB_PAIR, W_COH, W_SEG_START_E, W_SEG_N_E,
    W_SEG_TILE0_E, W_TILE_SEG_E, W_PLANE_HIST, W_NORMALS, W_COUNT
};
static_assert(W_COUNT <= 40, "gsap_sap::ws_size holds 40 entries");

static Buf* ws_buf(gsap_sap* h, int id) {
    switch (id) {
        case W_KEYS0: return &h->keys[0];
        case W_KEYS1: return &h->keys[1];
        case W_VALS0: return &h->vals[0];
        case W_VALS1: return &h->vals[1];
        case W_REC: return &h->rec;
        case W_GRP: return &h->grp;
        case W_GRP_END: return &h->grp_end;
        case W_ROWSTART: return &h->rowstart;
        case W_HAS: return &h->has;
        case W_BLOCK_SUMS: return &h->block_sums;
        case W_PARTIALS: return &h->partials;
        case W_LB_OBJ: return &h->lb_obj;
        case W_VR_PARTIALS: return &h->vr_partials;
        case W_TILE_SEG: return &h->tile_seg;
        case W_TILE_SUMS: return &h->tile_sums;
        case W_AXIS: return &h->axis;
        case W_HIST_OBJ: return &h->hist_obj;
        case W_SEG_START: return &h->seg_start;
        case W_SEG_N: return &h->seg_n;
        case W_SEG_TILE0: return &h->seg_tile0;
        case W_SEG_MOM: return &h->seg_mom;
        case W_FEATS: return &h->feats;
        case W_HIST_PAIR: return &h->hist_pair;
        case W_AXIS2: return &h->axis2;
        case W_SLABP: return &h->slabp;
        case W_SLAB_CNT: return &h->slab_cnt;
        case W_PTMP0: return &h->ptmp[0];
        case W_PTMP1: return &h->ptmp[1];
        case W_WSTART: return &h->wstart;
        case W_COH: return &h->coh;
        case W_SEG_START_E: return &h->seg_start_e;
        case W_SEG_N_E: return &h->seg_n_e;
        case W_SEG_TILE0_E: return &h->seg_tile0_e;
        case W_TILE_SEG_E: return &h->tile_seg_e;
        case W_PLANE_HIST: return &h->plane_hist;
        case W_NORMALS: return &h->normals;
        default: return &h->lb_pair;
    }
}

// Buffers that must start zeroed whenever their memory is (re)assigned:
// `has` (every row-start pass leaves it zeroed again) and the look-back
// arrays (their words are tagged with the handle's update epoch, so reused
// memory must not hold a tag an update could match).
static bool ws_zero_init(int id) { return id == W_HAS || id == W_LB_OBJ || id == W_LB_PAIR; }

static size_t ws_align(size_t b) { return (b + 255) & ~(size_t)255; }

struct WsReq {
    size_t b[W_COUNT] = {};
};

static size_t ws_total(const size_t* b) {
    size_t t = 0;
    for (int i = 0; i < W_COUNT; ++i) t += ws_align(b[i]);
    return t;
}

// Object side of an update of n objects in nseg scenes over ntiles sort
// tiles.  In slab mode the sort, gather and sweep buffers hold two entries
// per object.
static void req_obj(const gsap_sap* h, uint32_t n, uint32_t nseg, uint32_t ntiles, WsReq& r) {
    const size_t em = slab_mode(h) ? 2 : 1;
    const size_t want = std::max(n, h->desc.max_n);
    r.b[W_KEYS0] = r.b[W_KEYS1] = r.b[W_VALS0] = r.b[W_VALS1] = want * 4 * em;
    r.b[W_REC] = r.b[W_GRP] = want * 32 * em;
    r.b[W_GRP_END] = want * 4 * em;
    r.b[W_ROWSTART] = r.b[W_HAS] = want * 4;
    r.b[W_BLOCK_SUMS] = (want / 2048 + 2) * 4;
    r.b[W_PARTIALS] = (size_t)ntiles * kMomStride * sizeof(double);
    r.b[W_LB_OBJ] = ((size_t)ntiles * em + 1) * 256 * 8;
    r.b[W_VR_PARTIALS] = slab_mode(h) ? (size_t)ntiles * kVr * 4 : 0;
    r.b[W_TILE_SEG] = (size_t)ntiles * 4;
    r.b[W_TILE_SUMS] = (size_t)ntiles * kFeatSums * 4;
    r.b[W_AXIS] = (size_t)nseg * 16;
    r.b[W_HIST_OBJ] = (size_t)nseg * kMaxPasses * 256 * 4;
    r.b[W_SEG_START] = r.b[W_SEG_N] = (size_t)nseg * 4;
    r.b[W_SEG_TILE0] = (size_t)(nseg + 1) * 4;
    r.b[W_SEG_MOM] = (size_t)nseg * 9 * 8;
    r.b[W_FEATS] = (size_t)nseg * kFeat * 4;
    r.b[W_PLANE_HIST] = (size_t)nseg * kDensWords * 4;
    r.b[W_NORMALS] = (size_t)nseg * 4 * sizeof(float);
    // pair digit histograms, then the first pass's per-region counts and slot counters
    r.b[W_HIST_PAIR] = (kMaxPasses * 256 + kPairClaimWords + kPairDigitStartWords) * 4;
    if (slab_mode(h)) {  // per scene
        r.b[W_AXIS2] = (size_t)nseg * 4 * sizeof(float);
        r.b[W_SLABP] = (size_t)nseg * kSlabParams * sizeof(float);
        r.b[W_SLAB_CNT] = (size_t)nseg * 256 * sizeof(uint32_t);

        if (nseg > 1) {  // batched: the scene table of entries (2 per object)
            r.b[W_SEG_START_E] = r.b[W_SEG_N_E] = (size_t)nseg * 4;
            r.b[W_SEG_TILE0_E] = (size_t)(nseg + 1) * 4;
            r.b[W_TILE_SEG_E] = ((size_t)ntiles * 2 + nseg) * 4;
        }
    }
    if ((h->desc.flags & GSAP_FLAG_COHERENT) && nseg == 1) r.b[W_COH] = kCohSmallBytes;
}

// Internal pair buffers.  The sweep writes the pairs it finds densely into
// pair_regions regions of the first buffer, each able to hold the caller's
// capacity; the sort reads them as one array (regions padded to a whole tile)
// and ping-pongs with the second buffer, which holds the padded total.
static constexpr uint64_t kPairTileMax = 256 * 16;  // largest pair-sort tile
static int32_t req_pairs(gsap_sap* h, uint64_t cap, WsReq& r) {
    const uint64_t rcap = cap > 0 ? cap : 1;
    // pair positions are 32-bit; the margin keeps every sort index (padding,
    // unrolled histogram rounds) below 2^32
    if (rcap >= (1ull << 32) - (1ull << 22))
        return fail(h, GSAP_ERR_UNSUPPORTED, "pair capacity too large (pair indices are 32-bit)");
    const uint64_t nreg = (uint64_t)h->pair_regions;
    const uint64_t stride = (rcap + kPairTileMax - 1) / kPairTileMax * kPairTileMax + kPairTileMax;
    const uint64_t sorted = rcap + (nreg + 1) * kPairTileMax;  // padded total, upper bound
    const uint64_t tiles = (sorted + sort_tile(h->pair_items) - 1) / sort_tile(h->pair_items) + 1;
    r.b[W_PTMP0] = (size_t)(nreg * stride) * 8;
    r.b[W_PTMP1] = (size_t)sorted * 8;
    // window starts of the local stage, or the long-row list of the row sort
    r.b[W_WSTART] = (size_t)std::max(pair_local_wstart_words(rcap), pair_rows_table_words(rcap)) * 4;
    r.b[W_LB_PAIR] = (size_t)tiles * 256 * 8;
    h->slot_cap = rcap;
    h->region_stride = stride;
    h->sort_cap = sorted;
    return GSAP_OK;
}

// Zero the zero-initialised buffers of the current carve on st (a fill
// kernel, so a caller may record the call into a graph).
static int32_t ws_clear(gsap_sap* h, hipStream_t st) {
    for (int id = 0; id < W_COUNT; ++id) {
        Buf* b = ws_buf(h, id);
        if (ws_zero_init(id) && b->p && b->bytes)
            HIPCHK(h, launch_zero_u64((unsigned long long*)b->p, (uint32_t)((b->bytes + 7) / 8), st));
    }
    return GSAP_OK;
}

// Assigns memory to every buffer `base` points into, from sizes sz[], and
// clears the zero-initialised ones on st.
static int32_t ws_carve(gsap_sap* h, char* base, const size_t* sz, hipStream_t st) {
    size_t off = 0;
    for (int id = 0; id < W_COUNT; ++id) {
        Buf* b = ws_buf(h, id);
        release(*b);
        b->p = sz[id] ? base + off : nullptr;
        b->bytes = sz[id];
        h->ws_size[id] = sz[id];
        off += ws_align(sz[id]);
    }
    ++h->buf_gen;
    return ws_clear(h, st);
}

// Makes every buffer of req available.  Internal mode grows allocations
// (they never shrink); with a caller workspace the current carve is kept when
// it covers req, else the workspace is carved again for the union of the two
// (or for req alone when the union does not fit).
static int32_t ws_ensure(gsap_sap* h, const WsReq& req, hipStream_t st) {
    if (!h->ws_base) {
        for (int id = 0; id < W_COUNT; ++id) {
            Buf* b = ws_buf(h, id);
            if (req.b[id] == 0 || (b->p && req.b[id] <= b->bytes)) continue;
            HIPCHK(h, grow(*b, req.b[id], &h->buf_gen));
            if (ws_zero_init(id)) HIPCHK(h, hipMemset(b->p, 0, b->bytes));
        }
        return GSAP_OK;
    }
    bool covers = h->ws_carved;
    size_t merged[W_COUNT];
    for (int id = 0; id < W_COUNT; ++id) {
        covers = covers && req.b[id] <= h->ws_size[id];
        merged[id] = std::max(req.b[id], h->ws_carved ? h->ws_size[id] : (size_t)0);
    }
    if (covers) return GSAP_OK;
    const size_t* use = ws_total(merged) <= h->ws_bytes ? merged : req.b;
    const size_t need = ws_total(use);
    if (need > h->ws_bytes) {
        char buf[160];
        snprintf(buf, sizeof buf, "workspace of %llu bytes too small: %llu required",
                 (unsigned long long)h->ws_bytes, (unsigned long long)ws_total(req.b));
        return fail(h, GSAP_ERR_CAPACITY, buf);
    }
    h->ws_carved = true;
    return ws_carve(h, (char*)h->ws_base, use, st);
}

// Requirement of a single-scene update of n objects (pairs: cap > 0).
static int32_t req_single(gsap_sap* h, uint32_t n, uint64_t cap, bool count_only, WsReq& r) {
    const uint32_t tile = sort_tile(h->items);
    req_obj(h, n, 1, (n + tile - 1) / tile, r);
    if (!count_only) return req_pairs(h, cap, r);
    return GSAP_OK;
}

static void mark(gsap_sap* h, hipStream_t st) {
    if (h->timing && h->nev < gsap_sap::kEv) (void)hipEventRecord(h->ev[h->nev++], st);
}

// Enqueue one complete update (all stages) on st.  seg == nullptr: one scene.
static int32_t enqueue_update(gsap_sap* h, hipStream_t st, const void* in, uint32_t n,
                              const SegTable& seg, uint32_t ntiles, const uint32_t* group,
                              void* pairs, uint64_t cap, int32_t* entries, bool count_only,
                              unsigned long long* d_counts, unsigned long long* scene_out,
                              bool allow_coherent = false) {
    const int kind = h->desc.input_kind;
    const bool sym = h->desc.output_kind == GSAP_OUT_SYM_CSR;
    const uint32_t tile = sort_tile(h->items);
    const uint32_t bbits = bits_for(n);
    // Pair sort.  Whole-key onesweep: ceil(2 bbits / pair_bits) passes.  With
    // the local stage: 8-bit onesweep passes over key bits [L, 2 bbits) with
    // 8 <= L <= 24, then every run of equal key >> L sorted by the low L bits.
    const uint32_t kbits = 2 * bbits;
    const bool local = h->pair_local == 1 && kbits >= 16;
    // By rows: a counting sort on the first index, then each row by the
    // second (8 <= bbits <= 24; other sizes take the whole-key passes).
    const bool rows = !count_only && h->pair_local == 2 && bbits >= 8 && bbits <= 24;
    const uint32_t npass_g = !local ? 0 : (kbits > 24 ? (kbits - 24 + 7) / 8 : 1);
    const uint32_t shift0 = local ? kbits - 8 * npass_g : 0;
    const uint32_t pair_bits = local ? 8 : (uint32_t)h->pair_bits;
    const uint32_t npass_pair = local ? npass_g : (kbits + pair_bits - 1) / pair_bits;
    // object keys: 32-bit flipped floats, or 24-bit slab keys
    const uint32_t obj_key_bits = slab_mode(h) ? kSlabKeyBits : 32u;
    const uint32_t npass_obj = (obj_key_bits + h->obj_bits - 1) / h->obj_bits;
    Ctrl* ctrl = (Ctrl*)h->ctrl.p;
    h->nev = 0;
    mark(h, st);

    ObjPipeArgs a{};
    a.in = in;
    a.kind = kind;
    a.n = n;
    a.seg = seg;
    a.tile = tile;
    a.ntiles = ntiles;
    a.axis_mode = h->desc.axis_mode;
    for (int k = 0; k < 3; ++k) a.fixed_axis[k] = h->desc.fixed_axis[k];
    a.axis = (float*)h->axis.p;
    a.partials = (double*)h->partials.p;
    a.ctrl = ctrl;
    a.hist_obj = (uint32_t*)h->hist_obj.p;
    a.npass_obj = npass_obj;
    a.bits_obj = h->obj_bits;
    a.hist_pair = (uint32_t*)h->hist_pair.p;
    a.npass_pair = count_only ? 0 : npass_pair;
    a.pair_claim = a.hist_pair + kMaxPasses * 256;
    a.keys = (uint32_t*)h->keys[0].p;
    a.vals = (uint32_t*)h->vals[0].p;
    a.keys2 = (uint32_t*)h->keys[1].p;
    a.vals2 = (uint32_t*)h->vals[1].p;
    a.rec = h->rec.p;
    a.grp = h->grp.p;
    a.grp_end = (uint32_t*)h->grp_end.p;
    a.learned_axis = nullptr;
    a.seg_mom = (double*)h->seg_mom.p;
    a.plane_hist = (uint32_t*)h->plane_hist.p;
    a.seg_tests = h->lab_tests;
    a.seg_tests_stride = h->lab_stride;
    // slab mode: one scene only (the batched entry point refuses the flag)
    // slab mode: per scene slab axis, parameters and slab counts; the sort,
    // gather and sweep run over entries (2 per object), with the scene table
    // of entries when batched
    const bool slabs = slab_mode(h);
    a.slabs = slabs ? 1 : 0;
    a.axis2 = (float*)h->axis2.p;
    a.vr_partials = (float*)h->vr_partials.p;
    a.slab = (float*)h->slabp.p;
    a.slab_cnt = (uint32_t*)h->slab_cnt.p;
    const uint32_t n_e = slabs ? 2 * n : n;  // sorted entries
    SegTable seg_e = seg;
    uint32_t ntiles_e = ntiles;
    if (slabs && seg.tile_seg) {
        seg_e.seg_start = (const uint32_t*)h->seg_start_e.p;
        seg_e.seg_n = (const uint32_t*)h->seg_n_e.p;
        seg_e.seg_tile0 = (const uint32_t*)h->seg_tile0_e.p;
        seg_e.tile_seg = (const uint32_t*)h->tile_seg_e.p;
        seg_e.ntiles = h->ntiles_e;
        ntiles_e = h->ntiles_e;
    } else if (slabs) {
        ntiles_e = (n_e + tile - 1) / tile;
    }
    const bool learned = h->desc.axis_mode == GSAP_AXIS_LEARNED;
    // Coherent update (GSAP_FLAG_COHERENT): a single scene whose previous
    // update of the same object count left its sorted order and splitters in
    // these buffers.  Every kCohRefresh-th update refreshes the key function
    // (and sorts the regular way); the others keep it.
    const bool coherent = allow_coherent && (h->desc.flags & GSAP_FLAG_COHERENT) &&
                          seg.tile_seg == nullptr && !learned &&
                          (uint64_t)n_e <= (uint64_t)kCohB * kCohMaxB;
    const bool keep = coherent && h->coh_valid && h->coh_n == n && h->coh_slabs == (int)slabs &&
                      h->coh_gen == h->buf_gen && h->coh_age % kCohRefresh != 0;
    a.keep_key = keep ? 1 : 0;
    // the buffers the object sort ends in (the LSD passes alternate)
    uint32_t* const keysF = (npass_obj & 1u) ? a.keys2 : a.keys;
    uint32_t* const valsF = (npass_obj & 1u) ? a.vals2 : a.vals;
    CohArgs ca{};
    if (coherent) {
        char* cb = (char*)h->coh.p;
        ca.in = in;
        ca.kind = kind;
        ca.ne = n_e;
        ca.nb = (n_e + kCohB - 1) / kCohB;
        ca.slabs = slabs ? 1 : 0;
        ca.axis = a.axis;
        ca.axis2 = a.axis2;
        ca.slab = a.slab;
        ca.split = (unsigned long long*)cb;
        ca.far = ca.split + kCohMaxB;
        ca.kept = (uint32_t*)(ca.far + kCohFar);
        ca.fstart = ca.kept + kCohMaxB;
        ca.pstart = ca.fstart + kCohMaxB + 1;
        // kept entries go to the pair the LSD passes do not end in
        ca.keysX = (npass_obj & 1u) ? a.keys : a.keys2;
        ca.valsX = (npass_obj & 1u) ? a.vals : a.vals2;
        ca.keysF = keysF;
        ca.valsF = valsF;
        ca.slab_cnt = a.slab_cnt;
        ca.ctrl = ctrl;
    }
    if (learned) {
        if (!h->has_model)
            return fail(h, GSAP_ERR_UNSUPPORTED, "GSAP_AXIS_LEARNED needs gsap_sap_set_learned_model");
        a.axis_mode = GSAP_AXIS_PCA;  // the fallback; confident scenes are overwritten below
    }
    HIPCHK(h, launch_moments(a, st));
    if (learned) {
        HIPCHK(h, launch_features(a, (float*)h->tile_sums.p, (float*)h->feats.p, st));
        MlpWeights mw;
        mw.w1 = (const uint16_t*)h->mlp_wb.p;
        mw.w2 = mw.w1 + kFeat * kHid;
        mw.w3 = mw.w2 + kHid * kHid;
        const float* mb = (const float*)h->mlp_master.p;
        mw.b1 = mb + kFeat * kHid;
        mw.b2 = mb + kFeat * kHid + kHid + kHid * kHid;
        mw.b3 = mb + kFeat * kHid + kHid + kHid * kHid + kHid + kHid * kOut;
        // slab mode: the network names the sweep plane's normal; its axis pair
        // replaces the PCA pair of the confident scenes
        float* const nrm = slabs ? (float*)h->normals.p : nullptr;
        HIPCHK(h, launch_axis_infer((const float*)h->feats.p, seg.nseg, mw, (const float*)h->cand.p,
                                    h->conf_min, a.axis, nrm, st));
        if (slabs)
            HIPCHK(h, launch_plane_pairs(a.seg_mom, seg.seg_n, n, seg.nseg, nrm, a.axis, a.axis2, st));
    } else if (h->force_normal) {
        HIPCHK(h, launch_plane_pairs(a.seg_mom, seg.seg_n, n, seg.tile_seg ? seg.nseg : 1u,
                                     (const float*)h->normals.p, a.axis, slabs ? a.axis2 : nullptr, st));
    }
    if (slabs) HIPCHK(h, launch_slab_params(a, st));
    if (keep) {
        HIPCHK(h, launch_coh_block(ca, st));
    } else {
        HIPCHK(h, launch_project(a, st));
    }
    mark(h, st);

    SortArgs s{};
    s.n = n_e;
    s.seg = seg_e;
    s.ntiles = ntiles_e;
    s.hist = a.hist_obj;
    s.npass = npass_obj;
    s.lookback = (unsigned long long*)h->lb_obj.p;
    // at least kRtsMinTiles sort tiles: reduce-then-scan passes (no
    // look-back chain; below that the chain is short and the two extra
    // launches per pass cost more).  The coherent fallback keeps the
    // look-back passes, which it launches one workgroup per CU so that they
    // cost little when they return at once.
    s.rts = !keep && (seg_e.tile_seg ? ntiles_e : (n_e + tile - 1) / tile) >= kRtsMinTiles;
    s.ctrl = ctrl;
    s.ctr_base = 0;
    s.bits = h->obj_bits;
    s.items = h->items;
    s.grid = 0;
    int which = 0;
    if (keep) {
        // side list and placement; if the side list overflowed, the regular
        // projection (entry order; entries per slab counted already) and LSD
        // passes run instead
        HIPCHK(h, launch_coh_merge(ca, st));
        ObjPipeArgs af = a;
        af.run_if = &ctrl->coh_fallback;
        af.no_slab_count = 1;
        HIPCHK(h, launch_project(af, st));
        s.run_if = &ctrl->coh_fallback;
        // the passes claim their tiles, so one workgroup per CU sorts them
        // all; launched that small, they cost little when they return at once
        s.grid = h->cus;
    }
    HIPCHK(h, sort_kv32(s, a.keys, a.vals, a.keys2, a.vals2, &which, st));
    if (coherent) {
        HIPCHK(h, launch_coh_save(ca, st));
        h->coh_age = keep ? h->coh_age + 1 : 1;
        h->coh_valid = true;
        h->coh_n = n;
        h->coh_slabs = slabs ? 1 : 0;
        h->coh_gen = h->buf_gen;
    }
    mark(h, st);
    ObjPipeArgs ae = a;  // entry space
    ae.n = n_e;
    ae.seg = seg_e;
    ae.ntiles = ntiles_e;
    HIPCHK(h, launch_gather(ae, which ? a.keys2 : a.keys, which ? a.vals2 : a.vals, st));
    mark(h, st);

    SweepArgs w{};
    w.rec = h->rec.p;
    w.grp = h->grp.p;
    w.grp_end = (const uint32_t*)h->grp_end.p;
    w.n = n_e;
    w.seg = seg_e;
    w.tile = tile;
    w.ntiles = ntiles_e;
    w.slabs = slabs ? 1 : 0;
    w.xcd_chunk = (uint32_t)h->sweep_chunk;
    w.group = (h->desc.flags & GSAP_FLAG_SKIP_SAME_GROUP) ? group : nullptr;
    w.ctrl = ctrl;
    w.kind = kind;
    w.sym = sym;
    w.count_only = count_only;
    uint2* outp = (uint2*)pairs;
    uint2* t0p = (uint2*)h->ptmp[0].p;
    uint2* t1p = (uint2*)h->ptmp[1].p;
    w.out = count_only ? nullptr : t0p;
    w.slot_cap = count_only ? 0 : h->slot_cap;
    w.region_stride = h->region_stride;
    w.nregions = count_only ? 1u : (uint32_t)h->pair_regions;
    // with the local stage (at most kSweepHistPasses onesweep passes) the
    // sweep counts the pair sort's digits as it writes the pairs
    const bool sweep_hist = !count_only && local && npass_pair <= kSweepHistPasses;
    w.hist = sweep_hist ? a.hist_pair : nullptr;
    // with per-XCD pair regions the first pass claims its slots per region
    // (no look-back); the sweep then also counts that digit per region
    const bool msplit = sweep_hist && npass_pair > 1 && h->pair_regions == kMaxRegions;
    w.hist_reg = msplit ? a.pair_claim : nullptr;
    // two passes: the second one runs in parts (one per XCD), from the
    // per-part digit counts the sweep takes in place of its histogram
    uint32_t* const joint = msplit && npass_pair == 2 ? a.pair_claim + 2 * kMaxRegions * 256 : nullptr;
    w.joint = joint;
    w.hist_npass = (int)npass_pair;
    w.seg_pairs = count_only ? h->lab_pairs : nullptr;
    w.hist_shift0 = (int)shift0;
    w.bbits = (int)bbits;
    HIPCHK(h, launch_sweep(w, st));
    mark(h, st);
    const uint32_t nreg = (uint32_t)h->pair_regions;
    if (!count_only)
        HIPCHK(h, launch_pair_regions(ctrl, h->slot_cap, nreg, sort_tile(h->pair_items), st));
    if (rows) {
        // the row counters are `has` (zero between updates), the row starts
        // `rowstart`, the block totals `block_sums`; the long-row list lives
        // in the window-start table
        HIPCHK(h, launch_pair_rows(t0p, t1p, t0p, outp, ctrl, nreg, h->region_stride, h->sort_cap, cap,
                                   n, (int)bbits, (uint32_t*)h->has.p, (uint32_t*)h->rowstart.p,
                                   (uint32_t*)h->block_sums.p, sym ? entries : nullptr,
                                   (uint32_t*)h->wstart.p, st));
        mark(h, st);
        if (scene_out)
            HIPCHK(h, launch_scene_offsets(outp, ctrl, cap, seg.seg_start, seg.nseg, scene_out, st));
    } else if (!count_only) {
        // onesweep passes of the pair sort by reduce-then-scan (the digit
        // totals come from its own tile counts: no histogram pass) between
        // kRtsMinTiles and kRtsMaxPairTiles tiles of the capacity
        const uint64_t ptiles = (h->sort_cap + sort_tile(h->pair_items) - 1) / sort_tile(h->pair_items);
        const bool pair_rts = ptiles >= kRtsMinTiles && ptiles <= kRtsMaxPairTiles;
        if (!sweep_hist && !pair_rts)
            HIPCHK(h, launch_pair_hist(t0p, &ctrl->sort_n, h->sort_cap, a.hist_pair, npass_pair,
                                       bbits, pair_bits, shift0, ctrl->region_start,
                                       ctrl->region_valid, h->region_stride, nreg, st));
        mark(h, st);
        SortArgs p{};
        p.n_dev = &ctrl->sort_n;
        p.n_dev_rest = &ctrl->sort_real;
        p.cap = h->sort_cap;
        p.region_start = ctrl->region_start;
        p.region_valid = ctrl->region_valid;
        p.region_stride = h->region_stride;
        p.nregions = nreg;
        p.shift0 = (int)shift0;
        p.hist = a.hist_pair;
        p.hist_reg = msplit ? a.pair_claim : nullptr;
        p.claim = msplit ? a.pair_claim + kMaxRegions * 256 : nullptr;
        p.joint = joint;
        // after the per-part pass (keys of at most 40 bits): 32-bit items for
        // the local stage (it reads 4 bytes per pair; the top bits are the
        // pass's digit, implied by the position)
        const bool out32 = joint != nullptr;
        uint32_t* const dstarts = a.pair_claim + kPairClaimWords;
        uint32_t* const bst = out32 ? pair_local_bucket_starts((uint32_t*)h->wstart.p, h->slot_cap) : nullptr;
        p.st_out = out32 ? dstarts : nullptr;
        p.bst = bst;
        p.xcnt = out32 ? bst + kBucketWords : nullptr;
        p.xcur = out32 ? bst + kBucketWords + kRadixSq : nullptr;
        // the per-part pass: tiles assigned statically, many workgroups
        // per CU for the balance
        p.xgrid = 12 * h->cus;
        p.npass = npass_pair;
        p.lookback = (unsigned long long*)h->lb_pair.p;
        p.ctrl = ctrl;
        p.ctr_base = kMaxPasses;
        p.bits = (int)pair_bits;
        p.items = h->pair_items;
        p.bbits = (int)bbits;
        p.grid = (uint32_t)h->pair_grid;
        p.last_n = local ? nullptr : &ctrl->num_pairs;
        p.rts = pair_rts ? 1 : 0;
        p.last_cap = cap;
        const bool pev = h->timing && h->nev + (int)npass_pair + 3 <= gsap_sap::kEv;
        p.pass_ev = pev ? &h->ev[h->nev] : nullptr;
        HIPCHK(h, sort_pairs(p, t0p, t1p, local ? nullptr : outp, st));
        if (pev) h->nev += (int)npass_pair;
        if (local) {
            // the onesweep passes leave the bucket-sorted pairs in t1p after an
            // odd number of passes, in t0p after an even one
            uint2* x = (npass_pair & 1u) ? t1p : t0p;
            uint2* y = (npass_pair & 1u) ? t0p : t1p;
            // after the per-part pass its bucket starts give the local
            // stage's window starts without a search
            HIPCHK(h, launch_pair_local(x, y, outp, &ctrl->sort_real, &ctrl->num_pairs, cap,
                                        h->slot_cap, (uint32_t*)h->wstart.p, &ctrl->local_big,
                                        (int)bbits, (int)shift0, out32 ? dstarts : nullptr, bst, st));
            mark(h, st);
        }
        if (sym)
            HIPCHK(h, launch_entries(outp, &ctrl->num_pairs, cap, n, (uint32_t*)h->rowstart.p,
                                     (uint32_t*)h->has.p, (uint32_t*)h->block_sums.p, entries,
                                     ctrl, st));
        if (scene_out)
            HIPCHK(h, launch_scene_offsets(outp, ctrl, cap, seg.seg_start, seg.nseg, scene_out, st));
    }
    mark(h, st);
    // the synchronous entry points read the counts and the axis back in one copy
    unsigned long long* const own = (unsigned long long*)h->counts.p;
    HIPCHK(h, launch_finish(ctrl, count_only ? ~0ull : cap, d_counts, sym, (const float*)h->axis.p,
                            d_counts == own ? own + 5 : nullptr, st));
    mark(h, st);
    return GSAP_OK;
}

static int32_t check_common(gsap_sap* h, const void* in, uint32_t n, const uint32_t* group) {
    if (!h) return fail(nullptr, GSAP_ERR_INVALID_ARG, "null handle");
    if (n > 0 && !in) return fail(h, GSAP_ERR_INVALID_ARG, "d_in is NULL");
    if (n > 0x7FFFFFFFu) return fail(h, GSAP_ERR_INVALID_ARG, "n exceeds 2^31-1");
    if ((h->desc.flags & GSAP_FLAG_SKIP_SAME_GROUP) && n > 0 && !group)
        return fail(h, GSAP_ERR_INVALID_ARG, "GSAP_FLAG_SKIP_SAME_GROUP needs d_group");
    if (slab_mode(h) && n > (1u << 30))
        return fail(h, GSAP_ERR_UNSUPPORTED, "GSAP_FLAG_SLABS supports n <= 2^30");
    return GSAP_OK;
}

extern "C" int32_t gsap_sap_collide_async(gsap_sap* h, void* stream, const void* d_in, uint32_t n,
                                          const uint32_t* d_group, void* d_pairs,
                                          uint64_t pair_capacity, int32_t* d_entries,
                                          uint64_t* d_counts) {
    int32_t rc = check_common(h, d_in, n, d_group);
    if (rc) return rc;
    if (!d_counts) return fail(h, GSAP_ERR_INVALID_ARG, "d_counts is NULL");
    hipStream_t st = (hipStream_t)stream;
    const bool count_only = (pair_capacity == 0 || d_pairs == nullptr);
    if (!count_only && h->desc.output_kind == GSAP_OUT_SYM_CSR && !d_entries)
        return fail(h, GSAP_ERR_INVALID_ARG, "SYM_CSR output needs d_entries");
    if (n == 0) {
        HIPCHK(h, launch_zero_u64((unsigned long long*)d_counts, 5, st));
        return GSAP_OK;
    }
    const uint32_t tile = sort_tile(h->items);
    const uint32_t ntiles = (n + tile - 1) / tile;
    WsReq req;
    if ((rc = req_single(h, n, pair_capacity, count_only, req))) return rc;
    if ((rc = ws_ensure(h, req, st))) return rc;
    SegTable seg{};
    seg.nseg = 1;
    return enqueue_update(h, st, d_in, n, seg, ntiles, d_group, d_pairs, pair_capacity, d_entries,
                          count_only, (unsigned long long*)d_counts, nullptr, true);
}

extern "C" int32_t gsap_sap_collide_noalloc(gsap_sap* h, void* stream, const void* d_in,
                                            uint32_t n, const uint32_t* d_group, void* d_pairs,
                                            uint64_t pair_capacity, int32_t* d_entries,
                                            uint64_t* d_counts) {
    int32_t rc = check_common(h, d_in, n, d_group);
    if (rc) return rc;
    if (!d_counts) return fail(h, GSAP_ERR_INVALID_ARG, "d_counts is NULL");
    hipStream_t st = (hipStream_t)stream;
    const bool count_only = (pair_capacity == 0 || d_pairs == nullptr);
    if (!count_only && h->desc.output_kind == GSAP_OUT_SYM_CSR && !d_entries)
        return fail(h, GSAP_ERR_INVALID_ARG, "SYM_CSR output needs d_entries");
    if (n == 0) {
        HIPCHK(h, launch_zero_u64((unsigned long long*)d_counts, 5, st));
        return GSAP_OK;
    }
    void* base = nullptr;
    size_t bytes = 0;
    HIPCHK(h, scratch_region(&base, &bytes));
    WsReq req;
    // the sizes of this call only (not the handle's max_n)
    const uint32_t max_n = h->desc.max_n;
    h->desc.max_n = 0;
    rc = req_single(h, n, pair_capacity, count_only, req);
    h->desc.max_n = max_n;
    if (rc) return rc;
    const size_t need = ws_total(req.b);
    if (need > bytes) {
        char buf[200];
        snprintf(buf, sizeof buf,
                 "gsap_sap_collide_noalloc: this update needs %llu bytes of scratch, the static "
                 "scratch holds %llu (GSAP_NOALLOC_SCRATCH_BYTES)",
                 (unsigned long long)need, (unsigned long long)bytes);
        return fail(h, GSAP_ERR_CAPACITY, buf);
    }
    // This call's buffers are carved from the scratch in place of the
    // handle's own, which are kept as they are and restored afterwards (no
    // buffer of the handle is freed or moved, so a captured batched update
    // stays valid).
    Buf saved[W_COUNT];
    size_t off = 0;
    for (int id = 0; id < W_COUNT; ++id) {
        Buf* b = ws_buf(h, id);
        saved[id] = *b;
        b->p = req.b[id] ? (char*)base + off : nullptr;
        b->bytes = req.b[id];
        b->owned = false;
        off += ws_align(req.b[id]);
    }
    // the scratch holds whatever the last call on this GPU left: clear the
    // buffers that must start zeroed, every call
    rc = ws_clear(h, st);
    if (!rc) {
        SegTable seg{};
        seg.nseg = 1;
        const uint32_t tile = sort_tile(h->items);
        rc = enqueue_update(h, st, d_in, n, seg, (n + tile - 1) / tile, d_group, d_pairs,
                            pair_capacity, d_entries, count_only, (unsigned long long*)d_counts,
                            nullptr);
    }
    for (int id = 0; id < W_COUNT; ++id) *ws_buf(h, id) = saved[id];
    return rc;
}

static int32_t read_counts(gsap_sap* h, hipStream_t st, uint64_t cap, bool count_only,
                           uint64_t* h_num_pairs, uint32_t* h_num_entries, gsap_stats* stats,
                           uint32_t n) {
    // counts [0, 5) and the axis words [5, 7) (written by the finish kernel)
    HIPCHK(h, hipMemcpyAsync(h->h_counts, h->counts.p, 7 * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    const unsigned long long* c = h->h_counts;
    float ax[4];
    memcpy(ax, c + 5, sizeof ax);
    if (stats) memcpy(stats->axis, ax, 3 * sizeof(float));
    const float src = ax[3];
    if (h->timing && h->nev >= 2) {
        for (int i = 0; i + 1 < h->nev; ++i)
            (void)hipEventElapsedTime(&h->stage_ms[i], h->ev[i], h->ev[i + 1]);
    }
    if (h_num_pairs) *h_num_pairs = c[0];
    if (h_num_entries) *h_num_entries = (uint32_t)c[1];
    if (stats) {
        const bool sym = h->desc.output_kind == GSAP_OUT_SYM_CSR;
        stats->num_pairs = sym ? c[0] / 2 : c[0];
        stats->sweep_tests = c[2];
        stats->axis_source = (int32_t)src;
        stats->n = n;
        stats->ms_total = 0.0f;
        if (h->timing && h->nev >= 2) (void)hipEventElapsedTime(&stats->ms_total, h->ev[0], h->ev[h->nev - 1]);
    }
    if (c[3] & kFlagTimeout) return fail(h, GSAP_ERR_HIP, "internal: sort look-back wait expired");
    if (c[3] & kFlagSlab) return fail(h, GSAP_ERR_HIP, "internal: a box spans more than two slabs");
    if (c[3] & kFlagNonFinite) return fail(h, GSAP_ERR_NONFINITE, "non-finite value in input boxes");
    if (!count_only && c[0] > cap) {
        char buf[160];
        snprintf(buf, sizeof buf, "pair capacity %llu too small: %llu required",
                 (unsigned long long)cap, (unsigned long long)c[0]);
        return fail(h, GSAP_ERR_CAPACITY, buf);
    }
    return GSAP_OK;
}

extern "C" int32_t gsap_sap_collide(gsap_sap* h, void* stream, const void* d_in, uint32_t n,
                                    const uint32_t* d_group, void* d_pairs, uint64_t pair_capacity,
                                    int32_t* d_entries, uint64_t* h_num_pairs,
                                    uint32_t* h_num_entries, gsap_stats* stats) {
    int32_t rc = check_common(h, d_in, n, d_group);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const bool count_only = (pair_capacity == 0 || d_pairs == nullptr);
    if (n == 0) {
        if (h_num_pairs) *h_num_pairs = 0;
        if (h_num_entries) *h_num_entries = 0;
        if (stats) memset(stats, 0, sizeof(*stats));
        return GSAP_OK;
    }
    rc = gsap_sap_collide_async(h, stream, d_in, n, d_group, d_pairs, pair_capacity, d_entries,
                                (uint64_t*)h->counts.p);
    if (rc) return rc;
    return read_counts(h, st, pair_capacity, count_only, h_num_pairs, h_num_entries, stats, n);
}

extern "C" int32_t gsap_sap_query_workspace(gsap_sap* h, uint32_t n, uint64_t pair_capacity,
                                            uint64_t* bytes) {
    if (!h || !bytes) return fail(h, GSAP_ERR_INVALID_ARG, "null argument");
    if (n > 0x7FFFFFFFu) return fail(h, GSAP_ERR_INVALID_ARG, "n exceeds 2^31-1");
    WsReq req;
    // req_pairs records the layout of that capacity in the handle; keep the
    // layout of the last update
    const uint64_t sc = h->slot_cap, rs = h->region_stride, so = h->sort_cap;
    const int32_t rc = req_single(h, n, pair_capacity, pair_capacity == 0, req);
    h->slot_cap = sc;
    h->region_stride = rs;
    h->sort_cap = so;
    if (rc) return rc;
    *bytes = ws_total(req.b);
    return GSAP_OK;
}

extern "C" int32_t gsap_sap_set_workspace(gsap_sap* h, void* d_ws, uint64_t bytes) {
    if (!h) return fail(nullptr, GSAP_ERR_INVALID_ARG, "null handle");
    if (d_ws && ((uintptr_t)d_ws & 255u))
        return fail(h, GSAP_ERR_INVALID_ARG, "workspace must be 256-byte aligned");
    if (d_ws && bytes == 0) return fail(h, GSAP_ERR_INVALID_ARG, "workspace of 0 bytes");
    // updates still in flight may use the current buffers
    HIPCHK(h, hipDeviceSynchronize());
    for (int id = 0; id < W_COUNT; ++id) {
        release(*ws_buf(h, id));
        h->ws_size[id] = 0;
    }
    h->ws_base = d_ws;
    h->ws_bytes = d_ws ? bytes : 0;
    h->ws_carved = false;
    h->coh_valid = false;
    ++h->buf_gen;
    return GSAP_OK;
}

// Scene tables for the batched path: tiles never straddle scenes.
struct SceneTables {
    std::vector<uint32_t> start, cnt, t0, tseg;
    uint32_t ntiles = 0;
    // slab mode: the same table over entries (scene s owns entries
    // [2 start, 2 (start + n)))
    std::vector<uint32_t> start_e, cnt_e, t0_e, tseg_e;
    uint32_t ntiles_e = 0;
};

static int32_t scene_tables(gsap_sap* h, const uint32_t* off, uint32_t ns, SceneTables& t) {
    const uint32_t tile = sort_tile(h->items);
    t.start.resize(ns);
    t.cnt.resize(ns);
    t.t0.resize(ns + 1);
    t.tseg.clear();
    uint32_t nt = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        if (off[s + 1] < off[s]) return fail(h, GSAP_ERR_INVALID_ARG, "scene offsets decrease");
        t.start[s] = off[s];
        t.cnt[s] = off[s + 1] - off[s];
        t.t0[s] = nt;
        const uint32_t k = (t.cnt[s] + tile - 1) / tile;
        for (uint32_t i = 0; i < k; ++i) t.tseg.push_back(s);
        nt += k;
    }
    t.t0[ns] = nt;
    t.ntiles = nt;
    if (slab_mode(h)) {
        t.start_e.resize(ns);
        t.cnt_e.resize(ns);
        t.t0_e.resize(ns + 1);
        t.tseg_e.clear();
        uint32_t ne = 0;
        for (uint32_t s = 0; s < ns; ++s) {
            t.start_e[s] = 2u * t.start[s];
            t.cnt_e[s] = 2u * t.cnt[s];
            t.t0_e[s] = ne;
            const uint32_t k = (t.cnt_e[s] + tile - 1) / tile;
            for (uint32_t i = 0; i < k; ++i) t.tseg_e.push_back(s);
            ne += k;
        }
        t.t0_e[ns] = ne;
        t.ntiles_e = ne;
    }
    return GSAP_OK;
}

// Writes the scene tables into the (already ensured) device buffers.  The
// tables' contents change, so a captured batched update is stale.
static int32_t upload_scenes(gsap_sap* h, const SceneTables& t) {
    const size_t ns = t.start.size();
    HIPCHK(h, hipMemcpy(h->seg_start.p, t.start.data(), ns * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->seg_n.p, t.cnt.data(), ns * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->seg_tile0.p, t.t0.data(), (ns + 1) * 4, hipMemcpyHostToDevice));
    if (t.ntiles) HIPCHK(h, hipMemcpy(h->tile_seg.p, t.tseg.data(), (size_t)t.ntiles * 4, hipMemcpyHostToDevice));
    if (!t.start_e.empty() && h->seg_start_e.p) {
        HIPCHK(h, hipMemcpy(h->seg_start_e.p, t.start_e.data(), ns * 4, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->seg_n_e.p, t.cnt_e.data(), ns * 4, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->seg_tile0_e.p, t.t0_e.data(), (ns + 1) * 4, hipMemcpyHostToDevice));
        if (t.ntiles_e)
            HIPCHK(h, hipMemcpy(h->tile_seg_e.p, t.tseg_e.data(), (size_t)t.ntiles_e * 4,
                                hipMemcpyHostToDevice));
    }
    h->ntiles_e = t.ntiles_e;
    ++h->buf_gen;
    return GSAP_OK;
}

// Scene tables of `off` on the device with the object buffers (and, when
// cap > 0, the pair buffers) of that batch ensured in the same step.
static int32_t prepare_scenes(gsap_sap* h, hipStream_t st, const uint32_t* off, uint32_t ns,
                              uint64_t cap, bool with_pairs, uint32_t* ntiles_out) {
    h->coh_valid = false;  // the object buffers will hold another object set
    SceneTables t;
    int32_t rc = scene_tables(h, off, ns, t);
    if (rc) return rc;
    WsReq req;
    req_obj(h, off[ns], ns, std::max(t.ntiles, 1u), req);
    if (with_pairs && (rc = req_pairs(h, cap, req))) return rc;
    if ((rc = ws_ensure(h, req, st))) return rc;
    // the tables are written by synchronous copies after the carve's clears
    HIPCHK(h, hipStreamSynchronize(st));
    if ((rc = upload_scenes(h, t))) return rc;
    *ntiles_out = t.ntiles;
    return GSAP_OK;
}

extern "C" int32_t gsap_sap_collide_batched(gsap_sap* h, void* stream, const void* d_in,
                                            const uint32_t* h_scene_offsets, uint32_t n_scenes,
                                            void* d_pairs, uint64_t pair_capacity,
                                            uint64_t* d_scene_pair_offsets, uint64_t* h_num_pairs,
                                            gsap_stats* stats) {
    if (!h) return fail(nullptr, GSAP_ERR_INVALID_ARG, "null handle");
    if (!h_scene_offsets || n_scenes == 0) return fail(h, GSAP_ERR_INVALID_ARG, "no scenes");
    if (h->desc.output_kind != GSAP_OUT_CANON)
        return fail(h, GSAP_ERR_UNSUPPORTED, "batched mode emits canonical pairs only");
    if (h->desc.flags & GSAP_FLAG_SKIP_SAME_GROUP)
        return fail(h, GSAP_ERR_UNSUPPORTED, "batched mode has no group filter");
    const uint32_t n = h_scene_offsets[n_scenes];
    if (h_scene_offsets[0] != 0) return fail(h, GSAP_ERR_INVALID_ARG, "scene offsets must start at 0");
    int32_t rc = check_common(h, d_in, n, nullptr);
    if (rc) return rc;
    const bool count_only = (pair_capacity == 0 || d_pairs == nullptr);
    if (!count_only && !d_scene_pair_offsets)
        return fail(h, GSAP_ERR_INVALID_ARG, "d_scene_pair_offsets is NULL");
    // An untuned handle's batched updates claim pair slots on one counter per
    // XCD: the sweep makes at least one claim per workgroup, and with
    // thousands of small scenes one shared counter serialises them (2,048
    // scenes: slab sweep 3.7 -> 2.6 ms).
    if (!h->tune_set && !count_only) h->pair_regions = kMaxRegions;
    hipStream_t st = (hipStream_t)stream;
    if (n == 0) {
        if (h_num_pairs) *h_num_pairs = 0;
        if (d_scene_pair_offsets)
            HIPCHK(h, launch_zero_u64((unsigned long long*)d_scene_pair_offsets, n_scenes + 1, st));
        return GSAP_OK;
    }
    const bool same = h->gexec && h->g_in == d_in && h->g_pairs == d_pairs &&
                      h->g_cap == pair_capacity &&
                      h->g_scene_out == (unsigned long long*)d_scene_pair_offsets &&
                      h->g_offsets.size() == n_scenes + 1 &&
                      std::equal(h->g_offsets.begin(), h->g_offsets.end(), h_scene_offsets) &&
                      h->g_gen == h->buf_gen && !h->timing;
    if (!same) {
        uint32_t nt = 0;
        HIPCHK(h, hipStreamSynchronize(st));
        rc = prepare_scenes(h, st, h_scene_offsets, n_scenes, pair_capacity, !count_only, &nt);
        if (rc) return rc;
        SegTable seg{};
        seg.seg_start = (const uint32_t*)h->seg_start.p;
        seg.seg_n = (const uint32_t*)h->seg_n.p;
        seg.seg_tile0 = (const uint32_t*)h->seg_tile0.p;
        seg.tile_seg = (const uint32_t*)h->tile_seg.p;
        seg.nseg = n_scenes;
        seg.ntiles = nt;
        if (h->gexec) {
            (void)hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
        if (h->timing || count_only) {
            rc = enqueue_update(h, st, d_in, n, seg, nt, nullptr, d_pairs, pair_capacity, nullptr,
                                count_only, (unsigned long long*)h->counts.p,
                                count_only ? nullptr : (unsigned long long*)d_scene_pair_offsets);
            if (rc) return rc;
        } else {
            if (!h->cap_stream) HIPCHK(h, hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
            HIPCHK(h, hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_update(h, h->cap_stream, d_in, n, seg, nt, nullptr, d_pairs, pair_capacity,
                                nullptr, false, (unsigned long long*)h->counts.p,
                                (unsigned long long*)d_scene_pair_offsets);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(h->cap_stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIPCHK(h, e);
            e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(h, e);
            h->g_in = d_in;
            h->g_pairs = d_pairs;
            h->g_cap = pair_capacity;
            h->g_scene_out = (unsigned long long*)d_scene_pair_offsets;
            h->g_offsets.assign(h_scene_offsets, h_scene_offsets + n_scenes + 1);
            h->g_gen = h->buf_gen;
            HIPCHK(h, hipGraphLaunch(h->gexec, st));
        }
    } else {
        HIPCHK(h, hipGraphLaunch(h->gexec, st));
    }
    return read_counts(h, st, pair_capacity, count_only, h_num_pairs, nullptr, stats, n);
}

extern "C" int32_t gsap_sap_set_timing(gsap_sap* h, int32_t enable) {
    if (!h) return fail(nullptr, GSAP_ERR_INVALID_ARG, "null handle");
    h->timing = enable != 0;
    return GSAP_OK;
}

extern "C" int32_t gsap_sap_stage_times(gsap_sap* h, float* h_ms, int32_t max_stages,
                                        int32_t* n_stages) {
    if (!h || !h_ms || !n_stages) return fail(h, GSAP_ERR_INVALID_ARG, "null argument");
    int k = h->nev > 1 ? h->nev - 1 : 0;
    if (k > max_stages) k = max_stages;
    for (int i = 0; i < k; ++i) h_ms[i] = h->stage_ms[i];
    *n_stages = k;
    return GSAP_OK;
}

// ---------------------------------------------------------------- radix sort
struct gsap_radix_sort {
    std::string err;
    uint32_t max_n = 0;
    int keys_only = 0;
    Buf tk, tv, dv, hist, lb, ctrl;
};

static int32_t rs_fail(gsap_radix_sort* s, int32_t code, const std::string& m) {
    if (s) s->err = m;
    g_err = m;
    return code;
}

extern "C" int32_t gsap_radix_sort_create(uint32_t max_elements, int32_t keys_only,
                                          gsap_radix_sort** out) {
    if (!out) return rs_fail(nullptr, GSAP_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    if (gsap_device_count() <= 0) return rs_fail(nullptr, GSAP_ERR_HIP, "no HIP device");
    gsap_radix_sort* s = new gsap_radix_sort();
    uint64_t gen = 0;  // the sorter captures no graph
    s->max_n = max_elements;
    s->keys_only = keys_only;
    const uint32_t tiles = (max_elements + sort_tile(8) - 1) / sort_tile(8) + 1;
    hipError_t e = grow(s->tk, (size_t)max_elements * 4, &gen);
    if (e == hipSuccess) e = grow(s->tv, (size_t)max_elements * 4, &gen);
    if (e == hipSuccess) e = grow(s->dv, (size_t)max_elements * 4, &gen);
    if (e == hipSuccess) e = grow(s->hist, kMaxPasses * 256 * 4, &gen);
    if (e == hipSuccess) e = grow(s->lb, (size_t)tiles * 256 * 8, &gen);
    if (e == hipSuccess) e = hipMemset(s->lb.p, 0, (size_t)tiles * 256 * 8);
    if (e == hipSuccess) e = grow(s->ctrl, sizeof(Ctrl), &gen);
    if (e == hipSuccess) e = hipMemset(s->ctrl.p, 0, sizeof(Ctrl));
    if (e != hipSuccess) {
        std::string m = std::string("gsap_radix_sort_create: ") + hipGetErrorString(e);
        gsap_radix_sort_destroy(s);
        return rs_fail(nullptr, e == hipErrorOutOfMemory ? GSAP_ERR_OOM : GSAP_ERR_HIP, m);
    }
    *out = s;
    return GSAP_OK;
}

extern "C" int32_t gsap_radix_sort_destroy(gsap_radix_sort* s) {
    if (!s) return GSAP_OK;
    (void)hipDeviceSynchronize();
    release(s->tk); release(s->tv); release(s->dv); release(s->hist); release(s->lb); release(s->ctrl);
    delete s;
    return GSAP_OK;
}

static int32_t rs_sort(gsap_radix_sort* s, hipStream_t st, uint32_t* keys, uint32_t* vals, uint32_t n,
                       uint32_t key_bits, bool is_float) {
    if (!s) return rs_fail(nullptr, GSAP_ERR_INVALID_ARG, "null sorter");
    if (n > s->max_n) return rs_fail(s, GSAP_ERR_CAPACITY, "n exceeds maxElements");
    if (n == 0) return GSAP_OK;
    if (!keys || (!vals && !s->keys_only)) return rs_fail(s, GSAP_ERR_INVALID_ARG, "null keys/values");
    if (key_bits == 0 || key_bits > 32) return rs_fail(s, GSAP_ERR_INVALID_ARG, "key_bits in 1..32");
    uint32_t* v = vals ? vals : (uint32_t*)s->dv.p;
    const uint32_t npass = (key_bits + 7) / 8;
    hipError_t e = hipSuccess;
#define RSCHK(x)                                                                 \
    do {                                                                         \
        e = (x);                                                                 \
        if (e != hipSuccess) return rs_fail(s, GSAP_ERR_HIP, hipGetErrorString(e)); \
    } while (0)
    if (is_float) RSCHK(launch_flip(keys, n, 0, st));
    RSCHK(launch_sort_prep((Ctrl*)s->ctrl.p, (uint32_t*)s->hist.p, npass * 256, 0, st));
    RSCHK(launch_key_hist(keys, n, (uint32_t*)s->hist.p, (int)npass, 8, 0, st));
    SortArgs a{};
    a.n = n;
    a.hist = (uint32_t*)s->hist.p;
    a.npass = npass;
    a.lookback = (unsigned long long*)s->lb.p;
    a.ctrl = (Ctrl*)s->ctrl.p;
    a.bits = 8;
    a.items = 8;
    // reduce-then-scan passes from kRtsMinTiles tiles (the look-back buffer,
    // 2 KB per tile, holds their tile tables)
    a.rts = (n + sort_tile(8) - 1) / sort_tile(8) >= kRtsMinTiles;
    int which = 0;
    RSCHK(sort_kv32(a, keys, v, (uint32_t*)s->tk.p, (uint32_t*)s->tv.p, &which, st));
    if (which) {
        RSCHK(hipMemcpyAsync(keys, s->tk.p, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        if (vals) RSCHK(hipMemcpyAsync(vals, s->tv.p, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    }
    if (is_float) RSCHK(launch_flip(keys, n, 1, st));
#undef RSCHK
    return GSAP_OK;
}

extern "C" int32_t gsap_radix_sort_u32(gsap_radix_sort* s, void* stream, uint32_t* d_keys,
                                       uint32_t* d_values, uint32_t n, uint32_t key_bits) {
    return rs_sort(s, (hipStream_t)stream, d_keys, d_values, n, key_bits, false);
}

extern "C" int32_t gsap_radix_sort_f32(gsap_radix_sort* s, void* stream, float* d_keys,
                                       uint32_t* d_values, uint32_t n, uint32_t key_bits,
                                       int32_t negative_keys) {
    (void)key_bits;
    (void)negative_keys;  // the order-preserving map handles either sign
    return rs_sort(s, (hipStream_t)stream, (uint32_t*)d_keys, d_values, n, 32, true);
}

// ---------------------------------------------------------------- learned-axis support
namespace gsap_internal_api {
int32_t attach_model(gsap_sap* h, const uint16_t* d_wb, size_t wb_bytes, const float* d_master,
                     size_t master_bytes, float conf_min) {
    if (!h) return fail(nullptr, GSAP_ERR_INVALID_ARG, "null handle");
    if (!(conf_min > 0.0f && conf_min <= 1.0f)) return fail(h, GSAP_ERR_INVALID_ARG, "conf_min in (0,1]");
    HIPCHK(h, grow(h->mlp_wb, wb_bytes, &h->buf_gen));
    HIPCHK(h, grow(h->mlp_master, master_bytes, &h->buf_gen));
    HIPCHK(h, grow(h->cand, kCand * 3 * 4, &h->buf_gen));
    HIPCHK(h, hipDeviceSynchronize());
    HIPCHK(h, hipMemcpy(h->mlp_wb.p, d_wb, wb_bytes, hipMemcpyDeviceToDevice));
    HIPCHK(h, hipMemcpy(h->mlp_master.p, d_master, master_bytes, hipMemcpyDeviceToDevice));
    float cand[kCand * 3];
    gsap_axis_candidates(cand);
    HIPCHK(h, hipMemcpy(h->cand.p, cand, sizeof cand, hipMemcpyHostToDevice));
    h->conf_min = conf_min;
    h->has_model = true;
    ++h->buf_gen;  // conf_min is a kernel argument of a captured batched update
    return GSAP_OK;
}
}  // namespace gsap_internal_api

static SegTable seg_table(gsap_sap* h, uint32_t ns, uint32_t nt) {
    SegTable seg{};
    seg.seg_start = (const uint32_t*)h->seg_start.p;
    seg.seg_n = (const uint32_t*)h->seg_n.p;
    seg.seg_tile0 = (const uint32_t*)h->seg_tile0.p;
    seg.tile_seg = (const uint32_t*)h->tile_seg.p;
    seg.nseg = ns;
    seg.ntiles = nt;
    return seg;
}

extern "C" int32_t gsap_sap_features_batched(gsap_sap* h, void* stream, const void* d_in,
                                             const uint32_t* h_off, uint32_t ns, float* d_features) {
    if (!h || !d_in || !h_off || ns == 0 || !d_features)
        return fail(h, GSAP_ERR_INVALID_ARG, "gsap_sap_features_batched: bad argument");
    if (h->desc.input_kind != GSAP_INPUT_AABB)
        return fail(h, GSAP_ERR_UNSUPPORTED, "features are defined for AABB scenes");
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(h, hipStreamSynchronize(st));
    uint32_t nt = 0;
    if (int32_t rc = prepare_scenes(h, st, h_off, ns, 0, false, &nt)) return rc;
    ObjPipeArgs a{};
    a.in = d_in;
    a.kind = h->desc.input_kind;
    a.n = h_off[ns];
    a.seg = seg_table(h, ns, nt);
    a.tile = sort_tile(h->items);
    a.ntiles = nt;
    a.axis_mode = GSAP_AXIS_PCA;
    a.axis = (float*)h->axis.p;
    a.partials = (double*)h->partials.p;
    a.ctrl = (Ctrl*)h->ctrl.p;
    a.hist_obj = (uint32_t*)h->hist_obj.p;
    a.npass_obj = 4;
    a.bits_obj = 8;
    a.hist_pair = (uint32_t*)h->hist_pair.p;
    a.npass_pair = 0;
    a.seg_mom = (double*)h->seg_mom.p;
    a.plane_hist = (uint32_t*)h->plane_hist.p;
    if (nt == 0) return fail(h, GSAP_ERR_INVALID_ARG, "all scenes empty");
    for (uint32_t s = 0; s < ns; ++s)
        if (h_off[s + 1] == h_off[s]) return fail(h, GSAP_ERR_INVALID_ARG, "empty scene");
    HIPCHK(h, launch_moments(a, st));
    HIPCHK(h, launch_features(a, (float*)h->tile_sums.p, d_features, st));
    // the moments kernel bumped the epoch and cleared counters it owns; reset
    // the accumulators the finish kernel normally resets
    HIPCHK(h, launch_finish((Ctrl*)h->ctrl.p, ~0ull, (unsigned long long*)h->counts.p, 0, nullptr, nullptr,
                            st));
    HIPCHK(h, hipStreamSynchronize(st));
    return GSAP_OK;
}

// Labels from the sweep's own counts, for scenes of any size (slab handles,
// and plain handles with a scene beyond the in-LDS judge).  For each
// candidate every scene sweeps along it (plain) or the plane normal to it
// (slabs) in one count-only batched update; the gather adds each scene's
// tests to its column and the first update's sweep each scene's pairs.
static int32_t pipeline_axis_tests(gsap_sap* h, hipStream_t st, const void* d_in, const uint32_t* h_off,
                               uint32_t ns, const float* h_axes, uint32_t naxes, uint64_t* d_tests,
                               uint64_t* d_pairs) {
    if (h_off[0] != 0) return fail(h, GSAP_ERR_INVALID_ARG, "scene offsets must start at 0");
    const uint32_t n = h_off[ns];
    if (int32_t rc = check_common(h, d_in, n, nullptr)) return rc;
    for (uint32_t k = 0; k < naxes; ++k) {
        const float* w = h_axes + 3 * (size_t)k;
        if (!(std::isfinite(w[0]) && std::isfinite(w[1]) && std::isfinite(w[2])) ||
            !(w[0] * w[0] + w[1] * w[1] + w[2] * w[2] > 0.5f))
            return fail(h, GSAP_ERR_INVALID_ARG, "axis tests: candidates must be unit vectors");
    }
    uint32_t nt = 0;
    if (int32_t rc = prepare_scenes(h, st, h_off, ns, 0, false, &nt)) return rc;
    SegTable seg{};
    seg.seg_start = (const uint32_t*)h->seg_start.p;
    seg.seg_n = (const uint32_t*)h->seg_n.p;
    seg.seg_tile0 = (const uint32_t*)h->seg_tile0.p;
    seg.tile_seg = (const uint32_t*)h->tile_seg.p;
    seg.nseg = ns;
    seg.ntiles = nt;
    HIPCHK(h, launch_zero_u64((unsigned long long*)d_tests, ns * naxes, st));
    if (d_pairs) HIPCHK(h, launch_zero_u64((unsigned long long*)d_pairs, ns, st));
    std::vector<float> nr((size_t)ns * 4);
    int32_t rc = GSAP_OK;
    for (uint32_t k = 0; k < naxes && rc == GSAP_OK; ++k) {
        for (uint32_t s = 0; s < ns; ++s) {
            nr[4 * (size_t)s] = h_axes[3 * (size_t)k];
            nr[4 * (size_t)s + 1] = h_axes[3 * (size_t)k + 1];
            nr[4 * (size_t)s + 2] = h_axes[3 * (size_t)k + 2];
            nr[4 * (size_t)s + 3] = 1.0f;
        }
        HIPCHK(h, hipStreamSynchronize(st));
        HIPCHK(h, hipMemcpy(h->normals.p, nr.data(), nr.size() * sizeof(float), hipMemcpyHostToDevice));
        h->force_normal = true;
        h->lab_tests = (unsigned long long*)d_tests + k;
        h->lab_stride = naxes;
        h->lab_pairs = k == 0 ? (unsigned long long*)d_pairs : nullptr;
        rc = enqueue_update(h, st, d_in, n, seg, nt, nullptr, nullptr, 0, nullptr, true,
                            (unsigned long long*)h->counts.p, nullptr);
        // the update's error flags, as after any count-only update
        if (rc == GSAP_OK) rc = read_counts(h, st, 0, true, nullptr, nullptr, nullptr, n);
    }
    h->force_normal = false;
    h->lab_tests = nullptr;
    h->lab_stride = 0;
    h->lab_pairs = nullptr;
    return rc;
}

extern "C" int32_t gsap_sap_axis_tests_batched(gsap_sap* h, void* stream, const void* d_in,
                                               const uint32_t* h_off, uint32_t ns, const float* h_axes,
                                               uint32_t naxes, uint64_t* d_tests, uint64_t* d_pairs) {
    if (!h || !d_in || !h_off || ns == 0 || !h_axes || naxes == 0 || !d_tests)
        return fail(h, GSAP_ERR_INVALID_ARG, "gsap_sap_axis_tests_batched: bad argument");
    bool big = false;
    for (uint32_t s = 0; s < ns; ++s) {
        if (h_off[s + 1] < h_off[s]) return fail(h, GSAP_ERR_INVALID_ARG, "scene offsets decrease");
        big = big || h_off[s + 1] - h_off[s] > 2048;
    }
    if (slab_mode(h) || big) {
        if (h->desc.input_kind != GSAP_INPUT_AABB || h->desc.output_kind != GSAP_OUT_CANON ||
            (h->desc.flags & GSAP_FLAG_SKIP_SAME_GROUP))
            return fail(h, GSAP_ERR_UNSUPPORTED,
                        "axis tests beyond 2048 objects or with slabs: AABB scenes, canonical output, no group filter");
        return pipeline_axis_tests(h, (hipStream_t)stream, d_in, h_off, ns, h_axes, naxes, d_tests, d_pairs);
    }
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(h, hipStreamSynchronize(st));
    uint32_t nt = 0;
    if (int32_t rc = prepare_scenes(h, st, h_off, ns, 0, false, &nt)) return rc;
    HIPCHK(h, grow(h->ax_tests_axes, (size_t)naxes * 12, &h->buf_gen));
    HIPCHK(h, hipMemcpy(h->ax_tests_axes.p, h_axes, (size_t)naxes * 12, hipMemcpyHostToDevice));
    HIPCHK(h, launch_axis_tests(d_in, h->desc.input_kind, (const uint32_t*)h->seg_start.p,
                                (const uint32_t*)h->seg_n.p, ns, (const float*)h->ax_tests_axes.p,
                                naxes, (unsigned long long*)d_tests, (unsigned long long*)d_pairs, st));
    HIPCHK(h, hipStreamSynchronize(st));
    return GSAP_OK;
}

// ------------------------------------------------------------------ autotuner
static bool valid_tune(const gsap_tune* t) {
    const bool si = t->sort_items == 4 || t->sort_items == 8 || t->sort_items == 12 ||
                    t->sort_items == 16;
    const bool pi = t->pair_items == 8 || t->pair_items == 16;
    return si && pi && t->sort_bits >= 6 && t->sort_bits <= 8 && t->pair_bits >= 6 &&
           t->pair_bits <= 8 && t->pair_grid >= 256 && t->pair_grid <= 16384 &&
           (t->slabs == 0 || t->slabs == 1) &&
           (t->sweep_chunk == 0 || t->sweep_chunk == 4 || t->sweep_chunk == 16 ||
            t->sweep_chunk == 64) &&
           (t->pair_regions == 1 || t->pair_regions == kMaxRegions) &&
           (t->pair_local >= 0 && t->pair_local <= 2);
}

static gsap_tune tune_of(const gsap_sap* h) {
    gsap_tune t{};
    t.sort_items = h->items;
    t.sort_bits = h->obj_bits;
    t.pair_items = h->pair_items;
    t.pair_bits = h->pair_bits;
    t.pair_grid = h->pair_grid;
    t.slabs = h->slabs;
    t.sweep_chunk = h->sweep_chunk;
    t.pair_regions = h->pair_regions;
    t.pair_local = h->pair_local;
    t.ms = h->tune_ms;
    t.from_cache = h->tune_cached;
    return t;
}

static void apply_tune(gsap_sap* h, const gsap_tune& t) {
    h->items = t.sort_items;
    h->obj_bits = t.sort_bits;
    h->pair_items = t.pair_items;
    h->pair_bits = t.pair_bits;
    h->pair_grid = t.pair_grid;
    h->slabs = t.slabs;
    h->sweep_chunk = t.sweep_chunk;
    h->pair_regions = t.pair_regions;
    h->pair_local = t.pair_local;
    h->tune_set = true;
    h->coh_valid = false;
    if (h->gexec) {  // the captured batched update used the old configuration
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    h->g_offsets.clear();
}

extern "C" int32_t gsap_sap_get_tune(gsap_sap* h, gsap_tune* out) {
    if (!h || !out) return fail(h, GSAP_ERR_INVALID_ARG, "null argument");
    *out = tune_of(h);
    return GSAP_OK;
}

extern "C" int32_t gsap_sap_set_tune(gsap_sap* h, const gsap_tune* t) {
    if (!h || !t) return fail(h, GSAP_ERR_INVALID_ARG, "null argument");
    if (!valid_tune(t))
        return fail(h, GSAP_ERR_INVALID_ARG,
                    "gsap_sap_set_tune: items 4/8/12/16 and 8/16, bits 6..8, grid 256..16384, "
                    "slabs 0/1, sweep_chunk 0/4/16/64, pair_regions 1/8, pair_local 0/1");
    apply_tune(h, *t);
    h->tune_ms = 0.0f;
    h->tune_cached = 0;
    return GSAP_OK;
}

// Cache key: device, what the update computes, and the size bucket (half
// octaves of the object count; batched: also of the scene count).
static std::string tune_key(gsap_sap* h, uint32_t n, uint32_t n_scenes, bool batched) {
    int dev = 0;
    hipDeviceProp_t prop{};
    (void)hipGetDevice(&dev);
    (void)hipGetDeviceProperties(&prop, dev);
    std::string arch = prop.gcnArchName;
    for (char& c : arch)
        if (c == ' ' || c == ':') c = '_';
    const int nb = n ? (int)std::lround(2.0 * std::log2((double)n)) : 0;
    const int sb = batched && n_scenes ? (int)std::lround(2.0 * std::log2((double)n_scenes)) : 0;
    char buf[256];
    snprintf(buf, sizeof buf, "%s|cu%d|in%d|out%d|axis%d|flags%u|%s|n%d|s%d", arch.c_str(),
             prop.multiProcessorCount, h->desc.input_kind, h->desc.output_kind, h->desc.axis_mode,
             h->desc.flags, batched ? "batched" : "single", nb, sb);
    return buf;
}

static bool cache_lookup(const char* path, const std::string& key, gsap_tune* t) {
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool hit = false;
    while (fgets(line, sizeof line, f)) {
        char k[300];
        gsap_tune c{};
        if (sscanf(line, "%299s %d %d %d %d %d %d %d %d %d %f", k, &c.sort_items, &c.sort_bits,
                   &c.pair_items, &c.pair_bits, &c.pair_grid, &c.slabs, &c.sweep_chunk,
                   &c.pair_regions, &c.pair_local, &c.ms) == 11 &&
            key == k && valid_tune(&c)) {
            *t = c;  // the last matching line wins
            hit = true;
        }
    }
    fclose(f);
    return hit;
}

static void cache_append(const char* path, const std::string& key, const gsap_tune& t) {
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %d %d %d %d %d %d %d %d %d %.4f\n", key.c_str(), t.sort_items, t.sort_bits,
            t.pair_items, t.pair_bits, t.pair_grid, t.slabs, t.sweep_chunk, t.pair_regions,
            t.pair_local, t.ms);
    fclose(f);
}

extern "C" int32_t gsap_sap_autotune(gsap_sap* h, void* stream, const void* d_in, uint32_t n,
                                     const uint32_t* d_group, const uint32_t* offs,
                                     uint32_t n_scenes, const char* cache_path, gsap_tune* out) {
    if (!h || !out) return fail(h, GSAP_ERR_INVALID_ARG, "null argument");
    const bool batched = offs != nullptr;
    if (batched) {
        if (n_scenes == 0) return fail(h, GSAP_ERR_INVALID_ARG, "no scenes");
        n = offs[n_scenes];
    }
    if (n == 0 || !d_in) return fail(h, GSAP_ERR_INVALID_ARG, "nothing to tune on");
    hipStream_t st = (hipStream_t)stream;
    const std::string key = tune_key(h, n, n_scenes, batched);
    gsap_tune best = tune_of(h);
    if (cache_path && cache_lookup(cache_path, key, &best)) {
        apply_tune(h, best);
        h->tune_ms = best.ms;
        h->tune_cached = 1;
        *out = tune_of(h);
        return GSAP_OK;
    }
    // Pair buffers for the measured updates (the pair count does not depend
    // on the configuration).
    const bool sym = h->desc.output_kind == GSAP_OUT_SYM_CSR;
    int32_t rc;
    uint64_t P = 0;
    if (batched)
        rc = gsap_sap_collide_batched(h, stream, d_in, offs, n_scenes, nullptr, 0, nullptr, &P,
                                      nullptr);
    else
        rc = gsap_sap_collide(h, stream, d_in, n, d_group, nullptr, 0, nullptr, &P, nullptr,
                              nullptr);
    if (rc) return rc;
    void* d_pairs = nullptr;
    void* d_extra = nullptr;
    const size_t extra = batched ? ((size_t)n_scenes + 1) * 8 : (sym ? (size_t)n * 4 : 0);
    HIPCHK(h, hipMalloc(&d_pairs, std::max<uint64_t>(P, 1) * 8));
    if (extra && hipMalloc(&d_extra, extra) != hipSuccess) {
        (void)hipFree(d_pairs);
        return fail(h, GSAP_ERR_OOM, "autotune: no memory for the scene offsets / entries");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const bool timing = h->timing;
    h->timing = false;
    auto run = [&]() -> int32_t {
        if (batched)
            return gsap_sap_collide_batched(h, stream, d_in, offs, n_scenes, d_pairs, P,
                                            (uint64_t*)d_extra, nullptr, nullptr);
        return gsap_sap_collide(h, stream, d_in, n, d_group, d_pairs, P, (int32_t*)d_extra,
                                nullptr, nullptr, nullptr);
    };
    // Median of `reps` timed updates after one untimed (graph capture, first
    // touch of newly grown buffers).
    auto measure = [&](const gsap_tune& t, float* ms, int reps = 3) -> int32_t {
        apply_tune(h, t);
        int32_t r = run();
        if (r) return r;
        float v[9];
        for (int k = 0; k < reps; ++k) {
            if (hipEventRecord(e0, st) != hipSuccess) return GSAP_ERR_HIP;
            r = run();
            if (r) return r;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                return GSAP_ERR_HIP;
            (void)hipEventElapsedTime(&v[k], e0, e1);
        }
        std::sort(v, v + reps);
        *ms = v[reps / 2];
        return GSAP_OK;
    };
    // every configuration measured, for the confirmation below
    std::vector<std::pair<float, gsap_tune>> seen;
    float best_ms = 0.0f;
    rc = measure(best, &best_ms);
    if (!rc) seen.push_back({best_ms, best});
    auto consider = [&](gsap_tune t) {
        if (rc) return;
        float ms = 0.0f;
        rc = measure(t, &ms);
        if (!rc) seen.push_back({ms, t});
        // with a caller workspace, a configuration that needs more than it
        // holds is not a candidate (the pair count is exact, so this is the
        // only source of a capacity failure here)
        if (rc == GSAP_ERR_CAPACITY && h->ws_base) {
            rc = GSAP_OK;
            return;
        }
        if (!rc && ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    };
    {  // stage 1: plain or slab sweep x sweep XCD chunking x pair regions
        const gsap_tune base = best;
        for (int sl = 0; sl < 2; ++sl)
            for (int ch : {0, 4, 16, 64})
                for (int rg : {1, kMaxRegions}) {
                    gsap_tune t = base;
                    t.slabs = sl;
                    t.sweep_chunk = ch;
                    t.pair_regions = rg;
                    if (t.slabs != base.slabs || ch != base.sweep_chunk || rg != base.pair_regions)
                        consider(t);
                }
    }
    {  // stage 2: object sort tile and digit width
        const gsap_tune base = best;
        for (int items : {4, 8, 12, 16})
            for (int bits : {6, 7, 8}) {
                gsap_tune t = base;
                t.sort_items = items;
                t.sort_bits = bits;
                if (items != base.sort_items || bits != base.sort_bits) consider(t);
            }
    }
    {  // stage 3: pair sort tile, digit width and grid; whole key or high bits + local stage
        const gsap_tune base = best;
        for (int loc : {0, 1})
            for (int items : {8, 16})
                for (int bits : {6, 7, 8})
                    for (int grid : {512, 1024, 2048, 4096}) {
                        if (loc && bits != base.pair_bits) continue;  // 8-bit digits with the local stage
                        gsap_tune t = base;
                        t.pair_local = loc;
                        t.pair_items = items;
                        t.pair_bits = bits;
                        t.pair_grid = grid;
                        if (loc != base.pair_local || items != base.pair_items ||
                            bits != base.pair_bits || grid != base.pair_grid)
                            consider(t);
                    }
        if (base.pair_local != 2) {  // by rows (digit parameters unused)
            gsap_tune t = base;
            t.pair_local = 2;
            consider(t);
        }
    }
    // Confirmation: the stages keep a candidate on a median of three, so
    // near-equal ones (within the update's ~1% spread) fall either way; the
    // four fastest seen are timed again, median of seven, and the fastest
    // of those is kept.
    if (!rc && seen.size() > 1) {
        std::stable_sort(seen.begin(), seen.end(),
                         [](const std::pair<float, gsap_tune>& x, const std::pair<float, gsap_tune>& y) {
                             return x.first < y.first;
                         });
        float cbest = 0.0f;
        gsap_tune ctune = best;
        const size_t top = std::min<size_t>(4, seen.size());
        for (size_t k = 0; k < top && !rc; ++k) {
            float ms = 0.0f;
            rc = measure(seen[k].second, &ms, 7);
            if (!rc && (k == 0 || ms < cbest)) {
                cbest = ms;
                ctune = seen[k].second;
            }
        }
        if (!rc) {
            best = ctune;
            best_ms = cbest;
        }
    }
    h->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_pairs);
    if (d_extra) (void)hipFree(d_extra);
    if (rc) {
        const std::string msg = "autotune: " + h->err;
        return fail(h, rc, msg);
    }
    best.ms = best_ms;
    apply_tune(h, best);
    h->tune_ms = best_ms;
    h->tune_cached = 0;
    if (cache_path) cache_append(cache_path, key, tune_of(h));
    *out = tune_of(h);
    return GSAP_OK;
}
