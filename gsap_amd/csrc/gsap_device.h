// gsap_device.h -- wave64 / CDNA device helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gsap_internal.h"

namespace gsap {

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = lane_id();
    return (l == 0) ? 0ull : (~0ull >> (64 - l));
}
// Lane mask of p.  The wave64 builtin takes the i1 directly, so a mask that
// is already in SGPRs (v_cmp + s_and) is used as is.
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// popcount of a 64-bit mask restricted to lanes below this one (v_mbcnt pair)
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Order-preserving float -> uint32 (flip the sign bit of positives, all bits
// of negatives).  -0 orders before +0.
__device__ __forceinline__ uint32_t flip_f32(float f) {
    uint32_t u = __float_as_uint(f);
    uint32_t mask = (uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u;
    return u ^ mask;
}
__device__ __forceinline__ float unflip_f32(uint32_t u) {
    uint32_t mask = ((u >> 31) - 1u) | 0x80000000u;
    return __uint_as_float(u ^ mask);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Inclusive scan of the wave's 64 lanes under an associative op with
// identity id, by DPP moves (VALU only; no LDS crossbar round trips):
// row_shr 1..3 of the inputs, row_shr 4 and 8 of the partial sums (upper
// banks), then row_bcast:15 and row_bcast:31 carry the row totals into the
// rows above.  Lanes a move does not write (and sources outside the row)
// take the identity.  All 64 lanes must be active.
template <int CTRL, int ROWS, int BANKS>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t id, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)src, CTRL, ROWS, BANKS, false);
}
template <class Op>
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t x, uint32_t id, Op op) {
    uint32_t v = op(x, dpp_mov<0x111, 0xF, 0xF>(id, x));  // row_shr:1
    v = op(v, dpp_mov<0x112, 0xF, 0xF>(id, x));           // row_shr:2
    v = op(v, dpp_mov<0x113, 0xF, 0xF>(id, x));           // row_shr:3
    v = op(v, dpp_mov<0x114, 0xF, 0xE>(id, v));           // row_shr:4, banks 1-3
    v = op(v, dpp_mov<0x118, 0xF, 0xC>(id, v));           // row_shr:8, banks 2-3
    v = op(v, dpp_mov<0x142, 0xA, 0xF>(id, v));           // row_bcast:15 into rows 1, 3
    v = op(v, dpp_mov<0x143, 0xC, 0xF>(id, v));           // row_bcast:31 into rows 2, 3
    return v;
}
// the wave's maximum, in every lane (all lanes active)
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    const uint32_t m = wave_scan_dpp(v, 0u, [](uint32_t p, uint32_t q) { return p > q ? p : q; });
    return (uint32_t)__builtin_amdgcn_readlane((int)m, 63);
}
// inclusive scan within the wave (all lanes active)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    return wave_scan_dpp(v, 0u, [](uint32_t p, uint32_t q) { return p + q; });
}

// Exclusive scan over a 256-thread workgroup; s_tmp needs kWaves words.
// Returns the exclusive prefix; *total receives the workgroup total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    uint32_t l = lane_id(), w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (l == 63) s_tmp[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
        uint32_t t = s_tmp[i];
        off += (i < (int)w) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// Agent-scope relaxed atomics used by the decoupled look-back.
__device__ __forceinline__ unsigned long long ld_acq(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup w runs on XCD
// w mod 8) and each XCD has its own L2.  xcd_block hands each XCD chunks of
// k consecutive logical blocks instead (chunks dealt round-robin, so uneven
// work still spreads over all XCDs), so neighbouring blocks -- which read
// overlapping candidate windows -- share one L2.  k == 0: identity.  A
// bijection on [0, g): the tail that does not fill a round of chunks keeps
// the identity mapping.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t w, uint32_t g, uint32_t k) {
    if (k == 0) return w;
    const uint32_t span = kXcds * k, full = (g / span) * span;
    if (w >= full) return w;
    const uint32_t x = w % kXcds, i = w / kXcds;  // the i-th workgroup on XCD x
    return ((i / k) * kXcds + x) * k + i % k;
}

// Scene of a tile and the tile's object range.
struct TileInfo {
    uint32_t seg, seg_start, seg_end, lo, hi;
};
__device__ __forceinline__ TileInfo tile_info(const SegTable& s, uint32_t n, uint32_t tile,
                                              uint32_t tile_size) {
    TileInfo t;
    if (s.tile_seg) {
        t.seg = s.tile_seg[tile];
        t.seg_start = s.seg_start[t.seg];
        t.seg_end = t.seg_start + s.seg_n[t.seg];
        t.lo = t.seg_start + (tile - s.seg_tile0[t.seg]) * tile_size;
    } else {
        t.seg = 0;
        t.seg_start = 0;
        t.seg_end = n;
        t.lo = tile * tile_size;
    }
    uint32_t h = t.lo + tile_size;
    t.hi = h < t.seg_end ? h : t.seg_end;
    return t;
}

// Conservative projection of one object onto axis u (identical evaluation
// order to the CPU oracle: products, ((t0 + t1) + t2), no contraction).
__device__ __forceinline__ void load_lohi(const void* in, int kind, uint32_t i, float lo[3],
                                          float hi[3], float c4[4]) {
    if (kind == 0) {
        const float2* b = reinterpret_cast<const float2*>(in) + 3 * (size_t)i;
        float2 p0 = b[0], p1 = b[1], p2 = b[2];
        lo[0] = p0.x; lo[1] = p0.y; lo[2] = p1.x;
        hi[0] = p1.y; hi[1] = p2.x; hi[2] = p2.y;
    } else {
        float4 c = reinterpret_cast<const float4*>(in)[i];
        c4[0] = c.x; c4[1] = c.y; c4[2] = c.z; c4[3] = c.w;
        lo[0] = c.x - c.w; lo[1] = c.y - c.w; lo[2] = c.z - c.w;
        hi[0] = c.x + c.w; hi[1] = c.y + c.w; hi[2] = c.z + c.w;
    }
}

__device__ __forceinline__ void project_lohi(int kind, const float lo[3], const float hi[3],
                                             const float c4[4], const float u[3], bool exact,
                                             float* plo, float* phi) {
#pragma clang fp contract(off)
    float a0 = u[0] * (u[0] >= 0.0f ? lo[0] : hi[0]);
    float a1 = u[1] * (u[1] >= 0.0f ? lo[1] : hi[1]);
    float a2 = u[2] * (u[2] >= 0.0f ? lo[2] : hi[2]);
    float b0 = u[0] * (u[0] >= 0.0f ? hi[0] : lo[0]);
    float b1 = u[1] * (u[1] >= 0.0f ? hi[1] : lo[1]);
    float b2 = u[2] * (u[2] >= 0.0f ? hi[2] : lo[2]);
    float l = (a0 + a1) + a2;
    float h = (b0 + b1) + b2;
    if (!exact) {
        float ml, mh;
        if (kind == 0) {
            ml = (fabsf(a0) + fabsf(a1)) + fabsf(a2);
            mh = (fabsf(b0) + fabsf(b1)) + fabsf(b2);
        } else {
            float r = c4[3];
            float m0 = fabsf(u[0]) * (fabsf(c4[0]) + r);
            float m1 = fabsf(u[1]) * (fabsf(c4[1]) + r);
            float m2 = fabsf(u[2]) * (fabsf(c4[2]) + r);
            ml = mh = (m0 + m1) + m2;
        }
        l = l - (ml * 0x1p-20f + 1.17549435e-38f);
        h = h + (mh * 0x1p-20f + 1.17549435e-38f);
    }
    *plo = l;
    *phi = h;
}

__device__ __forceinline__ bool axis_exact(int kind, const float u[3]) {
    int nz = 0;
    bool unit = false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (u[k] != 0.0f) { ++nz; unit = (u[k] == 1.0f || u[k] == -1.0f); }
    return kind == 0 && nz == 1 && unit;
}

// Slab-mode keys (slab parameters sp: v0, 1/W, S, u0, 1/du, W): the slab of
// a projection on the slab axis, and the kSlabQBits-bit quantisation of a
// lower bound on the sweep axis.
__device__ __forceinline__ uint32_t slab_of(float x, const float* sp) {
    const float q = floorf((x - sp[0]) * sp[1]);
    return (uint32_t)fminf(fmaxf(q, 0.0f), sp[2] - 1.0f);  // NaN -> 0
}
__device__ __forceinline__ uint32_t qkey(float x, const float* sp) {
    const float q = floorf((x - sp[3]) * sp[4]);
    return (uint32_t)fminf(fmaxf(q, 0.0f), kSlabQMax);
}

}  // namespace gsap
