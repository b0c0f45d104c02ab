// k_coherent.hip -- the coherent object sort of GSAP_FLAG_COHERENT handles.
//
// Temporal coherence (the reference carries state between frames only in
// Sap::collision's interval, Sap.cpp:226-243).  Successive updates of one
// object set keep the key function (sweep axis, slab axis and slab / key
// parameters) of the last refresh, so the previous update's sorted keys stay
// comparable with this update's.  The sorted order is the stable LSD sort's:
// by the composite (key << 32 | entry position).
//
// Between two steps of a moving scene almost every entry stays within a few
// places of where it was (swaps with its neighbours), and a few hundred jump
// far (a box that enters or leaves a slab).  So:
//  * k_coh_block: block k of kCohB entries of the previous order computes
//    the entries' new composites and keeps those that stay in the block's
//    previous range [split[k], split[k + 1]) -- the previous composites at
//    positions k kCohB and (k + 1) kCohB.  Kept entries of different blocks
//    are then in block order.  Within the block the kept entries are nearly
//    sorted: an entry whose composite exceeds everything before it and
//    undercuts everything after it is in place; the others form short
//    out-of-order runs, each of which holds exactly its own entries, ranked
//    by comparisons within the run.  Entries that left the range go to a
//    side list.
//  * k_coh_far_sort: one workgroup sorts the side list in LDS, then finds
//    each block's first side-list entry and first final position (the kept
//    entries of the blocks before it plus the side-list entries below its
//    range);
//  * k_coh_place: block k's kept entries and the side-list entries of its
//    range, merged, go to the block's final positions.
// A side list longer than kCohFar, or an out-of-order run reaching more than
// kCohRun entries to either side of one of its entries, sets
// ctrl->coh_fallback, and the regular projection and LSD passes (gated on
// that flag) run instead.
#include "gsap_device.h"

namespace gsap {

namespace {

constexpr uint32_t kCohPer = kCohB / kWG;  // entries per thread of a block

// Entry value <-> entry position (the order the LSD path starts from): in slab
// mode box i has entries 2i (value i | kFirstBit) and 2i + 1 (value i).
__device__ __forceinline__ uint32_t entry_pos(uint32_t v, int slabs) {
    if (!slabs) return v;
    return (v & kFirstBit) ? 2u * (v & ~kFirstBit) : 2u * v + 1u;
}
__device__ __forceinline__ uint32_t entry_val(uint32_t pos, int slabs) {
    if (!slabs) return pos;
    return (pos & 1u) ? (pos >> 1) : ((pos >> 1) | kFirstBit);
}

// Key of entry v under the current key function: k_project's computation.
// *slab = the entry's slab, or 0xFFFFFFFF for an unused second entry and for
// plain keys.
__device__ __forceinline__ uint32_t entry_key(const CohArgs& a, uint32_t v, const float u[3],
                                              bool exact, uint32_t* slab) {
    const uint32_t box = a.slabs ? (v & ~kFirstBit) : v;
    float lo[3], hi[3], c4[4], pl, ph;
    load_lohi(a.in, a.kind, box, lo, hi, c4);
    project_lohi(a.kind, lo, hi, c4, u, exact, &pl, &ph);
    if (!a.slabs) {
        *slab = 0xFFFFFFFFu;
        return flip_f32(pl);
    }
    const float w[3] = {a.axis2[0], a.axis2[1], a.axis2[2]};
    float ql, qh;
    project_lohi(a.kind, lo, hi, c4, w, false, &ql, &qh);
    const uint32_t s0 = slab_of(ql, a.slab);
    uint32_t s1 = slab_of(qh, a.slab);
    if (s1 > s0 + 1u) {  // a box wider than a slab: fail loudly (as k_project)
        atomicOr(&a.ctrl->flags, kFlagSlab);
        s1 = s0 + 1u;
    }
    const uint32_t q = qkey(pl, a.slab);
    if (v & kFirstBit) {
        *slab = s0;
        return (s0 << kSlabQBits) | q;
    }
    if (s1 > s0) {
        *slab = s1;
        return (s1 << kSlabQBits) | q;
    }
    *slab = 0xFFFFFFFFu;
    return kSlabNull;
}

__device__ __forceinline__ unsigned long long composite(uint32_t key, uint32_t v, int slabs) {
    return ((unsigned long long)key << 32) | entry_pos(v, slabs);
}

__device__ __forceinline__ unsigned long long u64max(unsigned long long x, unsigned long long y) {
    return x > y ? x : y;
}
__device__ __forceinline__ unsigned long long u64min(unsigned long long x, unsigned long long y) {
    return x < y ? x : y;
}

// Exclusive prefix maximum / suffix minimum over the kWG threads' values
// (identities 0 and ~0); s_x, s_n: kWaves words each.
__device__ __forceinline__ void block_pmax_smin(unsigned long long mx, unsigned long long mn,
                                                unsigned long long* s_x, unsigned long long* s_n,
                                                unsigned long long* pmax, unsigned long long* smin) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    unsigned long long ix = mx, in = mn;  // inclusive within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long tx = __shfl_up(ix, o, 64);
        const unsigned long long tn = __shfl_down(in, o, 64);
        if (lane >= (uint32_t)o) ix = u64max(ix, tx);
        if (lane + (uint32_t)o < 64u) in = u64min(in, tn);
    }
    if (lane == 63u) s_x[w] = ix;
    if (lane == 0u) s_n[w] = in;
    __syncthreads();
    unsigned long long ex = __shfl_up(ix, 1, 64), en = __shfl_down(in, 1, 64);
    if (lane == 0u) ex = 0ull;
    if (lane == 63u) en = ~0ull;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
        if (k < (int)w) ex = u64max(ex, s_x[k]);
        if (k > (int)w) en = u64min(en, s_n[k]);
    }
    *pmax = ex;
    *smin = en;
}

// Bitonic sort of 256 composite keys held by one wave, 4 per lane (element
// r * 64 + lane in x[r]): distances of 64 and 128 are register pairs, the
// others cross-lane shuffles.  No LDS, no barriers.
__device__ __forceinline__ void wave_sort256(unsigned long long x[4], uint32_t lane) {
#pragma unroll
    for (uint32_t k = 2u; k <= 256u; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0u; j >>= 1) {
            if (j >= 64u) {
                const uint32_t rj = j >> 6;
#pragma unroll
                for (uint32_t r = 0; r < 4u; ++r) {
                    const uint32_t r2 = r ^ rj;
                    if (r2 > r) {
                        const bool asc = (((r << 6) | lane) & k) == 0u;
                        const unsigned long long a = x[r], b = x[r2];
                        const bool sw = asc ? (a > b) : (a < b);
                        x[r] = sw ? b : a;
                        x[r2] = sw ? a : b;
                    }
                }
            } else {
                const bool lower = (lane & j) == 0u;
#pragma unroll
                for (uint32_t r = 0; r < 4u; ++r) {
                    const bool asc = (((r << 6) | lane) & k) == 0u;
                    const unsigned long long y = __shfl_xor(x[r], (int)j, 64);
                    const bool keep_min = lower == asc;
                    x[r] = keep_min ? (x[r] < y ? x[r] : y) : (x[r] > y ? x[r] : y);
                }
            }
        }
    }
}

// first index in [b, e) of the sorted array d whose value is >= x
__device__ __forceinline__ uint32_t lower_bound64(const unsigned long long* d, uint32_t b, uint32_t e,
                                                  unsigned long long x) {
    while (b < e) {
        const uint32_t mid = b + ((e - b) >> 1);
        if (d[mid] < x) b = mid + 1u; else e = mid;
    }
    return b;
}

// Appends the entries c[q] of mask bits q to the side list (one claim per
// wave); past kCohFar entries the fallback is set instead.
__device__ __forceinline__ void far_append(const CohArgs& a, const unsigned long long c[kCohPer],
                                           uint32_t mask) {
    const uint32_t nf = (uint32_t)__popc(mask);
    const uint32_t inc = wave_incl_scan(nf);
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    uint32_t base = 0u;
    if (lane_id() == 0u && tot) base = atomicAdd(&a.ctrl->coh_far, tot);
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base) + inc - nf;
    if (base + nf > kCohFar) {
        if (nf) a.ctrl->coh_fallback = 1u;
        return;
    }
#pragma unroll
    for (uint32_t q = 0; q < kCohPer; ++q)
        if ((mask >> q) & 1u) a.far[base++] = c[q];
}

constexpr int kFarT = 1024;                  // threads of the side-list sort
constexpr uint32_t kFarQ = kCohFar / kFarT;  // keys per thread in its merges

}  // namespace

// ---------------------------------------------------------------- kernels
__global__ __launch_bounds__(kWG) void k_coh_block(CohArgs a) {
    __shared__ unsigned long long s_c[kCohB];  // kept composites, previous order
    __shared__ unsigned long long s_o[kCohB];  // the same, sorted
    __shared__ unsigned long long s_x[kWaves], s_n[kWaves];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_sc[256];             // entries per slab
    __shared__ uint8_t s_dirty[kCohB];
    __shared__ uint32_t s_bad;
    const uint32_t k = blockIdx.x, tid = threadIdx.x;
    const uint32_t p0 = k * kCohB;
    const uint32_t m = min(kCohB, a.ne - p0);
    const unsigned long long lo = k ? a.split[k] : 0ull;
    const unsigned long long hi = k + 1u < a.nb ? a.split[k + 1u] : ~0ull;
    s_sc[tid] = 0u;
    if (tid == 0) s_bad = 0u;
    const float u[3] = {a.axis[0], a.axis[1], a.axis[2]};
    const bool exact = axis_exact(a.kind, u);
    __syncthreads();

    // this thread's kCohPer consecutive entries of the previous order
    const uint32_t i0 = tid * kCohPer;
    uint32_t v[kCohPer];
    if (i0 + kCohPer <= m) {
        const uint4* src = reinterpret_cast<const uint4*>(a.valsF + p0 + i0);
#pragma unroll
        for (uint32_t q = 0; q < kCohPer / 4; ++q) {
            const uint4 t = src[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (uint32_t q = 0; q < kCohPer; ++q) v[q] = i0 + q < m ? a.valsF[p0 + i0 + q] : 0u;
    }
    unsigned long long c[kCohPer];
    uint32_t keepm = 0u, nk = 0u, run_slab = 0xFFFFFFFFu, run_n = 0u;
#pragma unroll
    for (uint32_t q = 0; q < kCohPer; ++q) {
        c[q] = ~0ull;
        if (i0 + q < m) {
            uint32_t slab;
            const uint32_t key = entry_key(a, v[q], u, exact, &slab);
            c[q] = composite(key, v[q], a.slabs);
            if (lo <= c[q] && c[q] < hi) {
                keepm |= 1u << q;
                ++nk;
            }
            // entries per slab, one LDS atomic per run of equal slabs
            if (slab != run_slab) {
                if (run_n && run_slab != 0xFFFFFFFFu) atomicAdd(&s_sc[run_slab], run_n);
                run_slab = slab;
                run_n = 0u;
            }
            ++run_n;
        }
    }
    if (run_n && run_slab != 0xFFFFFFFFu) atomicAdd(&s_sc[run_slab], run_n);

    // entries that left the block's range: the side list
    far_append(a, c, ~keepm & ((1u << (m > i0 ? min(kCohPer, m - i0) : 0u)) - 1u));

    // kept entries, compacted in order into s_c[0, M)
    uint32_t M;
    {
        const uint32_t ko = block_excl_scan(nk, s_tmp, &M);
        uint32_t o = ko;
#pragma unroll
        for (uint32_t q = 0; q < kCohPer; ++q)
            if ((keepm >> q) & 1u) s_c[o++] = c[q];
    }
    __syncthreads();
    if (s_sc[tid]) atomicAdd(&a.slab_cnt[tid], s_sc[tid]);  // slab mode only (else all zero)

    // An entry out of order with the one kCohDist places before or after it
    // moved that far within the block: the side list too (with, at most,
    // the one entry it is compared with), so the runs below stay short.
    unsigned long long x[kCohPer];
    {
        uint32_t mvm = 0u, nk2 = 0u;
#pragma unroll
        for (uint32_t q = 0; q < kCohPer; ++q) {
            const uint32_t i = i0 + q;
            x[q] = i < M ? s_c[i] : ~0ull;
            if (i < M) {
                const bool mv = (i + kCohDist < M && x[q] > s_c[i + kCohDist]) ||
                                (i >= kCohDist && s_c[i - kCohDist] > x[q]);
                mvm |= mv ? 1u << q : 0u;
                nk2 += mv ? 0u : 1u;
            }
        }
        far_append(a, x, mvm);
        const uint32_t M1 = M;
        const uint32_t ko = block_excl_scan(nk2, s_tmp, &M);
        uint32_t o = ko;
#pragma unroll
        for (uint32_t q = 0; q < kCohPer; ++q)
            if (i0 + q < M1 && !((mvm >> q) & 1u)) s_o[o++] = x[q];
    }
    __syncthreads();

    // in place: above everything before it, below everything after it
    unsigned long long mx = 0ull, mn = ~0ull;
#pragma unroll
    for (uint32_t q = 0; q < kCohPer; ++q) {
        const uint32_t i = i0 + q;
        x[q] = i < M ? s_o[i] : ~0ull;
        if (i < M) {
            mx = u64max(mx, x[q]);
            mn = u64min(mn, x[q]);
        }
    }
    unsigned long long pm, sm;
    block_pmax_smin(mx, mn, s_x, s_n, &pm, &sm);
    bool dirty[kCohPer];
    {
        unsigned long long run = pm;
#pragma unroll
        for (uint32_t q = 0; q < kCohPer; ++q) {
            dirty[q] = !(run < x[q]);
            run = u64max(run, x[q]);
        }
        run = sm;
#pragma unroll
        for (int q = (int)kCohPer - 1; q >= 0; --q) {
            if (i0 + (uint32_t)q < M) {
                dirty[q] = dirty[q] || !(x[q] < run);
                run = u64min(run, x[q]);
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kCohPer; ++q)
            if (i0 + q < M) s_dirty[i0 + q] = dirty[q] ? 1u : 0u;
    }
    __syncthreads();
    // an out-of-order run [b, e) holds exactly its own entries: rank within it
#pragma unroll
    for (uint32_t q = 0; q < kCohPer; ++q) {
        const uint32_t i = i0 + q;
        if (i >= M) continue;
        if (!dirty[q]) {
            s_c[i] = x[q];
            continue;
        }
        uint32_t b = i, e = i + 1u;  // the run, at most kCohRun entries on each side
        while (b > 0u && s_dirty[b - 1u] && i - b < kCohRun) --b;
        while (e < M && s_dirty[e] && e - i < kCohRun) ++e;
        if ((b > 0u && s_dirty[b - 1u]) || (e < M && s_dirty[e])) {
            s_bad = 1u;
            continue;
        }
        uint32_t r = b;
        for (uint32_t j = b; j < e; ++j) r += s_o[j] < x[q] ? 1u : 0u;
        s_c[r] = x[q];
    }
    __syncthreads();
    if (s_bad) {  // a long out-of-order run: the regular sort runs instead
        if (tid == 0) a.ctrl->coh_fallback = 1u;
        return;
    }
    for (uint32_t i = tid; i < M; i += kWG) {
        const unsigned long long y = s_c[i];
        a.keysX[p0 + i] = (uint32_t)(y >> 32);
        a.valsX[p0 + i] = entry_val((uint32_t)y, a.slabs);
    }
    if (tid == 0) a.kept[k] = M;
}

// The side list in LDS: each wave sorts 256-key chunks in registers, then
// sorted runs are merged pairwise (an element's place in the merged run is
// its index plus the count of smaller keys in the partner run).
__global__ __launch_bounds__(kFarT) void k_coh_far_sort(CohArgs a) {
    if (a.ctrl->coh_fallback) return;
    __shared__ unsigned long long s_a[kCohFar], s_b[kCohFar];
    __shared__ uint32_t s_w[kFarT / 64];
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const uint32_t m = a.ctrl->coh_far;  // <= kCohFar (else the fallback is set)
    for (uint32_t c0 = w * 256u; c0 < m; c0 += (kFarT / 64) * 256u) {
        unsigned long long x[4];
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
            const uint32_t i = c0 + r * 64u + lane;
            x[r] = i < m ? a.far[i] : ~0ull;
        }
        wave_sort256(x, lane);
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
            const uint32_t i = c0 + r * 64u + lane;
            if (i < m) s_a[i] = x[r];  // the padding sorts last
        }
    }
    __syncthreads();
    unsigned long long* src = s_a;
    unsigned long long* dst = s_b;
    for (uint32_t L = 256u; L < m; L <<= 1) {
        unsigned long long xv[kFarQ];
        uint32_t lo[kFarQ], hi[kFarQ], oth[kFarQ], dpos[kFarQ];
#pragma unroll
        for (uint32_t q = 0; q < kFarQ; ++q) {
            const uint32_t e = tid + q * kFarT;
            const uint32_t run = e / L, pos = e - run * L;
            const uint32_t other = (run ^ 1u) * L;  // start of the partner run
            xv[q] = e < m ? src[e] : 0ull;
            oth[q] = other;
            lo[q] = 0u;
            hi[q] = (e < m && other < m) ? min(L, m - other) : 0u;
            dpos[q] = (run & ~1u) * L + pos;
        }
        for (uint32_t step = 0; (1u << step) <= L; ++step) {
#pragma unroll
            for (uint32_t q = 0; q < kFarQ; ++q) {
                if (lo[q] < hi[q]) {
                    const uint32_t mid = (lo[q] + hi[q]) >> 1;
                    if (src[oth[q] + mid] < xv[q]) lo[q] = mid + 1u; else hi[q] = mid;
                }
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kFarQ; ++q)
            if (tid + q * kFarT < m) dst[dpos[q] + lo[q]] = xv[q];
        __syncthreads();
        unsigned long long* t = src;
        src = dst;
        dst = t;
    }
    for (uint32_t i = tid; i < m; i += kFarT) a.far[i] = src[i];
    // per block: its first side-list entry (the side-list entries below its
    // range) and its first final position (plus the kept entries before it)
    const uint32_t nb = a.nb, per = (nb + kFarT - 1) / kFarT, k0 = tid * per;
    uint32_t t = 0u;
    for (uint32_t k = k0; k < min(k0 + per, nb); ++k) t += a.kept[k];
    const uint32_t inc = wave_incl_scan(t);
    if (lane == 63u) s_w[w] = inc;
    __syncthreads();
    uint32_t run = inc - t;
    for (uint32_t j = 0; j < w; ++j) run += s_w[j];
    for (uint32_t k = k0; k < min(k0 + per, nb); ++k) {
        const uint32_t f = k ? lower_bound64(src, 0u, m, a.split[k]) : 0u;
        a.fstart[k] = f;
        a.pstart[k] = run + f;
        run += a.kept[k];
    }
    if (tid == 0) a.fstart[nb] = m;
}

__global__ __launch_bounds__(kWG) void k_coh_place(CohArgs a) {
    if (a.ctrl->coh_fallback) return;
    const uint32_t k = blockIdx.x, tid = threadIdx.x;
    const uint32_t p0 = k * kCohB;
    const uint32_t f0 = a.fstart[k], f1 = a.fstart[k + 1u];
    const uint32_t M = a.kept[k], base = a.pstart[k];
    for (uint32_t i = tid; i < M; i += kWG) {
        const uint32_t key = a.keysX[p0 + i], v = a.valsX[p0 + i];
        const uint32_t r = f0 < f1 ? lower_bound64(a.far, f0, f1, composite(key, v, a.slabs)) - f0 : 0u;
        const uint32_t o = base + i + r;  // < ne: the blocks' ranges tile [0, ne)
        if (o < a.ne) {
            a.keysF[o] = key;
            a.valsF[o] = v;
        }
    }
    for (uint32_t j = f0 + tid; j < f1; j += kWG) {
        const unsigned long long d = a.far[j];
        uint32_t b = 0u, e = M;  // kept entries below d
        while (b < e) {
            const uint32_t mid = b + ((e - b) >> 1);
            if (composite(a.keysX[p0 + mid], a.valsX[p0 + mid], a.slabs) < d) b = mid + 1u; else e = mid;
        }
        const uint32_t o = base + (j - f0) + b;
        if (o < a.ne) {
            a.keysF[o] = (uint32_t)(d >> 32);
            a.valsF[o] = entry_val((uint32_t)d, a.slabs);
        }
    }
}

__global__ void k_coh_save(const uint32_t* __restrict__ valsF, const uint32_t* __restrict__ keysF,
                           unsigned long long* split, uint32_t nb, int slabs) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 1u && k < nb) split[k] = composite(keysF[k * kCohB], valsF[k * kCohB], slabs);
}

hipError_t launch_coh_block(const CohArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_coh_block, dim3(a.nb), dim3(kWG), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_coh_merge(const CohArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_coh_far_sort, dim3(1), dim3(kFarT), 0, st, a);
    hipLaunchKernelGGL(k_coh_place, dim3(a.nb), dim3(kWG), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_coh_save(const CohArgs& a, hipStream_t st) {
    if (a.nb > 1u)
        hipLaunchKernelGGL(k_coh_save, dim3((a.nb + 255u) / 256u), dim3(256), 0, st, a.valsF,
                           a.keysF, a.split, a.nb, a.slabs);
    return hipGetLastError();
}

}  // namespace gsap
