// k_coherent.hip -- the coherent object sort of GSAP_FLAG_COHERENT handles.
//
// Temporal coherence (the reference carries state between frames only in
// Sap::collision's interval, Sap.cpp:226-243).  Successive updates of one
// object set keep the key function (sweep axis, slab axis and slab / key
// parameters) of the last refresh, so the previous update's sorted keys stay
// comparable with this update's.  Every kCohB-th sorted entry of the previous
// update is a splitter (composite key << 32 | entry position, sorted by
// construction); this update projects the entries in the previous sorted
// order (so a workgroup's entries fall into one or two buckets), counts the
// buckets, scans, scatters each entry into its bucket with one atomic claim
// per run of equal buckets in a wave, and sorts each bucket in LDS by the
// composite key -- the order of the stable LSD sort of the keys, which breaks
// ties by entry position.  The splitters only balance the buckets; any
// splitters give the sorted order.  A bucket of more than kCohCap entries
// sets ctrl->coh_fallback, and the regular projection and LSD passes (gated
// on that flag) run instead.
#include "gsap_device.h"

namespace gsap {

constexpr uint32_t kCohProjT = 4096u;   // positions per projection workgroup
constexpr uint32_t kCohWin = 32u;       // bucket counters kept in LDS per workgroup

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
// *slab = the entry's slab (kWG = 256 slabs), or 0xFFFFFFFF for an unused
// second entry and for plain keys.
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

// Bucket of composite c at previous position p: the number of splitters <= c.
// The guess p / kCohB is right for entries that kept their neighbourhood; a
// binary search settles the others.
__device__ __forceinline__ uint32_t bucket_of(const CohArgs& a, unsigned long long c, uint32_t p) {
    const uint32_t m = a.nb - 1u;  // splitters
    uint32_t g = min(p / kCohB, m);
    const bool lo_ok = g == 0u || a.split_key[g - 1u] <= c;
    const bool hi_ok = g == m || c < a.split_key[g];
    if (lo_ok && hi_ok) return g;
    uint32_t b = 0u, e = m;  // first j with split_key[j] > c
    while (b < e) {
        const uint32_t mid = b + ((e - b) >> 1);
        if (a.split_key[mid] <= c) b = mid + 1u; else e = mid;
    }
    return b;
}

__global__ void k_coh_reset(CohArgs a) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.nb; i += gridDim.x * blockDim.x)
        a.bcount[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl->coh_fallback = 0u;
}

// Entries in the previous sorted order: keys, bucket counts, entries per slab.
__global__ __launch_bounds__(kWG) void k_coh_project(CohArgs a) {
    __shared__ uint32_t s_b[kCohWin];
    __shared__ uint32_t s_sc[256];
    const uint32_t tid = threadIdx.x;
    const uint32_t p0 = blockIdx.x * kCohProjT;
    const uint32_t wb = p0 / kCohB > 8u ? p0 / kCohB - 8u : 0u;  // first bucket counted in LDS
    if (tid < kCohWin) s_b[tid] = 0u;
    s_sc[tid] = 0u;
    const float u[3] = {a.axis[0], a.axis[1], a.axis[2]};
    const bool exact = axis_exact(a.kind, u);
    __syncthreads();
    const uint32_t pe = min(p0 + kCohProjT, a.ne);
    for (uint32_t p = p0 + tid; p < pe; p += kWG) {
        const uint32_t v = a.prev_vals[p];
        uint32_t slab;
        const uint32_t key = entry_key(a, v, u, exact, &slab);
        a.keys0[p] = key;
        a.vals0[p] = v;
        const uint32_t b = bucket_of(a, composite(key, v, a.slabs), p);
        if (b - wb < kCohWin) atomicAdd(&s_b[b - wb], 1u);
        else atomicAdd(&a.bcount[b], 1u);  // an entry that moved far
        if (slab != 0xFFFFFFFFu) atomicAdd(&s_sc[slab], 1u);
    }
    __syncthreads();
    if (tid < kCohWin && s_b[tid] && wb + tid < a.nb) atomicAdd(&a.bcount[wb + tid], s_b[tid]);
    if (a.slabs && s_sc[tid]) atomicAdd(&a.slab_cnt[tid], s_sc[tid]);
}

// Exclusive scan of the bucket counts (one workgroup of 1024 threads, up to
// 8 buckets per thread); any bucket above kCohCap selects the fallback.
__global__ __launch_bounds__(1024) void k_coh_scan(CohArgs a) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_over;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    if (tid == 0) s_over = 0u;
    __syncthreads();
    uint32_t c[8], sum = 0u;
    bool over = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t i = tid * 8u + (uint32_t)k;
        c[k] = i < a.nb ? a.bcount[i] : 0u;
        over |= c[k] > kCohCap;
        sum += c[k];
    }
    // inclusive scan of the thread sums within the wave
    uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += t;
    }
    if (lane == 63u) s_w[w] = inc;
    if (over) atomicOr(&s_over, 1u);
    __syncthreads();
    uint32_t wpre = 0u;
    for (uint32_t k = 0; k < w; ++k) wpre += s_w[k];
    uint32_t run = wpre + inc - sum;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t i = tid * 8u + (uint32_t)k;
        if (i < a.nb) {
            a.bstart[i] = run;
            a.bcursor[i] = run;
        }
        run += c[k];
    }
    if (tid == 1023u) a.bstart[a.nb] = run;
    if (tid == 0 && s_over) a.ctrl->coh_fallback = 1u;
}

// Each entry to its bucket.  Lanes of a wave hold consecutive previous
// positions, so equal buckets come in runs: the first lane of a run claims
// the run's slots with one atomic.
__global__ __launch_bounds__(kWG) void k_coh_scatter(CohArgs a) {
    if (a.ctrl->coh_fallback) return;
    const uint32_t lane = lane_id();
    const uint32_t p = blockIdx.x * kWG + threadIdx.x;
    const bool valid = p < a.ne;
    uint32_t key = 0u, v = 0u, b = 0xFFFFFFFFu;
    if (valid) {
        key = a.keys0[p];
        v = a.vals0[p];
        b = bucket_of(a, composite(key, v, a.slabs), p);
    }
    const uint32_t prev = __shfl_up(b, 1, 64);
    const unsigned long long starts = ballot(valid && (lane == 0u || prev != b));
    const unsigned long long nvalid = ballot(valid);
    if (!valid) return;
    const unsigned long long le = starts & (lane == 63u ? ~0ull : ((2ull << lane) - 1ull));
    const uint32_t first = 63u - (uint32_t)__builtin_clzll(le);  // this run's first lane
    const unsigned long long after = starts & ~((2ull << lane) - 1ull) & (lane == 63u ? 0ull : ~0ull);
    const uint32_t last = after ? (uint32_t)__builtin_ctzll(after)
                                : 64u - (uint32_t)__builtin_clzll(nvalid);  // one past the run
    uint32_t base = 0u;
    if (lane == first) base = atomicAdd(&a.bcursor[b], last - first);
    base = __shfl(base, (int)first, 64);
    const uint32_t o = base + (lane - first);
    a.keys1[o] = key;
    a.vals1[o] = v;
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

// One bucket per workgroup: each wave sorts 256-key chunks in registers,
// then sorted runs are merged pairwise in LDS (an element's place in the
// merged run is its index plus the count of smaller keys in the partner
// run), and keys and values are written out in order.
__global__ __launch_bounds__(kWG) void k_coh_sort(CohArgs a) {
    if (a.ctrl->coh_fallback) return;
    __shared__ unsigned long long s_a[kCohCap], s_b[kCohCap];
    const uint32_t b = blockIdx.x, tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const uint32_t s0 = a.bstart[b], m = a.bstart[b + 1u] - s0;
    if (m == 0u) return;
    if (m > kCohCap) {  // the scan selects the fallback first; never sort out of bounds
        if (tid == 0) atomicOr(&a.ctrl->flags, kFlagCoherent);
        return;
    }
    for (uint32_t c0 = w * 256u; c0 < m; c0 += kWaves * 256u) {
        unsigned long long x[4];
#pragma unroll
        for (uint32_t r = 0; r < 4u; ++r) {
            const uint32_t i = c0 + r * 64u + lane;
            x[r] = i < m ? composite(a.keys1[s0 + i], a.vals1[s0 + i], a.slabs) : ~0ull;
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
    // Each thread's (up to kQ) elements search their partner runs in
    // lockstep, so their LDS loads overlap.
    constexpr uint32_t kQ = kCohCap / kWG;
    for (uint32_t L = 256u; L < m; L <<= 1) {
        unsigned long long xv[kQ];
        uint32_t lo[kQ], hi[kQ], oth[kQ], dpos[kQ];
#pragma unroll
        for (uint32_t q = 0; q < kQ; ++q) {
            const uint32_t e = tid + q * kWG;
            const uint32_t run = e / L, pos = e - run * L;
            const uint32_t other = (run ^ 1u) * L;  // start of the partner run
            xv[q] = e < m ? src[e] : 0ull;
            oth[q] = other;
            lo[q] = 0u;
            hi[q] = (e < m && other < m) ? min(L, m - other) : 0u;
            dpos[q] = (run & ~1u) * L + pos;  // start of the merged run + own index
        }
        // partner runs hold at most L <= kCohCap / 2 keys: log2 + 1 halvings
        for (uint32_t step = 0; (1u << step) <= L; ++step) {
#pragma unroll
            for (uint32_t q = 0; q < kQ; ++q) {
                if (lo[q] < hi[q]) {
                    const uint32_t mid = (lo[q] + hi[q]) >> 1;
                    if (src[oth[q] + mid] < xv[q]) lo[q] = mid + 1u; else hi[q] = mid;
                }
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kQ; ++q)
            if (tid + q * kWG < m) dst[dpos[q] + lo[q]] = xv[q];
        __syncthreads();
        unsigned long long* t = src;
        src = dst;
        dst = t;
    }
    for (uint32_t i = tid; i < m; i += kWG) {
        const unsigned long long c = src[i];
        a.keysF[s0 + i] = (uint32_t)(c >> 32);
        a.valsF[s0 + i] = entry_val((uint32_t)c, a.slabs);
    }
}

__global__ void k_coh_save(const uint32_t* __restrict__ valsF, const uint32_t* __restrict__ keysF,
                           unsigned long long* split_key, uint32_t nb, int slabs) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j + 1u < nb) {
        const uint32_t p = kCohB * (j + 1u);
        split_key[j] = composite(keysF[p], valsF[p], slabs);
    }
}

hipError_t launch_coh_project(const CohArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_coh_reset, dim3((a.nb + 1023u) / 1024u), dim3(1024), 0, st, a);
    const uint32_t g = (a.ne + kCohProjT - 1u) / kCohProjT;
    hipLaunchKernelGGL(k_coh_project, dim3(g ? g : 1u), dim3(kWG), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_coh_sort(const CohArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_coh_scan, dim3(1), dim3(1024), 0, st, a);
    hipLaunchKernelGGL(k_coh_scatter, dim3((a.ne + kWG - 1u) / kWG), dim3(kWG), 0, st, a);
    hipLaunchKernelGGL(k_coh_sort, dim3(a.nb), dim3(kWG), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_coh_save(const CohArgs& a, hipStream_t st) {
    if (a.nb > 1u)
        hipLaunchKernelGGL(k_coh_save, dim3((a.nb + 255u) / 256u), dim3(256), 0, st, a.valsF,
                           a.keysF, a.split_key, a.nb, a.slabs);
    return hipGetLastError();
}

}  // namespace gsap
