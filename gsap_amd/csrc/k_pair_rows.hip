// k_pair_rows.hip -- the pair sort by rows: a counting sort on the first index.
//
// The canonical pair list is ordered by (first, second) index, the order of
// the reference's pair list (Sap.cpp:269-330, SapKernel.cu:2249-2360).  When
// the list is short next to the object count (a rigid-body particle step
// finds about one directed pair per particle), digit passes over the whole
// key are latency-bound: each pass is a chain of tile look-backs over a few
// megabytes.  A counting sort on the first index needs no chain:
//  * k_row_count: pairs per row, on n row counters (one atomic per row a
//    wave holds);
//  * k_row_sums: per block of kRowB rows, the pair total and the non-empty
//    rows;
//  * k_row_starts: each block's prefix (the earlier blocks' totals, summed in
//    parallel), then every row's start.  The symmetric output's row starts
//    (the first position of every non-empty row) come out of the same scan,
//    and the counters are zeroed again for the next update;
//  * k_row_scatter: each pair into its row, at slots a wave claims by one
//    atomic per row on the row's start (which leaves the row's end there:
//    row r then spans [start[r - 1], start[r]));
//  * k_row_sort: a pair of a row of at most kRowSmall pairs goes to its row's
//    start plus the number of the row's pairs with a smaller second index;
//    longer rows are listed for the local stage's big-window kernel (one
//    workgroup per row: an LDS sort, or LSD passes through global memory).
#include "gsap_device.h"

namespace gsap {

namespace {

constexpr uint32_t kRowB = 4096;             // rows per block of the row scan
constexpr uint32_t kRowPer = kRowB / kWG;    // consecutive rows per thread
constexpr uint32_t kRowSmall = 64;           // longest row ranked by comparisons
constexpr uint32_t kRowGrid = 2048;          // workgroups of the pair-streaming kernels

// Lanes of the wave that hold the same row x (valid lanes; all lanes of the
// wave call it together): the sweep writes a pair's both orientations and a
// box's pairs together, so a row recurs within a wave, and atomics on one
// address serialise.  Up to kMatchRounds groups are found by ballots (the
// lowest unmatched lane's row each round); lanes left after that form groups
// of their own.
constexpr int kMatchRounds = 16;
__device__ __forceinline__ unsigned long long row_group(uint32_t x, bool valid) {
    unsigned long long todo = ballot(valid);
    unsigned long long mine = 1ull << lane_id();
    for (int r = 0; r < kMatchRounds && todo; ++r) {
        const uint32_t l = (uint32_t)__builtin_ctzll(todo);
        const uint32_t xl = (uint32_t)__builtin_amdgcn_readlane((int)x, (int)l);
        const bool eq = valid && x == xl;
        const unsigned long long m = ballot(eq);
        if (eq) mine = m;
        todo &= ~m;
    }
    return mine;
}

// f(pair, valid) for every pair the sweep wrote, 64 positions per wave per
// iteration: region r holds region_valid[r] - region_start[r] pairs from
// position r * stride of P (one region: all of them from 0).  After a
// capacity overflow every region is empty.  Every lane of a wave runs the
// same iterations (the group search needs the whole wave).
template <class F>
__device__ __forceinline__ void for_each_pair_wave(const uint2* __restrict__ P, const Ctrl* ctrl,
                                                   uint32_t nreg, unsigned long long stride, F f) {
    const unsigned long long g = blockIdx.x * (unsigned long long)kWG + threadIdx.x;
    const unsigned long long gs = (unsigned long long)gridDim.x * kWG;
    const unsigned long long g0 = g - lane_id();  // the wave's first position
    for (uint32_t r = 0; r < nreg; ++r) {
        const unsigned long long m = ctrl->region_valid[r] - ctrl->region_start[r];
        const uint2* src = P + (unsigned long long)r * stride;
        for (unsigned long long k0 = g0; k0 < m; k0 += gs) {
            const unsigned long long k = k0 + lane_id();
            const bool valid = k < m;
            f(valid ? src[k] : make_uint2(0xFFFFFFFFu, 0u), valid);
        }
    }
}

__global__ __launch_bounds__(kWG) void k_row_count(const uint2* __restrict__ P, const Ctrl* ctrl,
                                                   uint32_t nreg, unsigned long long stride,
                                                   uint32_t n, uint32_t* __restrict__ cnt) {
    for_each_pair_wave(P, ctrl, nreg, stride, [&](uint2 p, bool valid) {
        const bool ok = valid && p.x < n;
        const unsigned long long grp = row_group(p.x, ok);
        // the group's lowest lane adds the group's size
        if (ok && (grp & lanemask_lt()) == 0ull) atomicAdd(&cnt[p.x], (uint32_t)__popcll(grp));
    });
}

// bsum[b] = pairs of rows [b kRowB, (b + 1) kRowB), bsum[nb + b] = non-empty rows
__global__ __launch_bounds__(kWG) void k_row_sums(const uint32_t* __restrict__ cnt, uint32_t n,
                                                  uint32_t* __restrict__ bsum, uint32_t nb) {
    __shared__ uint32_t s_a[kWaves], s_b[kWaves];
    const uint32_t b0 = blockIdx.x * kRowB, b1 = min(b0 + kRowB, n);
    uint32_t t = 0, ne = 0;
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += kWG) {
        const uint32_t c = cnt[i];
        t += c;
        ne += c != 0u ? 1u : 0u;
    }
    t = wave_sum(t);
    ne = wave_sum(ne);
    if (lane_id() == 0) {
        s_a[threadIdx.x >> 6] = t;
        s_b[threadIdx.x >> 6] = ne;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, c = 0;
        for (int w = 0; w < kWaves; ++w) {
            a += s_a[w];
            c += s_b[w];
        }
        bsum[blockIdx.x] = a;
        bsum[nb + blockIdx.x] = c;
    }
}

__global__ __launch_bounds__(kWG) void k_row_starts(uint32_t* __restrict__ cnt, uint32_t n,
                                                    const uint32_t* __restrict__ bsum, uint32_t nb,
                                                    uint32_t* __restrict__ rs,
                                                    int32_t* __restrict__ entries, Ctrl* ctrl) {
    __shared__ uint32_t s_tmp[kWaves], s_a[kWaves], s_b[kWaves];
    const uint32_t b = blockIdx.x, tid = threadIdx.x, w = tid >> 6;
    uint32_t pa = 0, pb = 0;
    for (uint32_t k = tid; k < b; k += kWG) {
        pa += bsum[k];
        pb += bsum[nb + k];
    }
    pa = wave_sum(pa);
    pb = wave_sum(pb);
    if (lane_id() == 0) {
        s_a[w] = pa;
        s_b[w] = pb;
    }
    __syncthreads();
    pa = 0;
    pb = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
        pa += s_a[k];
        pb += s_b[k];
    }
    const uint32_t r0 = b * kRowB + tid * kRowPer;
    uint32_t c[kRowPer], t = 0, ne = 0;
    if (r0 + kRowPer <= n) {
        const uint4* src = reinterpret_cast<const uint4*>(cnt + r0);
#pragma unroll
        for (uint32_t q = 0; q < kRowPer / 4; ++q) {
            const uint4 v = src[q];
            c[4 * q] = v.x;
            c[4 * q + 1] = v.y;
            c[4 * q + 2] = v.z;
            c[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kRowPer; ++k) c[k] = r0 + k < n ? cnt[r0 + k] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kRowPer; ++k) {
        t += c[k];
        ne += c[k] != 0u ? 1u : 0u;
    }
    uint32_t tot_a, tot_b;
    const uint32_t ea = block_excl_scan(t, s_tmp, &tot_a);
    const uint32_t eb = block_excl_scan(ne, s_tmp, &tot_b);
    uint32_t run = pa + ea, er = pb + eb;
    uint32_t st[kRowPer];
#pragma unroll
    for (uint32_t k = 0; k < kRowPer; ++k) {
        st[k] = run;
        if (c[k] != 0u && r0 + k < n) {
            cnt[r0 + k] = 0u;  // zero again for the next update
            if (entries) entries[er] = (int32_t)run;
            ++er;
        }
        run += c[k];
    }
    if (r0 + kRowPer <= n) {
        uint4* dst = reinterpret_cast<uint4*>(rs + r0);
#pragma unroll
        for (uint32_t q = 0; q < kRowPer / 4; ++q)
            dst[q] = make_uint4(st[4 * q], st[4 * q + 1], st[4 * q + 2], st[4 * q + 3]);
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kRowPer; ++k)
            if (r0 + k < n) rs[r0 + k] = st[k];
    }
    if (tid == 0) {
        if (b == nb - 1u && entries) ctrl->num_entries = pb + tot_b;
        if (b == 0) ctrl->local_big = 0u;  // counted by k_row_sort
    }
}

__global__ __launch_bounds__(kWG) void k_row_scatter(const uint2* __restrict__ P, const Ctrl* ctrl,
                                                     uint32_t nreg, unsigned long long stride,
                                                     uint32_t n, uint32_t* __restrict__ rs,
                                                     uint2* __restrict__ X) {
    for_each_pair_wave(P, ctrl, nreg, stride, [&](uint2 p, bool valid) {
        const bool ok = valid && p.x < n;
        const unsigned long long grp = row_group(p.x, ok);
        const unsigned long long below = grp & lanemask_lt();
        uint32_t base = 0u;
        if (ok && below == 0ull) base = atomicAdd(&rs[p.x], (uint32_t)__popcll(grp));
        // the group's slots follow its lowest lane's claim, in lane order
        const uint32_t lead = (uint32_t)__builtin_ctzll(grp);
        base = (uint32_t)__shfl((int)base, (int)lead, 64);
        if (ok) X[base + (uint32_t)__popcll(below)] = p;
    });
}

__global__ __launch_bounds__(kWG) void k_row_sort(const uint2* __restrict__ X, const Ctrl* ctrl,
                                                  unsigned long long cap, uint32_t nobj,
                                                  const uint32_t* __restrict__ rs,
                                                  uint2* __restrict__ out, uint32_t* __restrict__ wst,
                                                  uint32_t* __restrict__ big, uint32_t* big_ctr) {
    // the placed pairs: every row ends where the last one does (a pair with
    // an out-of-range first index, which the sweep never writes, was not
    // placed)
    const unsigned long long placed = rs[nobj - 1u];
    const unsigned long long n = min(ctrl->sort_real, placed);
    unsigned long long lim = n;
    if (ctrl->num_pairs < lim) lim = ctrl->num_pairs;
    if (cap < lim) lim = cap;
    const uint32_t* Xw = reinterpret_cast<const uint32_t*>(X);
    const unsigned long long gs = (unsigned long long)gridDim.x * kWG;
    for (unsigned long long p = blockIdx.x * (unsigned long long)kWG + threadIdx.x; p < n; p += gs) {
        const uint2 q = X[p];
        if (q.x >= nobj) continue;
        const uint32_t s = q.x ? rs[q.x - 1u] : 0u, e = rs[q.x];
        if (e - s <= kRowSmall) {
            uint32_t rank = 0;
            for (uint32_t k = s; k < e; ++k) rank += Xw[2 * (size_t)k + 1] < q.y ? 1u : 0u;
            if (s + rank < lim) out[s + rank] = q;
        } else if (p == s) {  // a long row: one entry of the big-window list
            const uint32_t k = atomicAdd(big_ctr, 1u);
            wst[2 * k] = s;
            wst[2 * k + 1] = e;
            big[k] = 2 * k;
        }
    }
}

}  // namespace

uint64_t pair_rows_table_words(uint64_t cap) { return 3 * (cap / (kRowSmall + 1) + 1) + 2; }

hipError_t launch_pair_rows(const uint2* P, uint2* X, uint2* Y, uint2* out, Ctrl* ctrl,
                            uint32_t nreg, unsigned long long stride,
                            unsigned long long sort_cap, unsigned long long cap, uint32_t n,
                            int bbits, uint32_t* cnt, uint32_t* rs, uint32_t* bsum,
                            int32_t* entries, uint32_t* table, hipStream_t st) {
    if (n == 0 || bbits < 8 || bbits > 24) return hipErrorInvalidValue;
    unsigned long long g = (sort_cap + kWG - 1) / kWG;
    const uint32_t grid = (uint32_t)(g > kRowGrid ? kRowGrid : (g == 0 ? 1 : g));
    const uint32_t nb = (n + kRowB - 1) / kRowB;
    const uint64_t K = cap / (kRowSmall + 1) + 1;  // rows longer than kRowSmall, at most
    uint32_t* const wst = table;
    uint32_t* const big = table + 2 * K;
    uint32_t* const big_ctr = &ctrl->local_big;
    hipLaunchKernelGGL(k_row_count, dim3(grid), dim3(kWG), 0, st, P, ctrl, nreg, stride, n, cnt);
    hipLaunchKernelGGL(k_row_sums, dim3(nb), dim3(kWG), 0, st, cnt, n, bsum, nb);
    hipLaunchKernelGGL(k_row_starts, dim3(nb), dim3(kWG), 0, st, cnt, n, bsum, nb, rs, entries, ctrl);
    hipLaunchKernelGGL(k_row_scatter, dim3(grid), dim3(kWG), 0, st, P, ctrl, nreg, stride, n, rs, X);
    hipLaunchKernelGGL(k_row_sort, dim3(grid), dim3(kWG), 0, st, X, ctrl, cap, n, rs, out, wst, big,
                       big_ctr);
    return launch_local_big_windows(X, Y, out, &ctrl->sort_real, &ctrl->num_pairs, cap, wst, big,
                                    big_ctr, bbits, bbits, st);
}

}  // namespace gsap
