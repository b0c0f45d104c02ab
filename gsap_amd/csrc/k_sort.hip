// k_sort.hip -- stable LSD radix sort for gfx950 ("onesweep" organisation).
//
// Replaces nvRadixSort::RadixSort (physicboxSap2.16_Fermi/sapDetection/
// radixsort.h:110-278, radixsort.cu:1116-1498).  The reference runs, per
// 4-bit digit, a block-local split sort, an offset pass, a CUDPP scan and a
// scatter (4 launches + scan per digit).  Here one histogram of every digit
// position is taken up front (fused into the projection kernel for the object
// sort), then each digit is ONE kernel:
//   tile load (striped, coalesced) -> per-tile digit counts (LDS atomics,
//   equal neighbours aggregated) published at once to a decoupled look-back
//   array -> wave64 ballot match ranking (stable), while predecessors publish
//   -> look-back for the exclusive prefix, four predecessors per load round
//   -> keys staged in LDS in digit order -> coalesced scatter.
// Tiles are claimed through an atomic counter so a workgroup only ever waits
// on tiles that were claimed before it (no deadlock with a persistent grid).
// Segments (batched scenes) restart the look-back chain at their first tile.
#include "gsap_device.h"

namespace gsap {

struct KU32 {
    using T = uint32_t;
    static constexpr bool kHasVals = true;
    __device__ static __forceinline__ uint32_t digit(T k, int shift, int, uint32_t mask) {
        return (k >> shift) & mask;
    }
};

// Packed pair {a, b} (a in the low word) ordered by the logical key (a << bbits) | b.
struct KPair {
    using T = uint2;
    static constexpr bool kHasVals = false;
    __device__ static __forceinline__ uint32_t digit(T k, int shift, int bbits, uint32_t mask) {
        unsigned long long L = ((unsigned long long)k.x << bbits) | (unsigned long long)k.y;
        return (uint32_t)(L >> shift) & mask;
    }
};

// Region of a logical position of the region concatenation (regions are
// tile-aligned, so a tile or a 64-key chunk lies in one region): the last
// region starting at or before `lo`.
__device__ __forceinline__ uint32_t region_of(const unsigned long long* start, uint32_t nregions,
                                              unsigned long long lo) {
    uint32_t r = 0;
    for (uint32_t q = 1; q < nregions; ++q)
        if (start[q] <= lo) r = q;
    return r;
}

__device__ __forceinline__ unsigned long long lb_word(uint32_t tag, bool inc, uint32_t v) {
    return ((unsigned long long)tag << 33) | (inc ? (1ull << 32) : 0ull) | (unsigned long long)v;
}

// A pair key of at most 40 bits without its 8-bit digit at `shift`: the bits
// above the digit move down to `shift`, the bits below stay (32 bits).
__device__ __forceinline__ uint32_t pair_mid32(uint2 k, int shift, int bbits) {
    const unsigned long long L = ((unsigned long long)k.x << bbits) | (unsigned long long)k.y;
    return (uint32_t)(((L >> (shift + 8)) << shift) | (L & ((1ull << shift) - 1ull)));
}

// Rank of a valid lane's item among the tile's items of digit d (in claim
// order; any order within a digit will do), counting them in s_cnt: a run of
// equal digits in neighbouring lanes claims its ranks with one LDS atomic by
// its first lane, the others take theirs from it, so the runs of equal keys
// the sweep writes do not serialise on one LDS address.  Valid lanes are a
// prefix of the wave; the whole wave calls it.
__device__ __forceinline__ uint32_t run_rank(uint32_t d, bool valid, uint32_t* s_cnt) {
    const uint32_t lane = lane_id();
    const uint32_t prev = (uint32_t)__shfl_up((int)d, 1u, 64);
    const bool head = valid && (lane == 0u || prev != d);
    const unsigned long long hm = ballot(head);
    const uint32_t nvalid = (uint32_t)__popcll(ballot(valid));
    const unsigned long long le = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);  // lanes <= mine
    uint32_t base = 0u;
    if (head) {
        const unsigned long long nxt = hm & ~le;
        const uint32_t end = nxt ? (uint32_t)__builtin_ctzll(nxt) : nvalid;
        base = atomicAdd(&s_cnt[d], end - lane);
    }
    const unsigned long long mine = hm & le;  // my run's head: the last head at or below me
    const uint32_t hl = mine ? 63u - (uint32_t)__builtin_clzll(mine) : 0u;
    return (uint32_t)__shfl((int)base, (int)hl, 64) + (lane - hl);
}

// The digit d whose range [st[d], st[d + 1]) of a pass's output holds
// position p (st: nondecreasing, st[0] = 0 <= p; the last such d).
__device__ __forceinline__ uint32_t digit_at(const uint32_t* st, uint32_t lo, uint32_t hi, uint32_t p) {
    while (lo < hi) {  // the last d in [lo, hi] with st[d] <= p
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (st[mid] <= p) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

template <class KT, int BITS, int ITEMS>
__global__ __launch_bounds__(kWG) void k_onesweep(const typename KT::T* __restrict__ kin,
                                                  typename KT::T* __restrict__ kout,
                                                  const uint32_t* __restrict__ vin,
                                                  uint32_t* __restrict__ vout, SortArgs a,
                                                  int pass) {
    constexpr uint32_t RADIX = 1u << BITS, MASK = RADIX - 1u, TILE = kWG * ITEMS;
    using K = typename KT::T;
    __shared__ uint32_t s_wh[kWaves][RADIX];
    __shared__ uint32_t s_tstart[RADIX];
    __shared__ uint32_t s_gbase[RADIX];
    __shared__ uint32_t s_tmp[2 * kWaves];
    __shared__ uint32_t s_cnt[RADIX];
    __shared__ K s_k[TILE];
    __shared__ uint32_t s_v[KT::kHasVals ? TILE : 1];
    __shared__ uint32_t s_tile;

    if (a.run_if && *a.run_if == 0u) return;  // a path the update did not take
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const int shift = a.shift0 + pass * BITS;
    uint32_t n = a.n;
    if (a.n_dev) {
        unsigned long long nd = (pass > 0 && a.n_dev_rest) ? *a.n_dev_rest : *a.n_dev;
        n = (uint32_t)(nd > a.cap ? a.cap : nd);
    }
    const uint32_t ntiles = a.seg.tile_seg ? a.ntiles : (n + TILE - 1) / TILE;
    const uint32_t ctr_idx = (uint32_t)(a.ctr_base + pass);
    const uint32_t tag = ((a.ctrl->epoch << 4) | ctr_idx) & 0x7FFFFFFFu;
    const unsigned long long ltm = lanemask_lt();
    const unsigned long long gtm = ~ltm & ~(1ull << lane);
    const unsigned long long above = lane == 63u ? 0ull : ~((2ull << lane) - 1ull);
    for (uint32_t i = tid; i < RADIX; i += kWG) s_cnt[i] = 0u;
    // No pass writes past the sorted range; the final pass of a pair sort
    // also stops at the caller's capacity.
    unsigned long long out_lim = n;
    if (a.last_n && (uint32_t)pass + 1u == a.npass) {
        const unsigned long long ln = *a.last_n;
        out_lim = ln < out_lim ? ln : out_lim;
        out_lim = a.last_cap < out_lim ? a.last_cap : out_lim;
    }

    for (;;) {
        if (tid == 0) s_tile = atomicAdd(&a.ctrl->tile_ctr[ctr_idx], 1u);
        for (uint32_t i = tid; i < kWaves * RADIX; i += kWG) (&s_wh[0][0])[i] = 0u;
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles) break;
        TileInfo ti = tile_info(a.seg, n, tile, TILE);
        const bool first = a.seg.tile_seg ? (tile == a.seg.seg_tile0[ti.seg]) : (tile == 0u);

        K key[ITEMS];
        uint32_t val[ITEMS], rank[ITEMS], dig[ITEMS];
        const uint32_t base = ti.lo + wave * 64u * ITEMS + lane;
        const K* src = kin;
        if (pass == 0 && a.nregions > 1) {
            // first pass of a pair sort: the sweep's regions, without their
            // padding (the tile's pairs are a prefix of it, so the valid items
            // stay a prefix of every wave's row)
            const uint32_t r = region_of(a.region_start, a.nregions, ti.lo);
            src = kin + ((unsigned long long)r * a.region_stride - a.region_start[r]);
            const unsigned long long ve = a.region_valid[r];
            if (ve < ti.hi) ti.hi = (uint32_t)ve;
        }
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t idx = base + 64u * k;
            if (idx < ti.hi) {
                key[k] = src[idx];
                if (KT::kHasVals) val[k] = vin[idx];
            }
        }
        // Tile digit counts first, so that the tile's aggregate is published
        // before the (much longer) stable ranking.  Valid lanes are a prefix
        // of each wave; a run of equal digits in neighbouring lanes is one
        // LDS atomic.
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const bool valid = (base + 64u * k) < ti.hi;
            const uint32_t d = valid ? KT::digit(key[k], shift, a.bbits, MASK) : 0u;
            dig[k] = d;
            const uint32_t prev = (uint32_t)__shfl_up((int)d, 1u, 64);
            const bool head = valid && (lane == 0u || prev != d);
            // both ballots in converged control flow (all lanes active)
            const unsigned long long hm = ballot(head);
            const uint32_t nvalid = (uint32_t)__popcll(ballot(valid));
            if (head) {
                const unsigned long long nxt = hm & above;
                const uint32_t end = nxt ? (uint32_t)__builtin_ctzll(nxt) : nvalid;
                atomicAdd(&s_cnt[d], end - lane);
            }
        }
        __syncthreads();
        uint32_t tcount = 0;
        if (tid < RADIX) {
            tcount = s_cnt[tid];
            s_cnt[tid] = 0u;
            st_rel(a.lookback + (size_t)tile * RADIX + tid, lb_word(tag, first, tcount));
        }
        // Stable ranking: items in tile order (wave, item, lane); peers with
        // the same digit found by BITS ballots.
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const bool valid = (base + 64u * k) < ti.hi;
            const uint32_t d = dig[k];
            unsigned long long m = ballot(valid);
#pragma unroll
            for (int b = 0; b < BITS; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long bb = ballot(bit);
                m &= bit ? bb : ~bb;
            }
            const uint32_t before = s_wh[wave][d];
            rank[k] = before + mbcnt(m);
            if (valid && (m & gtm) == 0ull) s_wh[wave][d] = before + (uint32_t)__popcll(m);
        }
        __syncthreads();

        uint32_t excl = 0;
        if (tid < RADIX) {
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = s_wh[w][tid];
                s_wh[w][tid] = run;
                run += c;
            }
            if (!first) {
                // Predecessors' words four at a time: sum the ready ones up to
                // the nearest inclusive prefix; stop at a not-yet-published one
                // and reload from there.  The scene's first tile is inclusive,
                // so the walk never leaves the scene.
                constexpr int WIN = 4;
                int p = (int)tile - 1;
                uint32_t spins = 0;
                for (;;) {
                    unsigned long long v[WIN];
#pragma unroll
                    for (int k = 0; k < WIN; ++k)
                        v[k] = p - k >= 0 ? ld_acq(a.lookback + (size_t)(p - k) * RADIX + tid) : 0ull;
                    uint32_t add = 0;
                    int used = 0;
                    bool done = false;
#pragma unroll
                    for (int k = 0; k < WIN; ++k) {
                        if (done || used != k || (uint32_t)(v[k] >> 33) != tag) continue;
                        add += (uint32_t)v[k];
                        used = k + 1;
                        done = (v[k] >> 32) & 1ull;
                    }
                    excl += add;
                    if (done) break;
                    if (used == 0) {
                        // Bounded wait: a predecessor publishes within microseconds;
                        // if it never does, flag the update and let the kernel end.
                        if (++spins > (1u << 22)) {
                            atomicOr(&a.ctrl->flags, kFlagTimeout);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    p -= used;
                }
                st_rel(a.lookback + (size_t)tile * RADIX + tid, lb_word(tag, true, excl + tcount));
            }
        }
        const uint32_t hv =
            tid < RADIX ? a.hist[((size_t)ti.seg * a.npass + (uint32_t)pass) * RADIX + tid] : 0u;
        uint32_t tot;
        const uint32_t gex = block_excl_scan(hv, s_tmp, &tot);
        const uint32_t tex = block_excl_scan(tcount, s_tmp + kWaves, &tot);
        if (tid < RADIX) {
            s_gbase[tid] = ti.seg_start + gex + excl;
            s_tstart[tid] = tex;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            if ((base + 64u * k) < ti.hi) {
                const uint32_t d = dig[k];
                const uint32_t lp = s_tstart[d] + s_wh[wave][d] + rank[k];
                s_k[lp] = key[k];
                if (KT::kHasVals) s_v[lp] = val[k];
            }
        }
        __syncthreads();
        const uint32_t tn = ti.hi - ti.lo;
        for (uint32_t i = tid; i < tn; i += kWG) {
            const K kk = s_k[i];
            const uint32_t d = KT::digit(kk, shift, a.bbits, MASK);
            const uint32_t g = s_gbase[d] + (i - s_tstart[d]);
            if (g < out_lim) {
                kout[g] = kk;
                if (KT::kHasVals) vout[g] = s_v[i];
            }
        }
        __syncthreads();
    }
}

// Sort pass by reduce-then-scan: no look-back chain.  With a few thousand
// tiles a decoupled look-back pass is the latency of its chain -- each tile
// reads its predecessors' words through the device-coherent level, four per
// round trip -- while the tiles' own work takes a fraction of that.  Here:
//  * k_rts_count: per tile digit counts -> C(d, t) = tab[d * ntm + t];
//  * k_rts_scan: one workgroup per digit d: the counts of d in the tiles
//    before t, over all tiles -> E(d, t) = tab[(RADIX + d) * ntm + t], and
//    d's total -> tab[2 RADIX ntm + d];
//  * k_rts_scatter: stable ranking within the tile (BITS ballots, items in
//    tile order), LDS staging in digit order, coalesced scatter at the
//    tile's offsets: its scene's start, plus the scene's digit start (the
//    scan of the scene's histogram, or of the totals for one scene), plus
//    E(d, t) - E(d, the scene's first tile).
// ntm: the tiles launched (the most the pass can have); with a device count
// the tiles past the pass's own return at once.  tab: 2 RADIX ntm + RADIX
// words (the pass's look-back buffer).  The first pass of a pair sort reads
// the sweep's regions without their padding, as k_onesweep does.
namespace {
// the count of a pass's keys and its tiles
__device__ __forceinline__ uint32_t rts_n(const SortArgs& a, int pass) {
    if (!a.n_dev) return a.n;
    const unsigned long long nd = (pass > 0 && a.n_dev_rest) ? *a.n_dev_rest : *a.n_dev;
    return (uint32_t)(nd > a.cap ? a.cap : nd);
}
template <uint32_t TILE>
__device__ __forceinline__ uint32_t rts_tiles(const SortArgs& a, uint32_t n) {
    return a.seg.tile_seg ? a.ntiles : (n + TILE - 1) / TILE;
}
// the tile's range; for the first pass of a pair sort, its region's keys
// (a prefix of the tile: the region's padding cut off)
template <class KT>
__device__ __forceinline__ const typename KT::T* rts_tile(const SortArgs& a, int pass, uint32_t n,
                                                          uint32_t tile, uint32_t size,
                                                          const typename KT::T* kin, TileInfo* ti) {
    *ti = tile_info(a.seg, n, tile, size);
    if (pass == 0 && a.nregions > 1) {
        const uint32_t r = region_of(a.region_start, a.nregions, ti->lo);
        const unsigned long long ve = a.region_valid[r];
        if (ve < ti->hi) ti->hi = ve > ti->lo ? (uint32_t)ve : ti->lo;
        return kin + ((unsigned long long)r * a.region_stride - a.region_start[r]);
    }
    return kin;
}
}  // namespace

template <class KT, int BITS, int ITEMS>
__global__ __launch_bounds__(kWG) void k_rts_count(const typename KT::T* __restrict__ kin, SortArgs a,
                                                   int pass, uint32_t ntm, uint32_t* __restrict__ tab) {
    constexpr uint32_t RADIX = 1u << BITS, MASK = RADIX - 1u, TILE = kWG * ITEMS;
    using K = typename KT::T;
    __shared__ uint32_t s_cnt[RADIX];
    if (a.run_if && *a.run_if == 0u) return;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const int shift = a.shift0 + pass * BITS;
    const uint32_t n = rts_n(a, pass), tile = blockIdx.x;
    if (tile >= rts_tiles<TILE>(a, n)) return;  // the whole workgroup
    const unsigned long long above = lane == 63u ? 0ull : ~((2ull << lane) - 1ull);
    for (uint32_t i = tid; i < RADIX; i += kWG) s_cnt[i] = 0u;
    __syncthreads();
    TileInfo ti;
    const K* src = rts_tile<KT>(a, pass, n, tile, TILE, kin, &ti);
    const uint32_t hi = ti.hi;
    const uint32_t base = ti.lo + wave * 64u * ITEMS + lane;
    K key[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t idx = base + 64u * k;
        key[k] = idx < hi ? src[idx] : K{};
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = (base + 64u * k) < hi;
        const uint32_t d = valid ? KT::digit(key[k], shift, a.bbits, MASK) : 0u;
        const uint32_t prev = (uint32_t)__shfl_up((int)d, 1u, 64);
        const bool head = valid && (lane == 0u || prev != d);
        const unsigned long long hm = ballot(head);
        const uint32_t nvalid = (uint32_t)__popcll(ballot(valid));
        if (head) {
            const unsigned long long nxt = hm & above;
            const uint32_t end = nxt ? (uint32_t)__builtin_ctzll(nxt) : nvalid;
            atomicAdd(&s_cnt[d], end - lane);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < RADIX; i += kWG) tab[i * ntm + tile] = s_cnt[i];
}

template <int BITS, int ITEMS>
__global__ __launch_bounds__(kWG) void k_rts_scan(SortArgs a, int pass, uint32_t ntm, uint32_t* __restrict__ tab) {
    constexpr uint32_t RADIX = 1u << BITS, TILE = kWG * ITEMS;
    __shared__ uint32_t s_tmp[kWaves];
    if (a.run_if && *a.run_if == 0u) return;
    const uint32_t tid = threadIdx.x, d = blockIdx.x;
    const uint32_t nt = rts_tiles<TILE>(a, rts_n(a, pass));
    uint32_t tot, run = 0u;
    const uint32_t* cnt = tab + (size_t)d * ntm;
    uint32_t* off = tab + (size_t)(RADIX + d) * ntm;
    for (uint32_t t0 = 0; t0 < nt; t0 += kWG) {  // uniform
        const uint32_t t = t0 + tid;
        const uint32_t v = t < nt ? cnt[t] : 0u;
        const uint32_t e = block_excl_scan(v, s_tmp, &tot);
        if (t < nt) off[t] = run + e;
        run += tot;
    }
    if (tid == 0) tab[(size_t)2 * RADIX * ntm + d] = run;
}

template <class KT, int BITS, int ITEMS>
__global__ __launch_bounds__(kWG) void k_rts_scatter(const typename KT::T* __restrict__ kin,
                                                     typename KT::T* __restrict__ kout,
                                                     const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                     SortArgs a, int pass, uint32_t ntm,
                                                     const uint32_t* __restrict__ tab) {
    constexpr uint32_t RADIX = 1u << BITS, MASK = RADIX - 1u, TILE = kWG * ITEMS;
    using K = typename KT::T;
    __shared__ uint32_t s_wh[kWaves][RADIX];
    __shared__ uint32_t s_tstart[RADIX];
    __shared__ uint32_t s_gbase[RADIX];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ K s_k[TILE];
    __shared__ uint32_t s_v[KT::kHasVals ? TILE : 1];
    if (a.run_if && *a.run_if == 0u) return;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const int shift = a.shift0 + pass * BITS;
    const uint32_t n = rts_n(a, pass), tile = blockIdx.x;
    if (tile >= rts_tiles<TILE>(a, n)) return;  // the whole workgroup
    const unsigned long long gtm = ~lanemask_lt() & ~(1ull << lane);
    for (uint32_t i = tid; i < kWaves * RADIX; i += kWG) (&s_wh[0][0])[i] = 0u;
    __syncthreads();
    TileInfo ti;
    const K* src = rts_tile<KT>(a, pass, n, tile, TILE, kin, &ti);
    const uint32_t lo = ti.lo, hi = ti.hi;
    const uint32_t t0 = a.seg.tile_seg ? a.seg.seg_tile0[ti.seg] : 0u;  // the scene's first tile
    // positions written: the scene's; a pair sort's last pass also stops at
    // the caller's capacity
    unsigned long long lim = ti.seg_end;
    if (a.last_n && (uint32_t)pass + 1u == a.npass) {
        const unsigned long long ln = *a.last_n;
        lim = ln < lim ? ln : lim;
        lim = a.last_cap < lim ? a.last_cap : lim;
    }
    const uint32_t base = lo + wave * 64u * ITEMS + lane;
    K key[ITEMS];
    uint32_t val[ITEMS], rank[ITEMS], dig[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t idx = base + 64u * k;
        key[k] = idx < hi ? src[idx] : K{};
        if (KT::kHasVals) val[k] = idx < hi ? vin[idx] : 0u;
    }
    // stable ranking: items in tile order (wave, item, lane); peers with the
    // same digit found by BITS ballots
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = (base + 64u * k) < hi;
        const uint32_t d = valid ? KT::digit(key[k], shift, a.bbits, MASK) : 0u;
        dig[k] = d;
        unsigned long long m = ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t before = s_wh[wave][d];
        rank[k] = before + mbcnt(m);
        if (valid && (m & gtm) == 0ull) s_wh[wave][d] = before + (uint32_t)__popcll(m);
    }
    __syncthreads();
    uint32_t tcount = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = s_wh[w][tid];
            s_wh[w][tid] = tcount;
            tcount += c;
        }
    }
    uint32_t tot;
    const uint32_t tex = block_excl_scan(tcount, s_tmp, &tot);
    // the scene's digit starts: its histogram, or (one scene) the totals
    uint32_t hv = 0u;
    if (tid < RADIX)
        hv = a.seg.tile_seg ? a.hist[((size_t)ti.seg * a.npass + (uint32_t)pass) * RADIX + tid]
                            : tab[(size_t)2 * RADIX * ntm + tid];
    const uint32_t gex = block_excl_scan(hv, s_tmp, &tot);
    if (tid < RADIX) {
        const uint32_t* E = tab + (size_t)(RADIX + tid) * ntm;
        s_gbase[tid] = ti.seg_start + gex + (E[tile] - E[t0]);
        s_tstart[tid] = tex;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if ((base + 64u * k) < hi) {
            const uint32_t d = dig[k];
            const uint32_t lp = s_tstart[d] + s_wh[wave][d] + rank[k];
            s_k[lp] = key[k];
            if (KT::kHasVals) s_v[lp] = val[k];
        }
    }
    __syncthreads();
    const uint32_t tn = hi > lo ? hi - lo : 0u;
    for (uint32_t i = tid; i < tn; i += kWG) {
        const K kk = s_k[i];
        const uint32_t d = KT::digit(kk, shift, a.bbits, MASK);
        const uint32_t g = s_gbase[d] + (i - s_tstart[d]);
        if (g < lim) {
            kout[g] = kk;
            if (KT::kHasVals) vout[g] = s_v[i];
        }
    }
}

// First pass of the pair sort without look-back.  An LSD sort's first pass
// need not be stable, so a tile does not wait for its predecessors: it counts
// its digits, claims its slots of each digit with one atomic per digit on the
// counter of (its pair region, digit), and scatters.  Digit d of region r
// occupies [start(d) + sum of d's counts in regions < r, + its count in r),
// from the per-region histograms the sweep counted as it wrote the pairs;
// regions are the sweep's per-XCD allocation units, so each counter is
// contended only by one region's tiles.  Region r's tiles go to the
// workgroups of XCD r (workgroup w runs on XCD w mod 8, so those with
// w mod nregions == r; the grid is a multiple of nregions): a digit's runs of
// one region then lie side by side in one L2, which merges the 128-byte lines
// that two neighbouring runs share before writing them back.  Lines written
// in parts from two L2s cost HBM about twice their bytes.  Items are staged
// in LDS in digit order (LDS-atomic ranks) for a coalesced scatter, never
// past the pair count.  MID32 (followed by k_xcount / k_xsplit, keys of at
// most 40 bits): the output position implies the digit, so each key is
// written as the 32 bits left without it (pair_mid32), and the bucket sizes
// of the second pass are zeroed here.
template <int BITS, int ITEMS, bool MID32>
__global__ __launch_bounds__(kWG) void k_msplit(const uint2* __restrict__ kin, uint2* __restrict__ kout,
                                                SortArgs a) {
    constexpr uint32_t RADIX = 1u << BITS, MASK = RADIX - 1u, TILE = kWG * ITEMS;
    static_assert(RADIX == kWG, "one digit per thread");
    __shared__ uint32_t s_cnt[RADIX];
    __shared__ uint32_t s_tstart[RADIX];
    __shared__ uint32_t s_gbase[RADIX];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint2 s_k[TILE];

    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const int shift = a.shift0;
    if constexpr (MID32)
        for (uint32_t i = blockIdx.x * kWG + tid; i < kRadixSq; i += gridDim.x * kWG) a.xcnt[i] = 0u;
    unsigned long long nd = *a.n_dev;
    const uint32_t n = (uint32_t)(nd > a.cap ? a.cap : nd);
    unsigned long long nr = *a.n_dev_rest;
    const uint32_t nreal = (uint32_t)(nr > a.cap ? a.cap : nr);
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    const unsigned long long above = lane == 63u ? 0ull : ~((2ull << lane) - 1ull);
    // digit tid's start, and its counts in the regions before each region
    uint32_t tot;
    const uint32_t dstart = block_excl_scan(a.hist[tid], s_tmp, &tot);
    uint32_t rpre[kMaxRegions];
    {
        uint32_t run = 0;
#pragma unroll
        for (int r = 0; r < kMaxRegions; ++r) {
            rpre[r] = run;
            if ((uint32_t)r < a.nregions) run += a.hist_reg[r * RADIX + tid];
        }
    }
    // this workgroup's region and its tiles (regions are tile-aligned)
    const uint32_t r = blockIdx.x % a.nregions;
    const uint32_t rt0 = (uint32_t)(a.region_start[r] / TILE);
    const uint32_t rt1 = min((uint32_t)((a.region_start[r + 1] + TILE - 1) / TILE), ntiles);
    const uint2* src = kin + ((unsigned long long)r * a.region_stride - a.region_start[r]);
    const unsigned long long ve = a.region_valid[r];
    const uint32_t gstep = gridDim.x / a.nregions;
    for (uint32_t tile = rt0 + blockIdx.x / a.nregions; tile < rt1; tile += gstep) {
        const uint32_t lo = tile * TILE;
        uint32_t hi = min(lo + TILE, n);
        if (ve < hi) hi = (uint32_t)ve;
        if (lo >= hi) continue;  // a tile of padding (uniform)
        s_cnt[tid] = 0u;
        __syncthreads();
        const uint32_t base = lo + wave * 64u * ITEMS + lane;
        uint2 key[ITEMS];
        uint32_t dig[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t idx = base + 64u * k;
            if (idx < hi) key[k] = src[idx];
        }
        // tile digit counts; a run of equal digits in neighbouring lanes is
        // one LDS atomic (valid lanes are a prefix of each row)
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const bool valid = (base + 64u * k) < hi;
            const uint32_t d = valid ? KPair::digit(key[k], shift, a.bbits, MASK) : 0u;
            dig[k] = d;
            const uint32_t prev = (uint32_t)__shfl_up((int)d, 1u, 64);
            const bool head = valid && (lane == 0u || prev != d);
            const unsigned long long hm = ballot(head);
            const uint32_t nvalid = (uint32_t)__popcll(ballot(valid));
            if (head) {
                const unsigned long long nxt = hm & above;
                const uint32_t end = nxt ? (uint32_t)__builtin_ctzll(nxt) : nvalid;
                atomicAdd(&s_cnt[d], end - lane);
            }
        }
        __syncthreads();
        const uint32_t c = s_cnt[tid];
        uint32_t g = 0;
        if (c) g = atomicAdd(&a.claim[r * RADIX + tid], c);  // read after the staging
        const uint32_t ts = block_excl_scan(c, s_tmp, &tot);
        s_tstart[tid] = ts;
        s_cnt[tid] = ts;  // running LDS position of the digit
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
            if ((base + 64u * k) < hi) s_k[atomicAdd(&s_cnt[dig[k]], 1u)] = key[k];
        s_gbase[tid] = dstart + rpre[r] + g;
        __syncthreads();
        const uint32_t tn = hi - lo;
        for (uint32_t i = tid; i < tn; i += kWG) {
            const uint2 kk = s_k[i];
            const uint32_t d = KPair::digit(kk, shift, a.bbits, MASK);
            const uint32_t gp = s_gbase[d] + (i - s_tstart[d]);
            if (gp < nreal) {
                if constexpr (MID32)
                    reinterpret_cast<uint32_t*>(kout)[gp] = pair_mid32(kk, shift, a.bbits);
                else
                    kout[gp] = kk;
            }
        }
        __syncthreads();
    }
}

// Two-pass pair sort, second pass without look-back.  The local stage
// orders the pairs inside each bucket of equal (d2, d1) -- d2 this pass's
// digit, d1 the first pass's -- so the pass need only put every key into its
// bucket, in any order within it:
//  * k_xcount: bucket sizes, from the first pass's 32-bit items (an item's
//    d1 is the first pass's digit range that holds its position): per chunk
//    of positions and d1, LDS digit counts, then one atomic per non-empty
//    bucket;
//  * k_xstarts: every bucket's start -- d2's start (the scan of the per-part
//    digit counts the sweep took) plus the sizes of the buckets (d2, d1')
//    with d1' < d1 -- for the local stage, and a copy as slot counters;
//  * k_xsplit: per tile and d1, each item's rank within its digit (one LDS
//    atomic per run of equal digits, run_rank), one slot claim per non-empty
//    bucket, the items staged in digit order with their final positions, a
//    coalesced scatter of each key's low 32 bits.
// Part x (first digits 32 x to 32 x + 31, a contiguous range of the first
// pass's output) runs on the workgroups of XCD x (w mod kMaxRegions == x, the
// grid a multiple of kMaxRegions), so a bucket's counters and its
// neighbouring runs stay in one L2.  Tables are [d2][d1] (bst[65536] = the
// key count); k_msplit zeroes the sizes.
namespace {
constexpr uint32_t kXRadix = 256u, kXParts = kMaxRegions, kXDpp = kXRadix / kXParts;

// the first pass's digit starts (clamped to n) into s_d1[0..256]
__device__ __forceinline__ void x_digit_starts(const SortArgs& a, uint32_t n, uint32_t* s_d1,
                                               uint32_t* s_tmp) {
    uint32_t tot;
    const uint32_t d1start = block_excl_scan(a.hist[threadIdx.x], s_tmp, &tot);
    s_d1[threadIdx.x] = min(d1start, n);
    if (threadIdx.x == 0) s_d1[kXRadix] = min(tot, n);
}

// Digit counts of items [sa, sb) (at most kWG * ITEMS, a wave's rows of 64
// consecutive items) into s_cnt; a run of equal digits in neighbouring lanes
// is one LDS atomic.  dig[k]: the digits (0 past sb).
template <int ITEMS>
__device__ __forceinline__ void x_count_rows(const uint32_t* __restrict__ kin, uint32_t sa, uint32_t sb,
                                             int sh, uint32_t* s_cnt, uint32_t item[ITEMS],
                                             uint32_t dig[ITEMS]) {
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long above = lane == 63u ? 0ull : ~((2ull << lane) - 1ull);
    const uint32_t base = sa + wave * 64u * ITEMS + lane;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t idx = base + 64u * k;
        item[k] = idx < sb ? kin[idx] : 0u;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = (base + 64u * k) < sb;
        const uint32_t d = valid ? (item[k] >> sh) & (kXRadix - 1u) : 0u;
        dig[k] = d;
        const uint32_t prev = (uint32_t)__shfl_up((int)d, 1u, 64);
        const bool head = valid && (lane == 0u || prev != d);
        const unsigned long long hm = ballot(head);
        const uint32_t nvalid = (uint32_t)__popcll(ballot(valid));
        if (head) {
            const unsigned long long nxt = hm & above;
            const uint32_t end = nxt ? (uint32_t)__builtin_ctzll(nxt) : nvalid;
            atomicAdd(&s_cnt[d], end - lane);
        }
    }
}

}  // namespace

template <int ITEMS>
__global__ __launch_bounds__(kWG) void k_xcount(const uint32_t* __restrict__ kin, SortArgs a) {
    constexpr uint32_t TILE = kWG * ITEMS, CHUNKS = 4;  // CHUNKS tiles per flush
    static_assert(kXRadix == kWG, "one digit per thread");
    __shared__ uint32_t s_d1[kXRadix + 1];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_cnt[kXRadix];
    const uint32_t tid = threadIdx.x;
    const unsigned long long nd = *a.n_dev_rest;
    const uint32_t n = (uint32_t)(nd > a.cap ? a.cap : nd);
    x_digit_starts(a, n, s_d1, s_tmp);
    s_cnt[tid] = 0u;
    __syncthreads();
    const uint32_t x = blockIdx.x % kXParts, gstep = gridDim.x / kXParts;
    const uint32_t plo = s_d1[x * kXDpp], phi = s_d1[(x + 1u) * kXDpp];
    const uint32_t dlast = x * kXDpp + kXDpp - 1u;
    for (unsigned long long lo64 = plo + (unsigned long long)(blockIdx.x / kXParts) * TILE * CHUNKS;
         lo64 < phi; lo64 += (unsigned long long)gstep * TILE * CHUNKS) {
        const uint32_t lo = (uint32_t)lo64;
        const uint32_t hi = (uint32_t)min((unsigned long long)phi, lo64 + TILE * CHUNKS);
        const uint32_t d1lo = digit_at(s_d1, x * kXDpp, dlast, lo);
        const uint32_t d1hi = digit_at(s_d1, d1lo, dlast, hi - 1u);
        for (uint32_t d1 = d1lo; d1 <= d1hi; ++d1) {  // uniform
            const uint32_t sa = max(lo, s_d1[d1]), sb = min(hi, s_d1[d1 + 1u]);
            if (sa >= sb) continue;
            for (uint32_t c0 = sa; c0 < sb; c0 += TILE) {
                uint32_t item[ITEMS], dig[ITEMS];
                x_count_rows<ITEMS>(kin, c0, min(sb, c0 + TILE), a.shift0, s_cnt, item, dig);
            }
            __syncthreads();
            const uint32_t c = s_cnt[tid];
            s_cnt[tid] = 0u;
            if (c) atomicAdd(&a.xcnt[tid * kXRadix + d1], c);
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kWG) void k_xstarts(SortArgs a) {
    static_assert(kXRadix == kWG, "one bucket per thread");
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_ds[kXRadix];
    const uint32_t tid = threadIdx.x, d2 = blockIdx.x;
    const unsigned long long nd = *a.n_dev_rest;
    const uint32_t n = (uint32_t)(nd > a.cap ? a.cap : nd);
    uint32_t all = 0, tot;
#pragma unroll
    for (uint32_t q = 0; q < kXParts; ++q) all += a.joint[q * kXRadix + tid];
    const uint32_t ds = block_excl_scan(all, s_tmp, &tot);
    s_ds[tid] = ds;
    if (d2 == 0) {  // the digit starts, for the readers of the output items
        a.st_out[tid] = ds;
        if (tid == 0) {
            a.st_out[kXRadix] = tot;
            a.bst[kXRadix * kXRadix] = n;
        }
    }
    __syncthreads();
    const uint32_t c = a.xcnt[d2 * kXRadix + tid];  // bucket (d2, d1 = tid)
    const uint32_t e = block_excl_scan(c, s_tmp, &tot);
    const uint32_t b = s_ds[d2] + e;
    a.bst[d2 * kXRadix + tid] = b;
    a.xcur[d2 * kXRadix + tid] = b;
}

template <int ITEMS>
__global__ __launch_bounds__(kWG) void k_xsplit(const uint32_t* __restrict__ kin, uint32_t* __restrict__ kout,
                                                SortArgs a) {
    constexpr uint32_t TILE = kWG * ITEMS;
    static_assert(kXRadix == kWG, "one digit per thread");
    __shared__ uint32_t s_d1[kXRadix + 1];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_cnt[kXRadix];
    __shared__ uint2 s_tg[kXRadix];  // per digit: tile start, first claimed slot
    __shared__ uint2 s_kp[TILE];     // (output item, position), in digit order
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    const int sh = a.shift0;
    const uint32_t lowm = (1u << sh) - 1u;
    const unsigned long long nd = *a.n_dev_rest;
    const uint32_t n = (uint32_t)(nd > a.cap ? a.cap : nd);
    x_digit_starts(a, n, s_d1, s_tmp);
    s_cnt[tid] = 0u;
    __syncthreads();
    const uint32_t x = blockIdx.x % kXParts, gstep = gridDim.x / kXParts;
    const uint32_t plo = s_d1[x * kXDpp], phi = s_d1[(x + 1u) * kXDpp];
    const uint32_t dlast = x * kXDpp + kXDpp - 1u;
    for (unsigned long long lo64 = plo + (unsigned long long)(blockIdx.x / kXParts) * TILE; lo64 < phi;
         lo64 += (unsigned long long)gstep * TILE) {
        const uint32_t lo = (uint32_t)lo64;
        const uint32_t hi = (uint32_t)min((unsigned long long)phi, lo64 + TILE);
        const uint32_t d1lo = digit_at(s_d1, x * kXDpp, dlast, lo);
        const uint32_t d1hi = digit_at(s_d1, d1lo, dlast, hi - 1u);
        for (uint32_t d1 = d1lo; d1 <= d1hi; ++d1) {  // uniform
            const uint32_t sa = max(lo, s_d1[d1]), sb = min(hi, s_d1[d1 + 1u]);
            if (sa >= sb) continue;
            // item k's digit (bits 16..23) and its rank (below; < TILE)
            uint32_t item[ITEMS], dr[ITEMS];
            const uint32_t base = sa + wave * 64u * ITEMS + lane;
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const uint32_t idx = base + 64u * k;
                item[k] = idx < sb ? kin[idx] : 0u;
            }
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const uint32_t d = (item[k] >> sh) & (kXRadix - 1u);
                dr[k] = (d << 16) | run_rank(d, (base + 64u * k) < sb, s_cnt);
            }
            __syncthreads();
            const uint32_t c = s_cnt[tid];
            uint32_t g = 0u, tot;
            if (c) g = atomicAdd(&a.xcur[tid * kXRadix + d1], c);
            const uint32_t ts = block_excl_scan(c, s_tmp, &tot);
            s_tg[tid] = make_uint2(ts, g);
            __syncthreads();
            // staged in digit order with their final positions
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                if ((base + 64u * k) < sb) {
                    const uint32_t d = dr[k] >> 16, rk = dr[k] & 0xFFFFu;
                    const uint2 tg = s_tg[d];
                    // the key's low 32 bits: d above d1 above the item's low bits
                    const unsigned long long L = ((unsigned long long)d << (sh + 8)) |
                                                 ((unsigned long long)d1 << sh) | (item[k] & lowm);
                    s_kp[tg.x + rk] = make_uint2((uint32_t)L, tg.y + rk);
                }
            }
            __syncthreads();
            const uint32_t tn = sb - sa;
            for (uint32_t i = tid; i < tn; i += kWG) {
                const uint2 kp = s_kp[i];
                if (kp.y < n) kout[kp.y] = kp.x;
            }
            s_cnt[tid] = 0u;
            __syncthreads();
        }
    }
}

// Histogram of every digit position (plain sorts and the pair sort).  Lanes
// hold consecutive keys; equal digits in neighbouring lanes are counted as one
// run (one LDS atomic per run), so the high digits of nearly sorted or narrow
// keys -- the same value across a whole wave -- do not serialise on one bin.
template <class KT, int BITS>
__global__ __launch_bounds__(kWG) void k_hist(const typename KT::T* __restrict__ keys, uint32_t n_host,
                                              const unsigned long long* __restrict__ n_dev,
                                              unsigned long long cap, uint32_t* __restrict__ hist,
                                              int npass, int bbits, int flip_float, int shift0,
                                              const unsigned long long* region_start,
                                              const unsigned long long* region_valid,
                                              unsigned long long region_stride,
                                              uint32_t nregions) {
    constexpr uint32_t RADIX = 1u << BITS, MASK = RADIX - 1u;
    __shared__ uint32_t s_h[kMaxPasses * RADIX];
    for (uint32_t i = threadIdx.x; i < (uint32_t)npass * RADIX; i += kWG) s_h[i] = 0u;
    __syncthreads();
    uint32_t n = n_host;
    if (n_dev) {
        unsigned long long nd = *n_dev;
        n = (uint32_t)(nd > cap ? cap : nd);
    }
    const uint32_t lane = lane_id();
    const unsigned long long above = lane == 63u ? 0ull : ~((2ull << lane) - 1ull);
    // Each wave takes UNROLL 64-key chunks per round (loads issued together,
    // so enough bytes are in flight to stream the keys at full bandwidth).
    constexpr uint32_t UNROLL = 4;
    const uint32_t stride = gridDim.x * kWG;
    for (uint32_t b0 = blockIdx.x * kWG + (threadIdx.x & ~63u); b0 < n; b0 += UNROLL * stride) {
        typename KT::T kk[UNROLL];
        uint32_t lim[UNROLL];  // end of the chunk's valid keys (a region's padding excluded)
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; ++u) {
            const uint32_t c0 = b0 + u * stride;  // no wrap: n + UNROLL * stride < 2^32 (n < 2^32 - 2^22)
            kk[u] = typename KT::T{};
            lim[u] = n;
            if (c0 < n) {
                const typename KT::T* src = keys;
                if (nregions > 1) {
                    const uint32_t r = region_of(region_start, nregions, c0);
                    src = keys + ((unsigned long long)r * region_stride - region_start[r]);
                    if (region_valid[r] < lim[u]) lim[u] = (uint32_t)region_valid[r];
                }
                if (c0 + lane < lim[u]) kk[u] = src[c0 + lane];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < UNROLL; ++u) {
            const uint32_t c0 = b0 + u * stride;
            if (c0 >= n) break;  // wave-uniform
            const bool valid = c0 + lane < lim[u];
            typename KT::T k = kk[u];
            if constexpr (KT::kHasVals) {
                if (flip_float) k = flip_f32(__uint_as_float(k));
            }
            const uint32_t nvalid = (uint32_t)__popcll(ballot(valid));
            for (int p = 0; p < npass; ++p) {
                const uint32_t d = KT::digit(k, shift0 + p * BITS, bbits, MASK);
                const uint32_t prev = (uint32_t)__shfl_up((int)d, 1u, 64);
                const bool head = valid && (lane == 0u || prev != d);
                const unsigned long long nxt = ballot(head) & above;
                if (head) {
                    const uint32_t end = nxt ? (uint32_t)__builtin_ctzll(nxt) : nvalid;
                    atomicAdd(&s_h[p * RADIX + d], end - lane);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < (uint32_t)npass * RADIX; i += kWG)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

__global__ void k_flip(uint32_t* keys, uint32_t n, int unflip) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint32_t k = keys[i];
        keys[i] = unflip ? __float_as_uint(unflip_f32(k)) : flip_f32(__uint_as_float(k));
    }
}

__global__ void k_sort_prep(Ctrl* ctrl, uint32_t* hist, uint32_t nhist, int ctr_base) {
    for (uint32_t i = threadIdx.x; i < nhist; i += blockDim.x) hist[i] = 0u;
    if (threadIdx.x < kMaxPasses) ctrl->tile_ctr[ctr_base + threadIdx.x] = 0u;
    if (threadIdx.x == 0) ctrl->epoch += 1u;
}

uint32_t sort_tile(int items) { return (uint32_t)(kWG * items); }

template <class KT, int BITS, int ITEMS>
static hipError_t run_passes(const SortArgs& a, typename KT::T* k0, uint32_t* v0,
                             typename KT::T* k1, uint32_t* v1, int* which, hipStream_t st,
                             typename KT::T* final_out = nullptr) {
    const uint32_t TILE = kWG * ITEMS;
    uint32_t grid = a.grid;
    if (!a.n_dev) {
        uint32_t nt = a.seg.tile_seg ? a.ntiles : (a.n + TILE - 1) / TILE;
        if (nt == 0) { *which = 0; return hipSuccess; }
        if (grid == 0 || grid > nt) grid = nt;
    }
    if (grid == 0) grid = 1;
    int cur = 0;
    for (uint32_t p = 0; p < a.npass; ++p) {
        typename KT::T* ki = cur ? k1 : k0;
        typename KT::T* ko = (final_out && p + 1 == a.npass) ? final_out : (cur ? k0 : k1);
        uint32_t* vi = cur ? v1 : v0;
        uint32_t* vo = cur ? v0 : v1;
        if constexpr (!KT::kHasVals && BITS == 8) {
            if (p == 0 && a.claim && a.npass > 1) {  // never the final pass: no capacity clamp
                // a whole number of workgroups per region; 32-bit items for
                // the per-part pass
                const uint32_t mg = (grid + a.nregions - 1) / a.nregions * a.nregions;
                if (a.joint && a.npass == 2)
                    hipLaunchKernelGGL((k_msplit<BITS, ITEMS, true>), dim3(mg), dim3(kWG), 0, st, ki, ko, a);
                else
                    hipLaunchKernelGGL((k_msplit<BITS, ITEMS, false>), dim3(mg), dim3(kWG), 0, st, ki, ko, a);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
                if (a.pass_ev) (void)hipEventRecord(a.pass_ev[p], st);
                cur ^= 1;
                continue;
            }
        }
        if constexpr (!KT::kHasVals && BITS == 8) {
            if (p == 1 && a.joint && a.npass == 2) {  // the last pass, per part
                const uint32_t xg = a.xgrid ? a.xgrid : grid;
                const uint32_t mg = (xg + kMaxRegions - 1) / kMaxRegions * kMaxRegions;
                if (!a.st_out || !a.bst || !a.xcnt || !a.xcur) return hipErrorInvalidValue;
                const uint32_t* xi = reinterpret_cast<const uint32_t*>(ki);
                hipLaunchKernelGGL((k_xcount<2 * ITEMS>), dim3(mg), dim3(kWG), 0, st, xi, a);
                hipLaunchKernelGGL(k_xstarts, dim3(kXRadix), dim3(kWG), 0, st, a);
                hipLaunchKernelGGL((k_xsplit<ITEMS>), dim3(mg), dim3(kWG), 0, st, xi,
                                   reinterpret_cast<uint32_t*>(ko), a);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return e;
                if (a.pass_ev) (void)hipEventRecord(a.pass_ev[p], st);
                cur ^= 1;
                continue;
            }
        }
        if (a.rts) {  // reduce-then-scan
            const uint32_t ntm = a.n_dev ? (uint32_t)((a.cap + TILE - 1) / TILE)
                                         : (a.seg.tile_seg ? a.ntiles : (a.n + TILE - 1) / TILE);
            if (ntm > 0) {
                uint32_t* tab = reinterpret_cast<uint32_t*>(a.lookback);
                hipLaunchKernelGGL((k_rts_count<KT, BITS, ITEMS>), dim3(ntm), dim3(kWG), 0, st, ki, a, (int)p,
                                   ntm, tab);
                hipLaunchKernelGGL((k_rts_scan<BITS, ITEMS>), dim3(1u << BITS), dim3(kWG), 0, st, a, (int)p, ntm,
                                   tab);
                hipLaunchKernelGGL((k_rts_scatter<KT, BITS, ITEMS>), dim3(ntm), dim3(kWG), 0, st, ki, ko, vi, vo,
                                   a, (int)p, ntm, tab);
            }
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            if (a.pass_ev) (void)hipEventRecord(a.pass_ev[p], st);
            cur ^= 1;
            continue;
        }
        hipLaunchKernelGGL((k_onesweep<KT, BITS, ITEMS>), dim3(grid), dim3(kWG), 0, st, ki, ko, vi, vo,
                           a, (int)p);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (a.pass_ev) (void)hipEventRecord(a.pass_ev[p], st);
        cur ^= 1;
    }
    *which = cur;
    return hipSuccess;
}

template <int BITS>
static hipError_t kv32_items(const SortArgs& a, uint32_t* k0, uint32_t* v0, uint32_t* k1,
                             uint32_t* v1, int* which, hipStream_t st) {
    switch (a.items) {
        case 4: return run_passes<KU32, BITS, 4>(a, k0, v0, k1, v1, which, st);
        case 8: return run_passes<KU32, BITS, 8>(a, k0, v0, k1, v1, which, st);
        case 12: return run_passes<KU32, BITS, 12>(a, k0, v0, k1, v1, which, st);
        case 16: return run_passes<KU32, BITS, 16>(a, k0, v0, k1, v1, which, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t sort_kv32(const SortArgs& a, uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                     int* which, hipStream_t st) {
    switch (a.bits) {
        case 6: return kv32_items<6>(a, k0, v0, k1, v1, which, st);
        case 7: return kv32_items<7>(a, k0, v0, k1, v1, which, st);
        case 8: return kv32_items<8>(a, k0, v0, k1, v1, which, st);
        default: return hipErrorInvalidValue;
    }
}

template <int BITS>
static hipError_t pair_items(const SortArgs& a, uint2* p0, uint2* p1, uint2* out, hipStream_t st) {
    int which = 0;
    switch (a.items) {
        case 8: return run_passes<KPair, BITS, 8>(a, p0, nullptr, p1, nullptr, &which, st, out);
        case 16: return run_passes<KPair, BITS, 16>(a, p0, nullptr, p1, nullptr, &which, st, out);
        default: return hipErrorInvalidValue;
    }
}

hipError_t sort_pairs(const SortArgs& a, uint2* p0, uint2* p1, uint2* out, hipStream_t st) {
    switch (a.bits) {
        case 6: return pair_items<6>(a, p0, p1, out, st);
        case 7: return pair_items<7>(a, p0, p1, out, st);
        case 8: return pair_items<8>(a, p0, p1, out, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_key_hist(const uint32_t* keys, uint32_t n, uint32_t* hist, int npass, int bits,
                           int flip_float, hipStream_t st) {
    uint32_t grid = (n + kWG * 16 - 1) / (kWG * 16);
    if (grid == 0) grid = 1;
    if (grid > 2048) grid = 2048;
    switch (bits) {
        case 6: hipLaunchKernelGGL((k_hist<KU32, 6>), dim3(grid), dim3(kWG), 0, st, keys, n, nullptr, 0ull, hist, npass, 0, flip_float, 0, nullptr, nullptr, 0ull, 1u); break;
        case 7: hipLaunchKernelGGL((k_hist<KU32, 7>), dim3(grid), dim3(kWG), 0, st, keys, n, nullptr, 0ull, hist, npass, 0, flip_float, 0, nullptr, nullptr, 0ull, 1u); break;
        case 8: hipLaunchKernelGGL((k_hist<KU32, 8>), dim3(grid), dim3(kWG), 0, st, keys, n, nullptr, 0ull, hist, npass, 0, flip_float, 0, nullptr, nullptr, 0ull, 1u); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pair_hist(const uint2* pairs, const unsigned long long* n_dev,
                            unsigned long long cap, uint32_t* hist, int npass, int bbits, int bits,
                            int shift0, const unsigned long long* rs,
                            const unsigned long long* rv, unsigned long long rstride,
                            uint32_t nreg, hipStream_t st) {
    unsigned long long g = (cap + kWG * 16 - 1) / (kWG * 16);
    uint32_t grid = (uint32_t)(g > 2048 ? 2048 : (g == 0 ? 1 : g));
    switch (bits) {
        case 6: hipLaunchKernelGGL((k_hist<KPair, 6>), dim3(grid), dim3(kWG), 0, st, pairs, 0u, n_dev, cap, hist, npass, bbits, 0, shift0, rs, rv, rstride, nreg); break;
        case 7: hipLaunchKernelGGL((k_hist<KPair, 7>), dim3(grid), dim3(kWG), 0, st, pairs, 0u, n_dev, cap, hist, npass, bbits, 0, shift0, rs, rv, rstride, nreg); break;
        case 8: hipLaunchKernelGGL((k_hist<KPair, 8>), dim3(grid), dim3(kWG), 0, st, pairs, 0u, n_dev, cap, hist, npass, bbits, 0, shift0, rs, rv, rstride, nreg); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_flip(uint32_t* keys, uint32_t n, int unflip, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_flip, dim3((n + 255) / 256), dim3(256), 0, st, keys, n, unflip);
    return hipGetLastError();
}

hipError_t launch_sort_prep(Ctrl* ctrl, uint32_t* hist, uint32_t nhist, int ctr_base,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_sort_prep, dim3(1), dim3(256), 0, st, ctrl, hist, nhist, ctr_base);
    return hipGetLastError();
}

}  // namespace gsap
