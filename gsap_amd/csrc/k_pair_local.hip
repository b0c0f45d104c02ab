// k_pair_local.hip -- local stage of the canonical pair sort.
//
// The canonical pair list is ordered by the logical key (a << bbits) | b, the
// order of the reference's pair list (Sap.cpp:269-330, SapKernel.cu:2249-2360).
// With this stage the onesweep passes sort only the key bits [L, 2 bbits); the
// pairs then lie in runs ("buckets") of equal key >> L, in bucket order.  Keys
// are unique, so what is left is to order every bucket by its low L bits,
// independently of every other bucket: work that needs no look-back and one
// trip through HBM instead of one per digit.
//  * k_pair_wstart: for each window j of kLocW positions, the first bucket
//    start at or after j * kLocW (a galloping 64-lane search over the sorted
//    pairs); wstart[J] = the pair count.
//  * k_pair_local: one workgroup per window sorts the buckets that start in
//    it, positions [wstart[j], wstart[j+1]).  Up to kLocC pairs whose high
//    bits lie within 2^(32-L) of the first one's: in LDS, on unique 32-bit
//    keys (high - first high) << L | low bits.  One counting pass on the 12
//    bits below the top bit of the window's largest possible key (its last
//    bucket with all low bits set) puts every key in its bin; a key's place
//    in the bin is the number of the bin's keys smaller than it.  Other windows (more than kLocC pairs, buckets too far apart
//    for a 32-bit key, or a bin of more than kBinMax keys) go to a list.
//  * k_pair_local_big: the listed windows, by stable LSD passes of wave64
//    ballot ranking through global memory, one chunk at a time,
//    ping-ponging over the same positions with the free sort buffer.  A
//    kernel of its own, so the common path's registers stay few.
#include "gsap_device.h"

namespace gsap {

namespace {

constexpr int kLT = 512;                // threads of a local-sort workgroup
constexpr int kLWaves = kLT / 64;
constexpr int kLocItems = kLocC / kLT;  // keys per thread of an LDS sort
constexpr uint32_t kBinBits = 12;       // counting pass of an LDS sort: 4096 bins
constexpr uint32_t kBins = 1u << kBinBits;
constexpr uint32_t kBinMax = 256;       // largest bin ranked by comparisons (each key walks its bin)
constexpr uint32_t kLocTail = 64;      // LDS words of the key array used for reductions
constexpr int kGlobItems = 8;           // keys per thread per chunk of the global path

// Exclusive scan over the LT threads of a workgroup (s_tmp: LT / 64 words).
// Without TAIL_SYNC it ends without the barrier that lets s_tmp be written
// again at once: for callers that pass another barrier first.
template <int LT, bool TAIL_SYNC = true>
__device__ __forceinline__ uint32_t lblock_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    constexpr int NW = LT / 64;
    const uint32_t l = lane_id(), w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (l == 63) s_tmp[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const uint32_t t = s_tmp[i];
        off += i < (int)w ? t : 0u;
        tot += t;
    }
    if constexpr (TAIL_SYNC) __syncthreads();
    *total = tot;
    return off + inc - v;
}

__device__ __forceinline__ unsigned long long pkey(uint2 p, int bbits) {
    return ((unsigned long long)p.x << bbits) | (unsigned long long)p.y;
}
__device__ __forceinline__ uint2 pkey_pair(unsigned long long k, int bbits) {
    return make_uint2((uint32_t)(k >> bbits), (uint32_t)(k & ((1ull << bbits) - 1ull)));
}

// Stable rank of a row of 64 keys among the earlier keys of this wave with
// the same digit: peers found by `bw` ballots, the highest lane of each peer
// group advances the wave's counter of the digit.
__device__ __forceinline__ uint32_t wave_rank(uint32_t d, bool valid, int bw, uint32_t* cnt,
                                              unsigned long long gtm) {
    unsigned long long m = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if (b < bw) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = ballot(bit);
            m &= bit ? bb : ~bb;
        }
    }
    const uint32_t before = cnt[d];
    if (valid && (m & gtm) == 0ull) cnt[d] = before + (uint32_t)__popcll(m);
    return before + mbcnt(m);
}

}  // namespace

// Pair key bits of the 32-bit local view.  khigh = key >> L and klow = key &
// (2^L - 1) for key = (a << bbits) | b, in 32-bit arithmetic; valid when the
// key has at most 32 + L bits (2 bbits - L <= 32, i.e. up to 2^28 objects).
__device__ __forceinline__ uint32_t khigh(uint2 p, int bbits, int L) {
    return L >= bbits ? (p.x >> (L - bbits)) : ((p.x << (bbits - L)) | (p.y >> L));
}
__device__ __forceinline__ uint32_t klow(uint2 p, int bbits, int L) {
    return L >= bbits ? (((p.x & ((1u << (L - bbits)) - 1u)) << bbits) | p.y)
                      : (p.y & ((1u << L) - 1u));
}
// The pair of key kbase + x: (key >> bbits, key's low bbits bits).  The key
// has at most 2 bbits <= 62 bits, so key >> bbits fits 32 bits and is the
// low word of the 64-bit shift (alignbit).
__device__ __forceinline__ uint2 kpair_at(unsigned long long kbase, uint32_t x, int bbits) {
    const unsigned long long k = kbase + x;
    return make_uint2(__builtin_amdgcn_alignbit((uint32_t)(k >> 32), (uint32_t)k, (uint32_t)bbits),
                      (uint32_t)k & ((1u << bbits) - 1u));
}

// 32-bit items: after a two-pass sort of keys of at most 40 bits whose
// second pass ran per part, an item is the key's low 32 bits.  The bits
// above are the top 8 - hb bits of the second digit d (hb = 40 - K: the
// digit's bits below bit 32), and d is known from the position: digit d
// occupies [st[d], st[d + 1]) (st, 257 words, written by that pass).
// load_st copies st into LDS (ends with a barrier).
template <int LT>
__device__ __forceinline__ void load_st(const uint32_t* __restrict__ st_g, uint32_t* st) {
    for (uint32_t i = threadIdx.x; i <= 256u; i += LT) st[i] = st_g[i];
    __syncthreads();
}
// The digit of position q < st[256]: the largest d with st[d] <= q.
__device__ __forceinline__ uint32_t d2_of(const uint32_t* st, uint32_t q) {
    uint32_t lo = 0, hi = 256;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (st[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}
// d2_of for a q that every lane of a full wave holds: lane l compares
// st[4l .. 4l + 3] with q (one 16-byte LDS read; st 16-byte aligned), and the
// digit is the number of d in [0, 256) with st[d] <= q, less one (st[0] = 0).
__device__ __forceinline__ uint32_t d2_of_wave(const uint32_t* st, uint32_t q) {
    const uint4 v = reinterpret_cast<const uint4*>(st)[lane_id()];
    return (uint32_t)(__builtin_popcountll(ballot(v.x <= q)) + __builtin_popcountll(ballot(v.y <= q)) +
                      __builtin_popcountll(ballot(v.z <= q)) + __builtin_popcountll(ballot(v.w <= q))) -
           1u;
}
// key >> L and key & (2^L - 1) of an item with second digit d (L <= 24).
__device__ __forceinline__ uint32_t ihigh(uint32_t v, uint32_t d, int hb, int L) {
    return ((d >> hb) << (32 - L)) | (v >> L);
}
__device__ __forceinline__ uint32_t ilow(uint32_t v, int L) { return v & ((1u << L) - 1u); }

// One item of the local stage: an 8-byte pair, or (X32) a 32-bit item with
// its second digit d.
template <bool X32>
struct LItem {
    using T = typename std::conditional<X32, uint32_t, uint2>::type;
    __device__ static __forceinline__ uint32_t high(T v, uint32_t d, int bbits, int hb, int L) {
        if constexpr (X32) return ihigh(v, d, hb, L);
        else return khigh(v, bbits, L);
    }
    __device__ static __forceinline__ uint32_t low(T v, int bbits, int L) {
        if constexpr (X32) return ilow(v, L);
        else return klow(v, bbits, L);
    }
};

// One wave per window boundary j in [0, J]: the smallest q >= j * kLocW with
// q == n or key[q] >> L != key[q-1] >> L.  The pairs are sorted by key >> L,
// so "key[q] >> L differs from key[t-1] >> L" is monotone in q >= t.  With
// 32-bit items (second digits from st) it holds from the end of key[t-1]'s
// digit bucket on, and before it when the items' bits from L up differ.
template <bool X32>
__global__ __launch_bounds__(kWG) void k_pair_wstart(const void* __restrict__ Pv,
                                                     const unsigned long long* __restrict__ n_dev,
                                                     int bbits, int L,
                                                     uint32_t* __restrict__ wstart,
                                                     uint32_t* __restrict__ big_ctr,
                                                     const uint32_t* __restrict__ st_g) {
    __shared__ __align__(16) uint32_t s_st[X32 ? 260 : 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) *big_ctr = 0u;  // read by the kernels after this one
    if constexpr (X32) load_st<kWG>(st_g, s_st);
    const uint2* const P = reinterpret_cast<const uint2*>(Pv);
    const uint32_t* const P32 = reinterpret_cast<const uint32_t*>(Pv);
    const uint32_t n = (uint32_t)*n_dev;
    const uint32_t J = (n + kLocW - 1) / kLocW;
    const uint32_t lane = lane_id();
    for (uint32_t j = blockIdx.x * kWaves + (threadIdx.x >> 6); j <= J;
         j += gridDim.x * kWaves) {
        const unsigned long long t = (unsigned long long)j * kLocW;
        unsigned long long res = t;
        if (t >= n) {
            res = n;
        } else if (t > 0) {
            unsigned long long h0 = 0, qend = n;
            if constexpr (X32) {
                const uint32_t d0 = d2_of_wave(s_st, (uint32_t)(t - 1));  // t: per wave
                if (s_st[d0 + 1] < qend) qend = s_st[d0 + 1];
                h0 = P32[t - 1] >> L;
            } else {
                h0 = pkey(P[t - 1], bbits) >> L;
            }
            auto differs = [&](unsigned long long q) {  // q < n
                if constexpr (X32) return q >= qend || (unsigned long long)(P32[q] >> L) != h0;
                else return (pkey(P[q], bbits) >> L) != h0;
            };
            // gallop: probes lo + lane * step, the step growing 64x per round
            unsigned long long lo = t, hi = n, step = 1;
            for (;;) {
                const unsigned long long q = lo + (unsigned long long)lane * step;
                const bool pr = q >= n || differs(q);
                const unsigned long long m = ballot(pr);
                if (m) {
                    const uint32_t i = (uint32_t)__builtin_ctzll(m);
                    if (i == 0) {
                        hi = lo;
                    } else {
                        hi = lo + (unsigned long long)i * step;
                        if (hi > n) hi = n;
                        lo = lo + (unsigned long long)(i - 1) * step + 1;
                    }
                    break;
                }
                lo += 63ull * step + 1;
                step *= 64;
            }
            // refine: the answer lies in [lo, hi] and hi satisfies the predicate
            while (lo < hi) {
                const unsigned long long stride = (hi - lo + 63) / 64;
                const unsigned long long q = lo + (unsigned long long)lane * stride;
                const bool pr = q >= hi || differs(q);
                const unsigned long long m = ballot(pr);
                if (m == 0) {
                    lo += 63ull * stride + 1;
                    continue;
                }
                const uint32_t i = (uint32_t)__builtin_ctzll(m);
                if (i == 0) {
                    hi = lo;
                    break;
                }
                const unsigned long long nhi = lo + (unsigned long long)i * stride;
                lo = lo + (unsigned long long)(i - 1) * stride + 1;
                if (nhi < hi) hi = nhi;
            }
            res = hi;
        }
        if (lane == 0) wstart[j] = (uint32_t)res;
    }
}

// Counting sort of one window in LDS by LT threads, ITEMS keys each (kk,
// the first nk rows valid, S keys in all; s_e holds at least S words).  The
// keys are unique 32-bit views (high - hs) << L | low bits, none above khi.
// One counting pass on the kBinBits bits below khi's top bit (the bits above
// are zero in every key, so the bins are in key order) puts every key in its
// bin; a key's place in its bin is the number of the bin's
// smaller keys (bins are uneven -- keys cluster by bucket and index -- so the
// work is spread over the keys rather than over the bins; the pairs, and so
// the keys, are unique, so no tie-break is needed).  Writes the pairs to
// out[s + rank] (below out_lim) and returns true, or returns false without
// writing when a bin holds more than kBinMax keys.  s_bin holds kBins + 1
// words: counts in [0, kBins), then bin b's start in word b + 1 (word 0 is
// 0), so after the scatter bin b is [s_bin[b], s_bin[b + 1]).  Uniform
// control flow; ends with a barrier, so the LDS may be reused at once.
template <int LT, int ITEMS>
__device__ __forceinline__ bool bin_sort(const uint32_t (&kk)[ITEMS], uint32_t nk, uint32_t S,
                                         uint32_t khi, uint32_t s, uint32_t hs,
                                         unsigned long long out_lim,
                                         uint2* __restrict__ out, uint32_t* s_e, uint32_t* s_bin,
                                         uint32_t* s_tmp, uint32_t* s_red, int bbits, int L) {
    constexpr uint32_t BPT = kBins / LT;  // bins per thread
    const uint32_t tid = threadIdx.x;
    const unsigned long long kbase = (unsigned long long)hs << L;
    for (uint32_t i = tid; i < kBins; i += LT) s_bin[i] = 0u;
    if (tid == 0) s_red[0] = 0u;  // set by the threads that see a bin over kBinMax
    // bins: the kBinBits bits below khi's top bit hb (x >> sd < kBins)
    const uint32_t hb = khi ? 31u - (uint32_t)__clz(khi) : 0u;
    const uint32_t sd = hb >= kBinBits - 1u ? hb - (kBinBits - 1u) : 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t q = tid + (uint32_t)LT * k;
        if ((uint32_t)k < nk && q < S) atomicAdd(&s_bin[(kk[k] >> sd) & (kBins - 1u)], 1u);
    }
    __syncthreads();
    // bin starts (exclusive scan, BPT bins per thread) and the largest bin
    uint32_t cb[BPT], sum = 0, mx = 0, all;
#pragma unroll
    for (uint32_t i = 0; i < BPT; ++i) {
        cb[i] = s_bin[tid * BPT + i];
        sum += cb[i];
        mx = cb[i] > mx ? cb[i] : mx;
    }
    if (mx > kBinMax) s_red[0] = 1u;
    // its first barrier orders every count read (and flag write) before the
    // writes below; s_tmp is next written after several more barriers
    uint32_t run = lblock_excl_scan<LT, false>(sum, s_tmp, &all);
#pragma unroll
    for (uint32_t i = 0; i < BPT; ++i) {
        s_bin[tid * BPT + i + 1u] = run;
        run += cb[i];
    }
    if (tid == 0) s_bin[0] = 0u;
    const bool over = s_red[0] != 0u;  // read before the barrier: the next call clears it after
    __syncthreads();
    if (over) return false;  // uniform; s_e and s_bin are free again
    // keys to their bins (any order inside a bin); a bin's counter ends at
    // the next bin's start
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t q = tid + (uint32_t)LT * k;
        if ((uint32_t)k < nk && q < S)
            s_e[atomicAdd(&s_bin[((kk[k] >> sd) & (kBins - 1u)) + 1u], 1u)] = kk[k];
    }
    // places below rlim are written (out_lim - s, clamped)
    const uint32_t rlim = out_lim <= s ? 0u
                        : out_lim - s >= 0xffffffffull ? 0xffffffffu : (uint32_t)(out_lim - s);
    __syncthreads();
    for (uint32_t pz = tid; pz < S; pz += LT) {
        const uint32_t x = s_e[pz];
        const uint32_t bn = (x >> sd) & (kBins - 1u);
        const uint32_t bs = s_bin[bn], be = s_bin[bn + 1u];
        // two keys per step; the word after the bin's last (i + 1 <= S <
        // the key array's size) is read but not counted
        uint32_t rank = 0;
        for (uint32_t i = bs; i < be; i += 2u) {
            const uint32_t y0 = s_e[i], y1 = s_e[i + 1u];
            rank += (y0 < x ? 1u : 0u) + (i + 1u < be && y1 < x ? 1u : 0u);
        }
        // the key is (hs + (x >> L)) << L | (x's low L bits) = (hs << L) + x
        if (bs + rank < rlim) out[s + bs + rank] = kpair_at(kbase, x, bbits);
    }
    __syncthreads();
    return true;
}

// LDS path: one workgroup per window j (grid-stride), sorting the buckets that
// start in it, positions [wstart[j], wstart[j+1]), by bin_sort: up to
// kLocC - kLocTail pairs whose high bits lie within 2^(32-L) of the first
// one's.  Windows that do not fit (too many pairs, buckets too far apart, a
// bin of more than kBinMax keys) are appended to `big` for k_pair_local_big.
// Software pipeline over the workgroup's windows j, j + G, ...: window starts
// are read two windows ahead, and the next window's pairs are loaded into
// registers while this window is sorted.
template <bool X32>
__global__ __launch_bounds__(kLT) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_pair_local(
    const void* __restrict__ Xv, uint2* __restrict__ out, const unsigned long long* __restrict__ n_dev,
    const unsigned long long* __restrict__ last_n, unsigned long long last_cap,
    const uint32_t* __restrict__ wstart, uint32_t* __restrict__ big, uint32_t* __restrict__ big_ctr,
    int bbits, int L, const uint32_t* __restrict__ st_g, int hb) {
    using IT = LItem<X32>;
    using XT = typename IT::T;
    const XT* __restrict__ X = reinterpret_cast<const XT*>(Xv);
    // 48 KB (+1 KB with 32-bit items), so that three workgroups fit a CU: the
    // reduction words live in the last kLocTail words of the key array, which
    // the keys never reach and which are dead while the keys are scattered
    // and ranked.
    __shared__ uint32_t s_e[kLocC];
    __shared__ uint32_t s_bin[kBins + 1];
    __shared__ __align__(16) uint32_t s_st[X32 ? 260 : 1];
    uint32_t* const s_tmp = s_e + (kLocC - kLocTail);
    uint32_t* const s_red = s_tmp + kLWaves;
    constexpr uint32_t kCap = kLocC - kLocTail;

    const uint32_t tid = threadIdx.x;
    if constexpr (X32) load_st<kLT>(st_g, s_st);
    const uint32_t n = (uint32_t)*n_dev;
    unsigned long long out_lim = n;
    {
        const unsigned long long ln = *last_n;
        if (ln < out_lim) out_lim = ln;
        if (last_cap < out_lim) out_lim = last_cap;
    }
    const uint32_t J = (n + kLocW - 1) / kLocW;
    const bool fits32 = 2 * bbits - L <= 32;
    const uint32_t G = gridDim.x;

    uint32_t j = blockIdx.x, s = 0, e = 0, s1 = 0, e1 = 0;
    if (j < J) {
        s = wstart[j];
        e = wstart[j + 1];
    }
    if (j + G < J) {
        s1 = wstart[j + G];
        e1 = wstart[j + G + 1];
    }
    // The prefetch holds the first kPre rows of a window (4,096 pairs, more
    // than most windows have); larger windows load the rest when sorted.
    constexpr int kPre = kLocItems / 2;
    XT raw[kPre], rfirst{}, rlast{};
    auto fetch = [&](uint32_t fs, uint32_t fe) {  // the pairs of an LDS-sized window
        const uint32_t fS = fe - fs;
        if (fS == 0 || fS > kCap || !fits32) return;
        const uint32_t fnk = (fS + kLT - 1) / kLT;
        rfirst = X[fs];
        rlast = X[fe - 1];
#pragma unroll
        for (int k = 0; k < kPre; ++k) {
            const uint32_t q = tid + (uint32_t)kLT * k;
            raw[k] = XT{};
            if ((uint32_t)k < fnk && q < fS) raw[k] = X[fs + q];
        }
    };
    fetch(s, e);
    for (; j < J; j += G) {
        uint32_t s2 = 0, e2 = 0;
        if (j + 2u * G < J) {
            s2 = wstart[j + 2u * G];
            e2 = wstart[j + 2u * G + 1u];
        }
        bool fetched = false;
        if (s < e) {  // some bucket starts in this window
            const uint32_t S = e - s;
            // 32-bit items: a window inside one second-digit bucket (all but
            // the few that cross a bucket start, which are handed on) has
            // keys (d << 32 | item) with d fixed, so its keys relative to the
            // first bucket are item - (first item with its low L bits clear)
            uint32_t ds = 0, hs = 0;
            bool ok = false;
            if constexpr (X32) {
                ds = d2_of_wave(s_st, s);  // s, e: the same in every lane
                hs = ihigh(rfirst, ds, hb, L);
                ok = S <= kCap && d2_of_wave(s_st, e - 1) == ds;
            } else {
                hs = khigh(rfirst, bbits, L);
                ok = fits32 && S <= kCap && khigh(rlast, bbits, L) - hs < (1u << (32 - L));
            }
            if (ok) {
                const uint32_t nk = (S + kLT - 1) / kLT;
                uint32_t kk[kLocItems];
#pragma unroll
                for (int k = 0; k < kLocItems; ++k) {
                    const uint32_t q = tid + (uint32_t)kLT * k;
                    XT p{};
                    if (k < kPre) {
                        p = raw[k < kPre ? k : 0];
                    } else if ((uint32_t)k < nk) {
                        if (q < S) p = X[s + q];
                    }
                    if constexpr (X32)
                        kk[k] = p - (rfirst & ~((1u << L) - 1u));
                    else
                        kk[k] = ((khigh(p, bbits, L) - hs) << L) | klow(p, bbits, L);
                }
                // every key is at most the last bucket's with all low bits set
                uint32_t khi;
                if constexpr (X32)
                    khi = (rlast | ((1u << L) - 1u)) - (rfirst & ~((1u << L) - 1u));
                else
                    khi = ((khigh(rlast, bbits, L) - hs) << L) | ((1u << L) - 1u);
                fetch(s1, e1);
                fetched = true;
                ok = bin_sort<kLT, kLocItems>(kk, nk, S, khi, s, hs, out_lim, out, s_e, s_bin,
                                              s_tmp, s_red, bbits, L);
            }
            if (!ok && tid == 0) big[atomicAdd(big_ctr, 1u)] = j;
        }
        if (!fetched) fetch(s1, e1);
        s = s1;
        e = e1;
        s1 = s2;
        e1 = e2;
    }
}

// The windows k_pair_local handed on, one 1024-thread workgroup each (grid-
// stride): bin_sort with twice the capacity (kBigC keys) where they fit,
// otherwise stable LSD passes through global memory X -> Y -> X ... with the
// last one into out, over keys relative to the window's first bucket, one
// chunk of kBigT * kGlobItems keys at a time (wave64 ballot ranking).
constexpr int kBigT = 1024;
constexpr int kBigItems = 16;
constexpr uint32_t kBigC = kBigT * kBigItems;
template <bool X32>
__global__ __launch_bounds__(kBigT) void k_pair_local_big(const void* Xv, uint2* Z, uint2* Y,
                                                          uint2* __restrict__ out,
                                                          const unsigned long long* __restrict__ n_dev,
                                                          const unsigned long long* __restrict__ last_n,
                                                          unsigned long long last_cap,
                                                          const uint32_t* __restrict__ wstart,
                                                          const uint32_t* __restrict__ big,
                                                          const uint32_t* __restrict__ big_ctr, int bbits,
                                                          int L, const uint32_t* __restrict__ st_g, int hb) {
    using IT = LItem<X32>;
    using XT = typename IT::T;
    const XT* const Xin = reinterpret_cast<const XT*>(Xv);
    // the global path's 8-byte pairs: the input itself, or (32-bit items) the
    // window expanded into the same positions of a free buffer Z
    uint2* const X = X32 ? Z : const_cast<uint2*>(reinterpret_cast<const uint2*>(Xv));
    constexpr int NW = kBigT / 64;
    constexpr uint32_t kChunk = kBigT * kGlobItems;
    // one LDS block, 80 KB: keys | bins (the LSD path's counters reuse it)
    __shared__ uint32_t s_mem[kBigC + kBins + 1];
    __shared__ uint32_t s_st[X32 ? 257 : 1];
    uint32_t* const s_e = s_mem;
    uint32_t* const s_bin = s_mem + kBigC;
    uint32_t* const s_tmp = s_e + (kBigC - kLocTail);
    uint32_t* const s_red = s_tmp + NW;
    uint32_t(*const s_wh)[256] = reinterpret_cast<uint32_t(*)[256]>(s_mem);  // [NW][256]
    uint32_t* const s_base = s_mem + NW * 256;
    uint32_t* const s_tmp2 = s_base + 256;
    constexpr uint32_t kCap = kBigC - kLocTail;

    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
    if constexpr (X32) load_st<kBigT>(st_g, s_st);
    const unsigned long long gtm = ~lanemask_lt() & ~(1ull << lane);
    const uint32_t n = (uint32_t)*n_dev;
    unsigned long long out_lim = n;
    {
        const unsigned long long ln = *last_n;
        if (ln < out_lim) out_lim = ln;
        if (last_cap < out_lim) out_lim = last_cap;
    }
    const bool fits32 = 2 * bbits - L <= 32;
    const uint32_t lowmask = (1u << L) - 1u;
    const uint32_t nbig = *big_ctr;
    for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const uint32_t j = big[bi];
        const uint32_t s = wstart[j], e = wstart[j + 1], S = e - s;
        uint32_t ds = 0;  // 32-bit items: the second digit of the window's first item
        if constexpr (X32) ds = d2_of(s_st, s);
        if (fits32 && S <= kCap) {
            uint32_t de = 0;
            if constexpr (X32) de = d2_of(s_st, e - 1);
            const uint32_t hs = IT::high(Xin[s], ds, bbits, hb, L);
            const uint32_t hspan = IT::high(Xin[e - 1], de, bbits, hb, L) - hs;
            if (hspan < (1u << (32 - L))) {
                const uint32_t nk = (S + kBigT - 1) / kBigT;
                uint32_t kk[kBigItems], d = ds;
#pragma unroll
                for (int k = 0; k < kBigItems; ++k) {
                    XT p{};
                    const uint32_t q = tid + (uint32_t)kBigT * k;
                    if ((uint32_t)k < nk && q < S) p = Xin[s + q];
                    if constexpr (X32)  // positions rise with k: advance the digit
                        while (d < 255u && s + q >= s_st[d + 1]) ++d;
                    kk[k] = ((IT::high(p, d, bbits, hb, L) - hs) << L) | IT::low(p, bbits, L);
                }
                if (bin_sort<kBigT, kBigItems>(kk, nk, S, (hspan << L) | lowmask, s, hs, out_lim,
                                               out, s_e, s_bin, s_tmp, s_red, bbits, L))
                    continue;
            }
        }
        if constexpr (X32) {  // expand the window to 8-byte pairs at X[s, e)
            for (uint32_t q = tid; q < S; q += kBigT) {
                const uint32_t dq = d2_of(s_st, s + q);
                const unsigned long long key =
                    ((unsigned long long)(dq >> hb) << 32) | (unsigned long long)Xin[s + q];
                X[s + q] = pkey_pair(key, bbits);
            }
            __syncthreads();
        }
        const unsigned long long hs = pkey(X[s], bbits) >> L;
        const unsigned long long he = pkey(X[e - 1], bbits) >> L;
        const unsigned long long kb = hs << L;
        const unsigned long long kmax = ((he - hs) << L) | lowmask;
        const uint32_t T = 64u - (uint32_t)__clzll((long long)kmax);
        const uint32_t npass = (T + 7u) / 8u;
        for (uint32_t p = 0; p < npass; ++p) {
            const bool last = p + 1 == npass;
            const uint2* src = (p & 1u) ? Y : X;
            uint2* dst = last ? out : ((p & 1u) ? X : Y);
            const uint32_t sh = 8u * p;
            const int bw = (int)min(8u, T - sh);
            const uint32_t mask = (1u << bw) - 1u;
            if (tid < 256u) s_base[tid] = 0u;
            __syncthreads();
            for (uint32_t q = tid; q < S; q += kBigT)
                atomicAdd(&s_base[(uint32_t)((pkey(src[s + q], bbits) - kb) >> sh) & mask], 1u);
            __syncthreads();
            uint32_t all;
            const uint32_t c0v = tid < 256u ? s_base[tid] : 0u;
            const uint32_t ex = lblock_excl_scan<kBigT>(c0v, s_tmp2, &all);
            if (tid < 256u) s_base[tid] = s + ex;
            __syncthreads();
            for (uint32_t c0 = 0; c0 < S; c0 += kChunk) {
                const uint32_t cs = min(S - c0, kChunk);
                const uint32_t nk = (cs + kBigT - 1) / kBigT;
                const uint32_t w0 = c0 + wave * 64u * nk;
                for (uint32_t i = lane; i < 256u; i += 64u) s_wh[wave][i] = 0u;
                unsigned long long kk[kGlobItems];
                uint32_t rk[kGlobItems];
#pragma unroll
                for (int k = 0; k < kGlobItems; ++k) {
                    const uint32_t q = w0 + 64u * k + lane;
                    kk[k] = 0ull;
                    if ((uint32_t)k < nk && q < c0 + cs) kk[k] = pkey(src[s + q], bbits) - kb;
                }
#pragma unroll
                for (int k = 0; k < kGlobItems; ++k) {
                    if ((uint32_t)k < nk) {
                        const uint32_t q = w0 + 64u * k + lane;
                        rk[k] = wave_rank((uint32_t)(kk[k] >> sh) & mask, q < c0 + cs, bw, s_wh[wave], gtm);
                    }
                }
                __syncthreads();
                if (tid < 256u) {
                    uint32_t run = s_base[tid];
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        const uint32_t cw = s_wh[w][tid];
                        s_wh[w][tid] = run;
                        run += cw;
                    }
                    s_base[tid] = run;
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < kGlobItems; ++k) {
                    if ((uint32_t)k < nk) {
                        const uint32_t q = w0 + 64u * k + lane;
                        if (q < c0 + cs) {
                            const uint32_t g = s_wh[wave][(uint32_t)(kk[k] >> sh) & mask] + rk[k];
                            // g lies in [s, e) when the window's pairs stay put
                            // between the count and the scatter; the bound keeps
                            // every store in the window whatever the data
                            if (g < e && (!last || g < out_lim)) dst[g] = pkey_pair(kk[k] + kb, bbits);
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

// Window starts (J + 2 words) followed by the list of handed-on windows.
static uint64_t wstart_half(uint64_t cap) { return cap / kLocW + 2; }
// window starts, the list of oversized windows, then the bucket starts,
// sizes and slot counters of a two-pass sort with 32-bit items
uint64_t pair_local_wstart_words(uint64_t cap) { return 2 * wstart_half(cap) + kBucketWords + 2 * kRadixSq; }
uint32_t* pair_local_bucket_starts(uint32_t* wstart, uint64_t cap) { return wstart + 2 * wstart_half(cap); }

// Window starts from bucket starts (k_xstarts, non-decreasing): window
// j starts at the first bucket start at or after j * kLocW.  Thread k writes
// the windows j with B[k-1] < j * kLocW <= B[k] (B clamped to the key count
// n; k = 65536 covers the windows after the last bucket start); window J =
// ceil(n / kLocW) is n.
__global__ void k_window_starts(const uint32_t* __restrict__ bst, const unsigned long long* n_dev,
                                uint32_t* __restrict__ wstart, uint32_t* __restrict__ big_ctr) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = (uint32_t)*n_dev;
    const uint32_t J = (n + kLocW - 1) / kLocW;
    if (k == 0) {
        *big_ctr = 0u;  // read by the kernels after this one
        wstart[0] = 0u;
        wstart[J] = n;
    }
    if (k > kBucketWords - 1u) return;
    const uint32_t hiB = k < kBucketWords - 1u ? min(bst[k], n) : n;
    const uint32_t loB = k == 0 ? 0u : min(bst[k - 1u], n);
    const uint32_t j1 = min(hiB / kLocW, J);
    for (uint32_t j = loB / kLocW + 1u; j <= j1; ++j) wstart[j] = hiB;
}

hipError_t launch_local_big_windows(uint2* x, uint2* y, uint2* out, const unsigned long long* n_dev,
                                    const unsigned long long* last_n, unsigned long long last_cap,
                                    const uint32_t* wst, const uint32_t* big, const uint32_t* big_ctr,
                                    int bbits, int L, hipStream_t st) {
    if (L < 8 || L > 24) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pair_local_big<false>, dim3(256), dim3(kBigT), 0, st, x, x, y, out, n_dev,
                       last_n, last_cap, wst, big, big_ctr, bbits, L, nullptr, 0);
    return hipGetLastError();
}

hipError_t launch_pair_local(uint2* x, uint2* y, uint2* out, const unsigned long long* n_dev,
                             const unsigned long long* last_n, unsigned long long last_cap,
                             unsigned long long cap, uint32_t* wstart, uint32_t* big_ctr, int bbits,
                             int L, const uint32_t* stg, const uint32_t* bst, hipStream_t st) {
    if (L < 8 || L > 24) return hipErrorInvalidValue;
    const bool x32 = stg != nullptr;
    const int kbits = 2 * bbits;
    if (x32 && (kbits > 40 || kbits - L != 16)) return hipErrorInvalidValue;
    const int hb = x32 ? 40 - kbits : 0;  // second-digit bits below bit 32
    const unsigned long long J = (cap + kLocW - 1) / kLocW;  // windows, at most
    uint32_t* big = wstart + wstart_half(cap);
    uint2* const z = x + cap;  // 32-bit items: expansion space of oversized windows
    unsigned long long gw = (J + 1 + kWaves - 1) / kWaves;
    if (gw > 2048) gw = 2048;
    // local-sort workgroups: up to 8,192 (three fit a CU; more than one
    // round of them balances the uneven windows: at 150 M pairs 2,048 ->
    // 8,192 took the stage 0.586 -> 0.549 ms, 65,536 0.593 ms)
    unsigned long long gl = J == 0 ? 1 : J;
    if (gl > 8192) gl = 8192;
    if (x32) {
        if (bst)
            hipLaunchKernelGGL(k_window_starts, dim3((kBucketWords + 255) / 256), dim3(256), 0, st, bst,
                               n_dev, wstart, big_ctr);
        else
            hipLaunchKernelGGL(k_pair_wstart<true>, dim3((uint32_t)gw), dim3(kWG), 0, st, x, n_dev, bbits,
                               L, wstart, big_ctr, stg);
        hipLaunchKernelGGL(k_pair_local<true>, dim3((uint32_t)gl), dim3(kLT), 0, st, x, out, n_dev,
                           last_n, last_cap, wstart, big, big_ctr, bbits, L, stg, hb);
        hipLaunchKernelGGL(k_pair_local_big<true>, dim3(256), dim3(kBigT), 0, st, x, z, y, out, n_dev,
                           last_n, last_cap, wstart, big, big_ctr, bbits, L, stg, hb);
    } else {
        hipLaunchKernelGGL(k_pair_wstart<false>, dim3((uint32_t)gw), dim3(kWG), 0, st, x, n_dev, bbits,
                           L, wstart, big_ctr, stg);
        hipLaunchKernelGGL(k_pair_local<false>, dim3((uint32_t)gl), dim3(kLT), 0, st, x, out, n_dev,
                           last_n, last_cap, wstart, big, big_ctr, bbits, L, stg, hb);
        hipLaunchKernelGGL(k_pair_local_big<false>, dim3(256), dim3(kBigT), 0, st, x, x, y, out, n_dev,
                           last_n, last_cap, wstart, big, big_ctr, bbits, L, stg, hb);
    }
    return hipGetLastError();
}

}  // namespace gsap
