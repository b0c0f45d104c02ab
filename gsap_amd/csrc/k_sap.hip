// k_sap.hip -- the object stages of one SaP update on gfx950.
//
// Reference pipeline: Sap::collision, physicboxSap2.16_Fermi/sapDetection/
// Sap.cpp:211-335.  Stage map (reference -> here):
//   Deviation/reduceExt/calDev/axisVecMultiple (Sap.cpp:148-209,
//     SapKernel.cu:62-97, 1341-1403)          -> k_moments (fp64, one pass,
//                                               last tile of each scene
//                                               solves the 3x3 eigenproblem)
//   projectSubMultiple (SapKernel.cu:150-213)  -> k_project (+ digit histogram)
//   reorderSphere + CounterThreadSphPCATexSub
//     (SapKernel.cu:1429-1504)                 -> k_gather (sorted records +
//                                               sweep end + sweep-test count)
//   kernel_SphereTexCullPCAElong/Sub + assem-
//     PairsSphereSub (SapKernel.cu:1649-1921)  -> k_sweep (wave per 64 objects,
//                                               candidates through the scalar
//                                               cache, ballot/mbcnt compaction)
//   cullRepetitivePairs / setPairsEntries
//     (SapKernel.cu:2249-2360, 2819-2855)      -> k_rowflag + k_entries
#include "gsap_device.h"

namespace gsap {

// ---------------------------------------------------------------- moments
__device__ static void jacobi3(double A[3][3], double V[3][3], double ev[3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
        double scale = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
        if (off <= 1e-300 || off <= scale * 1e-18) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = A[p][q];
                if (apq == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    ev[0] = A[0][0];
    ev[1] = A[1][1];
    ev[2] = A[2][2];
}

// Axis of one scene from its fp64 moments (same rules as the CPU oracle).
__device__ static void choose_axis(const double m[9], uint32_t n, int mode, const float* fixed,
                                   const float* learned, float u[3], float* src) {
    if (mode == 2) {  // fixed
        u[0] = fixed[0]; u[1] = fixed[1]; u[2] = fixed[2];
        *src = 2.0f;
        return;
    }
    if (mode == 3 && learned && learned[3] > 0.0f) {  // learned, confident
        u[0] = learned[0]; u[1] = learned[1]; u[2] = learned[2];
        *src = 3.0f;
        return;
    }
    double inv = n ? 1.0 / (double)n : 0.0;
    double mu[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
    double C[3][3];
    C[0][0] = m[3] * inv - mu[0] * mu[0];
    C[0][1] = C[1][0] = m[4] * inv - mu[0] * mu[1];
    C[0][2] = C[2][0] = m[5] * inv - mu[0] * mu[2];
    C[1][1] = m[6] * inv - mu[1] * mu[1];
    C[1][2] = C[2][1] = m[7] * inv - mu[1] * mu[2];
    C[2][2] = m[8] * inv - mu[2] * mu[2];
    if (mode == 1) {  // reference coordinate rule (SapKernel.cu:62-97)
        int ax = C[0][0] >= C[1][1] ? 0 : 1;
        if (ax) ax = C[1][1] >= C[2][2] ? 1 : 2;
        else ax = C[0][0] >= C[2][2] ? 0 : 2;
        u[0] = u[1] = u[2] = 0.0f;
        u[ax] = 1.0f;
        *src = 1.0f;
        return;
    }
    double V[3][3], ev[3];
    jacobi3(C, V, ev);
    int best = 0;
    for (int k = 1; k < 3; ++k)
        if (ev[k] > ev[best]) best = k;
    double v[3] = {V[0][best], V[1][best], V[2][best]};
    int big = 0;
    for (int k = 1; k < 3; ++k)
        if (fabs(v[k]) > fabs(v[big])) big = k;
    double sg = v[big] < 0 ? -1.0 : 1.0;
    double nn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (!(nn > 0)) { v[0] = 1; v[1] = 0; v[2] = 0; nn = 1; sg = 1; }
    for (int k = 0; k < 3; ++k) u[k] = (float)(sg * v[k] / nn);
    *src = 0.0f;
}

__global__ __launch_bounds__(kWG) void k_moments(ObjPipeArgs a) {
    __shared__ double s_red[kWaves][9];
    __shared__ uint32_t s_last;
    const uint32_t tile = blockIdx.x, tid = threadIdx.x;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);

    // Per-update housekeeping for the later stages (which run after this kernel).
    {
        const uint32_t total = a.seg.nseg * a.npass_obj * 256u;
        const uint32_t chunk = (total + gridDim.x - 1) / gridDim.x;
        const uint32_t b0 = tile * chunk, b1 = min(b0 + chunk, total);
        for (uint32_t i = b0 + tid; i < b1; i += kWG) a.hist_obj[i] = 0u;
        if (tile == 0) {
            for (uint32_t i = tid; i < a.npass_pair * 256u; i += kWG) a.hist_pair[i] = 0u;
            if (tid < 2 * kMaxPasses) a.ctrl->tile_ctr[tid] = 0u;
            if (tid == 0) a.ctrl->epoch += 1u;
        }
    }

    double s[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) s[k] = 0.0;
    bool bad = false;
    for (uint32_t i = ti.lo + tid; i < ti.hi; i += kWG) {
        double c[3];
        if (a.kind == 0) {
            const float2* b = reinterpret_cast<const float2*>(a.in) + 3 * (size_t)i;
            float2 p0 = b[0], p1 = b[1], p2 = b[2];
            bad |= !(isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y) &&
                     isfinite(p2.x) && isfinite(p2.y));
            c[0] = ((double)p0.x + (double)p1.y) * 0.5;
            c[1] = ((double)p0.y + (double)p2.x) * 0.5;
            c[2] = ((double)p1.x + (double)p2.y) * 0.5;
        } else {
            float4 q = reinterpret_cast<const float4*>(a.in)[i];
            bad |= !(isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w));
            c[0] = q.x; c[1] = q.y; c[2] = q.z;
        }
        s[0] += c[0]; s[1] += c[1]; s[2] += c[2];
        s[3] += c[0] * c[0]; s[4] += c[0] * c[1]; s[5] += c[0] * c[2];
        s[6] += c[1] * c[1]; s[7] += c[1] * c[2]; s[8] += c[2] * c[2];
    }
    const uint32_t w = tid >> 6, l = lane_id();
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double v = wave_sum(s[k]);
        if (l == 0) s_red[w][k] = v;
    }
    if (ballot(bad) && l == 0) atomicOr(&a.ctrl->flags, kFlagNonFinite);
    __syncthreads();
    if (tid < 9) {
        double v = 0.0;
        for (int ww = 0; ww < kWaves; ++ww) v += s_red[ww][tid];
        a.partials[(size_t)tile * 10 + tid] = v;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) {
        const uint32_t nt = a.seg.tile_seg ? (a.seg.seg_tile0[ti.seg + 1] - a.seg.seg_tile0[ti.seg])
                                           : a.ntiles;
        const uint32_t t = atomicAdd(&a.seg_ticket[ti.seg], 1u);
        s_last = (t == nt - 1u);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (tid == 0) {
        const uint32_t t0 = a.seg.tile_seg ? a.seg.seg_tile0[ti.seg] : 0u;
        const uint32_t t1 = a.seg.tile_seg ? a.seg.seg_tile0[ti.seg + 1] : a.ntiles;
        double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t t = t0; t < t1; ++t)
            for (int k = 0; k < 9; ++k)
                m[k] += __hip_atomic_load(&a.partials[(size_t)t * 10 + k], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        float u[3], src;
        const float* learned = a.learned_axis ? a.learned_axis + 4 * (size_t)ti.seg : nullptr;
        choose_axis(m, ti.seg_end - ti.seg_start, a.axis_mode, a.fixed_axis, learned, u, &src);
        float* ax = a.axis + 4 * (size_t)ti.seg;
        ax[0] = u[0]; ax[1] = u[1]; ax[2] = u[2]; ax[3] = src;
        a.seg_ticket[ti.seg] = 0u;
    }
}

// ---------------------------------------------------------------- projection
__global__ __launch_bounds__(kWG) void k_project(ObjPipeArgs a) {
    __shared__ uint32_t s_h[4][256];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    for (uint32_t i = tid; i < 4 * 256; i += kWG) (&s_h[0][0])[i] = 0u;
    const float* ax = a.axis + 4 * (size_t)ti.seg;
    const float u[3] = {ax[0], ax[1], ax[2]};
    const bool exact = axis_exact(a.kind, u);
    __syncthreads();
    for (uint32_t i = ti.lo + tid; i < ti.hi; i += kWG) {
        float lo[3], hi[3], c4[4];
        load_lohi(a.in, a.kind, i, lo, hi, c4);
        float pl, ph;
        project_lohi(a.kind, lo, hi, c4, u, exact, &pl, &ph);
        const uint32_t key = flip_f32(pl);
        a.keys[i] = key;
        a.vals[i] = i;
        atomicAdd(&s_h[0][key & 255u], 1u);
        atomicAdd(&s_h[1][(key >> 8) & 255u], 1u);
        atomicAdd(&s_h[2][(key >> 16) & 255u], 1u);
        atomicAdd(&s_h[3][key >> 24], 1u);
    }
    __syncthreads();
    uint32_t* h = a.hist_obj + (size_t)ti.seg * a.npass_obj * 256u;
    for (uint32_t i = tid; i < 4 * 256; i += kWG) {
        const uint32_t v = (&s_h[0][0])[i];
        if (v) atomicAdd(&h[i], v);
    }
}

// ---------------------------------------------------------------- gather
__global__ __launch_bounds__(kWG) void k_gather(ObjPipeArgs a, const uint32_t* __restrict__ K,
                                                const uint32_t* __restrict__ V) {
    const uint32_t tile = blockIdx.x, tid = threadIdx.x;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    const float* ax = a.axis + 4 * (size_t)ti.seg;
    const float u[3] = {ax[0], ax[1], ax[2]};
    const bool exact = axis_exact(a.kind, u);
    unsigned long long tests = 0;
    for (uint32_t s = ti.lo + tid; s < ti.hi; s += kWG) {
        const uint32_t orig = V[s];
        float lo[3], hi[3], c4[4];
        load_lohi(a.in, a.kind, orig, lo, hi, c4);
        float pl, ph;
        project_lohi(a.kind, lo, hi, c4, u, exact, &pl, &ph);
        const uint32_t hk = flip_f32(ph);
        uint32_t b = s + 1, e = ti.seg_end;  // first t in [s+1, seg_end) with K[t] > hk
        while (b < e) {
            const uint32_t mid = b + ((e - b) >> 1);
            if (K[mid] <= hk) b = mid + 1; else e = mid;
        }
        tests += (unsigned long long)(b - s - 1);
        if (a.kind == 0) {
            RecAabb r;
            r.lo[0] = lo[0]; r.lo[1] = lo[1]; r.lo[2] = lo[2];
            r.hi[0] = hi[0]; r.hi[1] = hi[1]; r.hi[2] = hi[2];
            r.orig = orig;
            r.end = b;
            reinterpret_cast<RecAabb*>(a.rec)[s] = r;
        } else {
            RecSphere r;
            r.c[0] = c4[0]; r.c[1] = c4[1]; r.c[2] = c4[2]; r.c[3] = c4[3];
            r.orig = orig;
            r.end = b;
            r.pad[0] = r.pad[1] = 0u;
            reinterpret_cast<RecSphere*>(a.rec)[s] = r;
        }
    }
    tests = wave_sum(tests);
    if (lane_id() == 0 && tests) atomicAdd(&a.ctrl->tests, tests);
}

// ---------------------------------------------------------------- sweep
// One workgroup per 256 consecutive sorted objects of one scene.  Objects are
// re-ordered inside the workgroup by their sweep end so each wave's 64 lanes
// have similar candidate ranges; every lane is one object, and the wave walks
// the union of its lanes' candidate ranges one candidate at a time.  The
// candidate record is wave-uniform (scalar loads); each lane runs the exact
// predicate, hits are compacted with ballot + mbcnt into a per-wave LDS
// buffer, and flushed with one global atomic per batch.  Candidates past a
// lane's own sweep end can never overlap it (the projected intervals are
// conservative), so only the t > s condition needs a mask.
template <int KIND, bool GROUP, int MODE>  // MODE 0 canonical, 1 symmetric, 2 count
__global__ __launch_bounds__(kWG) void k_sweep(SweepArgs a) {
    constexpr uint32_t BUF = 1024;
    __shared__ uint32_t s_end[kSweepTile];
    __shared__ uint32_t s_ord[kSweepTile];
    __shared__ uint2 s_buf[MODE == 2 ? 1 : kWaves][MODE == 2 ? 1 : BUF];
    using Rec = typename std::conditional<KIND == 0, RecAabb, RecSphere>::type;
    // Records are read-only here; the constant address space lets uniform
    // candidate loads go through the scalar cache (s_load_dwordx8).
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) Rec CRec;
#else
    typedef const Rec CRec;
#endif
    CRec* __restrict__ R = (CRec*)(a.rec);

    const uint32_t subs = a.tile / kSweepTile;
    const uint32_t tile = blockIdx.x / subs, sub = blockIdx.x % subs;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    const uint32_t lo = ti.lo + sub * kSweepTile;
    if (lo >= ti.hi) return;
    const uint32_t hi = min(lo + kSweepTile, ti.hi), cnt = hi - lo;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;

    const uint32_t e0 = tid < cnt ? reinterpret_cast<const Rec*>(a.rec)[lo + tid].end : 0u;
    s_end[tid] = e0;
    __syncthreads();
    uint32_t rank = 0;
    for (uint32_t j = 0; j < kSweepTile; ++j) {
        const uint32_t ej = s_end[j];
        rank += (ej < e0) | ((ej == e0) & (j < tid));
    }
    s_ord[rank] = tid;
    __syncthreads();
    const uint32_t my = s_ord[tid];
    const bool act = my < cnt;
    const uint32_t pos = lo + my;
    Rec me = reinterpret_cast<const Rec*>(a.rec)[act ? pos : lo];
    const uint32_t myend = act ? me.end : 0u;
    const uint32_t wend = uniform(wave_max_u32(myend));
    uint32_t gme = 0;
    if (GROUP && act) gme = a.group[me.orig];

    unsigned long long wcount = 0;  // MODE 2
    uint32_t wcnt = 0;              // staged entries (uniform)
    unsigned long long base_dummy = 0;
    (void)base_dummy;

    auto flush = [&]() {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&a.ctrl->num_pairs, (unsigned long long)wcnt);
        base = __shfl(base, 0, 64);
        for (uint32_t k = lane; k < wcnt; k += 64) {
            if (base + k < a.cap) a.out[base + k] = s_buf[w][k];
        }
        wcnt = 0;
    };

    auto test = [&](const Rec& c) -> bool {
        bool hit;
        if constexpr (KIND == 0) {
            hit = (c.lo[0] <= me.hi[0]) & (me.lo[0] <= c.hi[0]) & (c.lo[1] <= me.hi[1]) &
                  (me.lo[1] <= c.hi[1]) & (c.lo[2] <= me.hi[2]) & (me.lo[2] <= c.hi[2]);
        } else {
#pragma clang fp contract(off)
            const float x0 = me.c[0] - c.c[0];
            const float y0 = me.c[1] - c.c[1];
            const float z0 = me.c[2] - c.c[2];
            const float r0 = me.c[3] + c.c[3];
            const float t1 = (x0 * x0 + y0 * y0) + z0 * z0;
            const float e1 = r0 * r0;
            hit = t1 < e1;
        }
        return hit;
    };

    auto handle = [&](bool hit, uint32_t corig) {
        if (GROUP) hit = hit && (a.group[corig] != gme);
        const unsigned long long m = ballot(hit);
        if (m) {
            if constexpr (MODE == 2) {  // pre-count: directed pairs when symmetric
                wcount += (unsigned long long)__popcll(m) << (a.sym ? 1 : 0);
            } else {
                if (hit) {
                    const uint32_t p = wcnt + (MODE == 1 ? 2u : 1u) * mbcnt(m);
                    if constexpr (MODE == 0) {
                        const uint32_t x = me.orig < corig ? me.orig : corig;
                        const uint32_t y = me.orig < corig ? corig : me.orig;
                        s_buf[w][p] = make_uint2(x, y);
                    } else {
                        s_buf[w][p] = make_uint2(me.orig, corig);
                        s_buf[w][p + 1] = make_uint2(corig, me.orig);
                    }
                }
                wcnt += (MODE == 1 ? 2u : 1u) * (uint32_t)__popcll(m);
                if (wcnt > BUF - 128u) flush();
            }
        }
    };

    // phase A: candidates inside this workgroup's own range (need t > s)
    const uint32_t a_end = min(hi, wend);
    for (uint32_t t = lo + 1; t < a_end; ++t) {
        const Rec c = R[t];
        const bool h = test(c) & act & (t > pos);
        handle(h, c.orig);
    }
    // phase B: candidates after the workgroup's range, four per step so the
    // scalar loads of the next candidates overlap the tests of the current.
    uint32_t t = hi;
    for (; t + 4 <= wend; t += 4) {
        const Rec c0 = R[t], c1 = R[t + 1], c2 = R[t + 2], c3 = R[t + 3];
        const bool h0 = test(c0) & act, h1 = test(c1) & act;
        const bool h2 = test(c2) & act, h3 = test(c3) & act;
        handle(h0, c0.orig);
        handle(h1, c1.orig);
        handle(h2, c2.orig);
        handle(h3, c3.orig);
    }
    for (; t < wend; ++t) {
        const Rec c = R[t];
        handle(test(c) & act, c.orig);
    }
    if constexpr (MODE == 2) {
        if (lane == 0 && wcount) atomicAdd(&a.ctrl->num_pairs, wcount);
    } else {
        if (wcnt) flush();
    }
    if (lane == 0 && wend > lo + 1)
        atomicAdd(&a.ctrl->exec_tests, (unsigned long long)(wend - lo - 1) * 64ull);
}

// ---------------------------------------------------------------- CSR entries
// Row starts of the sorted directed pair list (reference pairEntries).
__global__ void k_rowflag(const uint2* __restrict__ P, const unsigned long long* n_dev,
                          unsigned long long cap, uint32_t* rowstart, uint32_t* has) {
    unsigned long long m = *n_dev;
    if (m > cap) m = cap;
    for (unsigned long long k = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; k < m;
         k += (unsigned long long)gridDim.x * blockDim.x) {
        const uint32_t x = P[k].x;
        if (k == 0 || P[k - 1].x != x) {
            rowstart[x] = (uint32_t)k;
            has[x] = 1u;
        }
    }
}

// Two-phase compaction of `has` into entries (block b: objects [b*2048, ...)).
__global__ __launch_bounds__(kWG) void k_entries_count(const uint32_t* has, uint32_t n,
                                                       uint32_t* block_sums) {
    __shared__ uint32_t s_tmp[kWaves];
    const uint32_t b0 = blockIdx.x * 2048u;
    uint32_t c = 0;
    for (uint32_t i = b0 + threadIdx.x; i < min(b0 + 2048u, n); i += kWG) c += has[i];
    uint32_t tot;
    block_excl_scan(c, s_tmp, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kWG) void k_entries_write(uint32_t* has, const uint32_t* rowstart,
                                                       uint32_t n, const uint32_t* block_sums,
                                                       int32_t* entries, Ctrl* ctrl) {
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ uint32_t s_base;
    const uint32_t b0 = blockIdx.x * 2048u, tid = threadIdx.x;
    if (tid == 0) {
        uint32_t s = 0;
        for (uint32_t b = 0; b < blockIdx.x; ++b) s += block_sums[b];
        s_base = s;
        if (blockIdx.x == gridDim.x - 1) ctrl->num_entries = s + block_sums[blockIdx.x];
    }
    __syncthreads();
    uint32_t run = s_base;
    for (uint32_t c0 = b0; c0 < min(b0 + 2048u, n); c0 += kWG) {
        const uint32_t i = c0 + tid;
        const uint32_t h = (i < n) ? has[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(h, s_tmp, &tot);
        if (h) {
            entries[run + ex] = (int32_t)rowstart[i];
            has[i] = 0u;  // leave `has` zeroed for the next update
        }
        run += tot;
    }
}

__global__ void k_finish(Ctrl* ctrl, unsigned long long cap, unsigned long long* d_counts) {
    if (threadIdx.x == 0) {
        const unsigned long long np = ctrl->num_pairs;
        uint32_t fl = ctrl->flags;
        if (np > cap) fl |= kFlagCapacity;
        d_counts[0] = np;
        d_counts[1] = ctrl->num_entries;
        d_counts[2] = ctrl->tests;
        d_counts[3] = fl;
        d_counts[4] = ctrl->exec_tests;
        ctrl->num_pairs = 0;
        ctrl->num_entries = 0;
        ctrl->tests = 0;
        ctrl->exec_tests = 0;
        ctrl->flags = 0;
    }
}

// Per-scene ranges of the sorted (scene-major) canonical pair list.
__global__ void k_scene_offsets(const uint2* __restrict__ P, const Ctrl* ctrl, unsigned long long cap,
                                const uint32_t* __restrict__ seg_start, uint32_t nseg,
                                unsigned long long* out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nseg) return;
    unsigned long long m = ctrl->num_pairs;
    if (m > cap) m = cap;
    if (s == nseg) { out[s] = m; return; }
    const uint32_t key = seg_start[s];
    unsigned long long b = 0, e = m;
    while (b < e) {
        const unsigned long long mid = b + ((e - b) >> 1);
        if (P[mid].x < key) b = mid + 1; else e = mid;
    }
    out[s] = b;
}

// ---------------------------------------------------------------- launchers
hipError_t launch_moments(const ObjPipeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_moments, dim3(a.ntiles), dim3(kWG), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_project(const ObjPipeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_project, dim3(a.ntiles), dim3(kWG), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_gather(const ObjPipeArgs& a, const uint32_t* K, const uint32_t* V, hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3(a.ntiles), dim3(kWG), 0, st, a, K, V);
    return hipGetLastError();
}

template <int KIND, bool GROUP, int MODE>
static void sweep_launch(const SweepArgs& a, hipStream_t st) {
    const uint32_t blocks = a.ntiles * (a.tile / kSweepTile);
    hipLaunchKernelGGL((k_sweep<KIND, GROUP, MODE>), dim3(blocks), dim3(kWG), 0, st, a);
}

template <int KIND, bool GROUP>
static void sweep_mode(const SweepArgs& a, hipStream_t st) {
    if (a.count_only) sweep_launch<KIND, GROUP, 2>(a, st);
    else if (a.sym) sweep_launch<KIND, GROUP, 1>(a, st);
    else sweep_launch<KIND, GROUP, 0>(a, st);
}

hipError_t launch_sweep(const SweepArgs& a, hipStream_t st) {
    if (a.ntiles == 0) return hipSuccess;
    if (a.kind == 0) {
        if (a.group) sweep_mode<0, true>(a, st); else sweep_mode<0, false>(a, st);
    } else {
        if (a.group) sweep_mode<1, true>(a, st); else sweep_mode<1, false>(a, st);
    }
    return hipGetLastError();
}

hipError_t launch_entries(const uint2* P, const unsigned long long* n_dev, unsigned long long cap,
                          uint32_t nobj, uint32_t* rowstart, uint32_t* has, uint32_t* block_sums,
                          int32_t* entries, Ctrl* ctrl, hipStream_t st) {
    unsigned long long g = (cap + 255) / 256;
    const uint32_t grid = (uint32_t)(g > 4096 ? 4096 : (g == 0 ? 1 : g));
    hipLaunchKernelGGL(k_rowflag, dim3(grid), dim3(256), 0, st, P, n_dev, cap, rowstart, has);
    const uint32_t nb = (nobj + 2047) / 2048;
    if (nb == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_entries_count, dim3(nb), dim3(kWG), 0, st, has, nobj, block_sums);
    hipLaunchKernelGGL(k_entries_write, dim3(nb), dim3(kWG), 0, st, has, rowstart, nobj, block_sums,
                       entries, ctrl);
    return hipGetLastError();
}

hipError_t launch_finish(Ctrl* ctrl, unsigned long long cap, unsigned long long* d_counts, int sym,
                         hipStream_t st) {
    (void)sym;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, st, ctrl, cap, d_counts);
    return hipGetLastError();
}

hipError_t launch_scene_offsets(const uint2* P, Ctrl* ctrl, unsigned long long cap,
                                const uint32_t* seg_start, uint32_t nseg, unsigned long long* out,
                                hipStream_t st) {
    hipLaunchKernelGGL(k_scene_offsets, dim3((nseg + 1 + 255) / 256), dim3(256), 0, st, P, ctrl, cap,
                       seg_start, nseg, out);
    return hipGetLastError();
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace gsap
