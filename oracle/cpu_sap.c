/* cpu_sap.c -- the CPU baseline: a multi-threaded CPU restatement of the
 * reference's SaP update, timed beside the GPU (test infrastructure, like the
 * rest of oracle/; nothing in the product links it).
 *
 * Method, following physicboxSap2.16_Fermi/sapDetection/Sap.cpp:211-335:
 *   axis    the reference rule axisVecMultiple (SapKernel.cu:62-97): the
 *           coordinate axis of largest centre variance, ties to the lower
 *           index (fp64 moments here)
 *   keys    Herf's order-preserving float flip of each box's lower bound on
 *           that axis (radixsort.cu:110-138); on a coordinate axis the
 *           projection is exact
 *   sort    stable LSD radix sort of (key, index), 8-bit digits, 4 passes
 *           (the reference's RadixSort::sort, radixsort.cu:1384-1498)
 *   sweep   each sorted box tests the boxes after it up to the last lower
 *           key <= its upper key (the early exit of SapKernel.cu:1882), the
 *           AABB predicate of include/aabb.h:47-57 (inclusive)
 *   output  each unordered pair once as (min, max), sorted lexicographically
 *           by an LSD radix sort of the key i * 2^b + j (b = bits of n)
 * AABB input only (6 floats per box).  The reference's 4-cell split and its
 * cross pass are not restated (BASELINE's baseline is the plain sweep).
 * OpenMP: moments, projection, the radix passes (per-thread digit counts,
 * stable scatter) and the sweep (dynamic chunks, per-thread pair buffers).
 * The output equals or_sap's (checked in tests/test_oracle.py). */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double now_s(void) { return omp_get_wtime(); }

static inline uint32_t flip_f32(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t mask = (uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u;
    return u ^ mask;
}

/* Stable LSD radix sort of 32-bit keys with 32-bit values, 8-bit digits.
 * Each thread counts the digits of a contiguous chunk; the scatter places
 * chunk t of digit d after chunks < t of digit d, so the pass is stable. */
static void radix_kv32(uint32_t* k, uint32_t* v, uint32_t* k2, uint32_t* v2, size_t n, int nt) {
    size_t* cnt = (size_t*)calloc((size_t)nt * 256, sizeof(size_t));
    for (int shift = 0; shift < 32; shift += 8) {
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const size_t a = n * (size_t)t / (size_t)nt, b = n * (size_t)(t + 1) / (size_t)nt;
            size_t* c = cnt + (size_t)t * 256;
            memset(c, 0, 256 * sizeof(size_t));
            for (size_t i = a; i < b; ++i) c[(k[i] >> shift) & 255u]++;
#pragma omp barrier
#pragma omp single
            {
                size_t run = 0;
                for (int d = 0; d < 256; ++d)
                    for (int u = 0; u < nt; ++u) {
                        size_t x = cnt[(size_t)u * 256 + d];
                        cnt[(size_t)u * 256 + d] = run;
                        run += x;
                    }
            }
            for (size_t i = a; i < b; ++i) {
                const size_t o = c[(k[i] >> shift) & 255u]++;
                k2[o] = k[i];
                v2[o] = v[i];
            }
        }
        uint32_t* tk = k; k = k2; k2 = tk;
        uint32_t* tv = v; v = v2; v2 = tv;
    }
    /* four passes: the result is back in the caller's arrays */
    free(cnt);
}

/* Stable LSD radix sort of 64-bit keys over their low key_bits, 8-bit digits. */
static uint64_t* radix_u64(uint64_t* k, uint64_t* k2, size_t n, int key_bits, int nt) {
    size_t* cnt = (size_t*)calloc((size_t)nt * 256, sizeof(size_t));
    for (int shift = 0; shift < key_bits; shift += 8) {
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const size_t a = n * (size_t)t / (size_t)nt, b = n * (size_t)(t + 1) / (size_t)nt;
            size_t* c = cnt + (size_t)t * 256;
            memset(c, 0, 256 * sizeof(size_t));
            for (size_t i = a; i < b; ++i) c[(k[i] >> shift) & 255u]++;
#pragma omp barrier
#pragma omp single
            {
                size_t run = 0;
                for (int d = 0; d < 256; ++d)
                    for (int u = 0; u < nt; ++u) {
                        size_t x = cnt[(size_t)u * 256 + d];
                        cnt[(size_t)u * 256 + d] = run;
                        run += x;
                    }
            }
            for (size_t i = a; i < b; ++i) k2[c[(k[i] >> shift) & 255u]++] = k[i];
        }
        uint64_t* tk = k; k = k2; k2 = tk;
    }
    free(cnt);
    return k;
}

static size_t upper_bound_u32(const uint32_t* keys, size_t lo, size_t hi, uint32_t v) {
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (keys[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

typedef struct {
    uint64_t* p;
    size_t n, cap;
} pbuf;

static void push(pbuf* b, uint64_t x) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 4096;
        b->p = (uint64_t*)realloc(b->p, b->cap * sizeof(uint64_t));
    }
    b->p[b->n++] = x;
}

/* One update of n AABBs (6 floats each) on `threads` threads.  Writes up to
 * cap canonical pairs (uint32 x2) and returns the pair count; *tests_out is
 * the sweep-test count; times[0..3] = axis + projection, sort, sweep,
 * canonical sort (s).  group (optional) drops pairs with equal ids. */
int64_t cpu_sap(const float* boxes, size_t n, const uint32_t* group, int threads, uint32_t* pairs,
                uint64_t cap, uint64_t* tests_out, int* axis_out, double times[4]) {
    const int nt = threads > 0 ? threads : 1;
    if (n == 0) {
        if (tests_out) *tests_out = 0;
        return 0;
    }
    double t0 = now_s();
    /* coordinate axis of largest centre variance (fp64 sums per thread) */
    double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
#pragma omp parallel for num_threads(nt) reduction(+ : s1[:3], s2[:3]) schedule(static)
    for (size_t i = 0; i < n; ++i) {
        const float* b = boxes + 6 * i;
        for (int k = 0; k < 3; ++k) {
            const double c = ((double)b[k] + (double)b[k + 3]) * 0.5;
            s1[k] += c;
            s2[k] += c * c;
        }
    }
    double var[3];
    for (int k = 0; k < 3; ++k) {
        const double mu = s1[k] / (double)n;
        var[k] = s2[k] / (double)n - mu * mu;
    }
    int a = var[0] >= var[1] ? (var[0] >= var[2] ? 0 : 2) : (var[1] >= var[2] ? 1 : 2);
    const int ea = a == 0 ? 1 : 0, eb = 3 - a - ea;
    if (axis_out) *axis_out = a;
    uint32_t* keys = (uint32_t*)malloc(n * 4);
    uint32_t* perm = (uint32_t*)malloc(n * 4);
    uint32_t* k2 = (uint32_t*)malloc(n * 4);
    uint32_t* v2 = (uint32_t*)malloc(n * 4);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t i = 0; i < n; ++i) {
        keys[i] = flip_f32(boxes[6 * i + a]);
        perm[i] = (uint32_t)i;
    }
    double t1 = now_s();
    radix_kv32(keys, perm, k2, v2, n, nt);
    /* sorted structure of arrays: the two other axes' bounds, the upper key */
    float* la = (float*)malloc(n * 4);
    float* ha = (float*)malloc(n * 4);
    float* lb = (float*)malloc(n * 4);
    float* hb = (float*)malloc(n * 4);
    uint32_t* hik = k2;
#pragma omp parallel for num_threads(nt) schedule(static)
    for (size_t s = 0; s < n; ++s) {
        const float* b = boxes + 6 * (size_t)perm[s];
        la[s] = b[ea];
        ha[s] = b[ea + 3];
        lb[s] = b[eb];
        hb[s] = b[eb + 3];
        hik[s] = flip_f32(b[a + 3]);
    }
    double t2 = now_s();
    int bb = 1;
    while (bb < 32 && (1ull << bb) < n) ++bb;
    pbuf* bufs = (pbuf*)calloc((size_t)nt, sizeof(pbuf));
    uint64_t tests = 0;
#pragma omp parallel num_threads(nt) reduction(+ : tests)
    {
        pbuf* my = &bufs[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
        for (size_t s = 0; s < n; ++s) {
            const size_t end = upper_bound_u32(keys, s + 1, n, hik[s]);
            tests += end - s - 1;
            const float las = la[s], has = ha[s], lbs = lb[s], hbs = hb[s];
            const uint32_t i = perm[s];
            /* on the sweep axis every candidate overlaps (lo_s <= lo_t <= hi_s,
             * exact keys); test the other two axes, 32 candidates per mask */
            for (size_t t0 = s + 1; t0 < end; t0 += 32) {
                const size_t m = end - t0 < 32 ? end - t0 : 32;
                uint32_t mask = 0;
                for (size_t q = 0; q < m; ++q) {
                    const size_t t = t0 + q;
                    const int ok = (la[t] <= has) & (las <= ha[t]) & (lb[t] <= hbs) & (lbs <= hb[t]);
                    mask |= (uint32_t)ok << q;
                }
                while (mask) {
                    const int q = __builtin_ctz(mask);
                    mask &= mask - 1;
                    const uint32_t j = perm[t0 + (size_t)q];
                    if (group && group[i] == group[j]) continue;
                    const uint64_t lo = i < j ? i : j, hi = i < j ? j : i;
                    push(my, (lo << bb) | hi);
                }
            }
        }
    }
    double t3 = now_s();
    size_t P = 0;
    size_t* base = (size_t*)malloc(((size_t)nt + 1) * sizeof(size_t));
    for (int t = 0; t < nt; ++t) {
        base[t] = P;
        P += bufs[t].n;
    }
    uint64_t* pk = (uint64_t*)malloc((P ? P : 1) * 8);
    uint64_t* pk2 = (uint64_t*)malloc((P ? P : 1) * 8);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        if (bufs[t].n) memcpy(pk + base[t], bufs[t].p, bufs[t].n * 8);
        free(bufs[t].p);
    }
    const uint64_t* sorted = radix_u64(pk, pk2, P, 2 * bb, nt);
    const uint64_t mask = (1ull << bb) - 1;
    if (P <= cap) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (size_t k = 0; k < P; ++k) {
            pairs[2 * k] = (uint32_t)(sorted[k] >> bb);
            pairs[2 * k + 1] = (uint32_t)(sorted[k] & mask);
        }
    }
    double t4 = now_s();
    if (times) {
        times[0] = t1 - t0;
        times[1] = t2 - t1;
        times[2] = t3 - t2;
        times[3] = t4 - t3;
    }
    if (tests_out) *tests_out = tests;
    free(keys); free(perm); free(k2); free(v2); free(la); free(ha); free(lb); free(hb);
    free(bufs); free(base); free(pk); free(pk2);
    return (int64_t)P;
}

/* Threads OpenMP would use by default on this host. */
int cpu_sap_max_threads(void) { return omp_get_max_threads(); }
