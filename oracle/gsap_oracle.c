/*
 * gsap_oracle.c -- CPU restatement of the reference SaP broadphase and the
 * particle rigid-body step.  TEST INFRASTRUCTURE ONLY (see gsap_oracle.h).
 *
 * Reference paths are relative to physicboxSap2.16_Fermi/ (PB/).
 * Build: gcc -O2 -std=c99 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#define _POSIX_C_SOURCE 199309L
#include "gsap_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Phase times of the last or_sap call (setup = axis+project+sort, sweep,
 * canonicalisation), for the CPU baseline. */
static double g_times[3];
void or_last_times(double out[3]) { out[0] = g_times[0]; out[1] = g_times[1]; out[2] = g_times[2]; }

/* ------------------------------------------------------------------------ */
/* Order-preserving float <-> uint32 map (Herf "radix tricks"), applied by    */
/* the reference inside its first/last sort pass: PB/sapDetection/          */
/* radixsort.cu:110-138 (floatFlip / floatUnflip).                           */
/* ------------------------------------------------------------------------ */
uint32_t or_flip(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t mask = (uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u;
    return u ^ mask;
}

float or_unflip(uint32_t u) {
    uint32_t mask = ((u >> 31) - 1u) | 0x80000000u;
    uint32_t v = u ^ mask;
    float f;
    memcpy(&f, &v, 4);
    return f;
}

void or_flip_array(const float* in, uint32_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = or_flip(in[i]);
}

/* Box accessors: AABB = 6 floats (lo.xyz, hi.xyz); sphere = float4 (x,y,z,r). */
static inline void box_lohi(const float* boxes, int kind, size_t i, float lo[3], float hi[3]) {
    if (kind == OR_KIND_AABB) {
        const float* b = boxes + 6 * i;
        lo[0] = b[0]; lo[1] = b[1]; lo[2] = b[2];
        hi[0] = b[3]; hi[1] = b[4]; hi[2] = b[5];
    } else {
        const float* b = boxes + 4 * i;
        float r = b[3];
        for (int k = 0; k < 3; ++k) { lo[k] = b[k] - r; hi[k] = b[k] + r; }
    }
}

/* ------------------------------------------------------------------------ */
/* Moments.  The reference accumulates the mean and the six covariance terms */
/* in fp32 with a tree reduction (PB/sapDetection/Sap.cpp:148-209,           */
/* SapKernel.cu:1341-1403).  Here the sums are fp64 (the one FP64 use of the */
/* path); centre = (lo+hi)/2 for boxes, (x,y,z) for spheres.                 */
/* ------------------------------------------------------------------------ */
void or_moments(const float* boxes, int kind, size_t n, double out[9]) {
    double s[9] = {0};
    for (size_t i = 0; i < n; ++i) {
        double c[3];
        if (kind == OR_KIND_AABB) {
            const float* b = boxes + 6 * i;
            for (int k = 0; k < 3; ++k) c[k] = ((double)b[k] + (double)b[k + 3]) * 0.5;
        } else {
            const float* b = boxes + 4 * i;
            for (int k = 0; k < 3; ++k) c[k] = (double)b[k];
        }
        s[0] += c[0]; s[1] += c[1]; s[2] += c[2];
        s[3] += c[0] * c[0]; s[4] += c[0] * c[1]; s[5] += c[0] * c[2];
        s[6] += c[1] * c[1]; s[7] += c[1] * c[2]; s[8] += c[2] * c[2];
    }
    memcpy(out, s, sizeof(s));
}

static void cov_from_moments(const double m[9], size_t n, double C[3][3]) {
    double inv = n ? 1.0 / (double)n : 0.0;
    double mu[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
    C[0][0] = m[3] * inv - mu[0] * mu[0];
    C[0][1] = C[1][0] = m[4] * inv - mu[0] * mu[1];
    C[0][2] = C[2][0] = m[5] * inv - mu[0] * mu[2];
    C[1][1] = m[6] * inv - mu[1] * mu[1];
    C[1][2] = C[2][1] = m[7] * inv - mu[1] * mu[2];
    C[2][2] = m[8] * inv - mu[2] * mu[2];
}

/* Cyclic Jacobi eigen-solve of a symmetric 3x3 (fp64).  Replaces the
 * reference's coordinate-axis choice (SapKernel.cu:62-97) in PCA mode. */
static void jacobi3(double A[3][3], double V[3][3], double ev[3]) {
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
    ev[0] = A[0][0]; ev[1] = A[1][1]; ev[2] = A[2][2];
}

void or_pca_axis(const double mom[9], size_t n, float axis[3], double evals[3]) {
    double C[3][3], V[3][3], ev[3];
    cov_from_moments(mom, n, C);
    jacobi3(C, V, ev);
    int best = 0;
    for (int k = 1; k < 3; ++k)
        if (ev[k] > ev[best]) best = k;
    double u[3] = {V[0][best], V[1][best], V[2][best]};
    int big = 0;
    for (int k = 1; k < 3; ++k)
        if (fabs(u[k]) > fabs(u[big])) big = k;
    double sg = u[big] < 0 ? -1.0 : 1.0;
    double nn = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (!(nn > 0)) { u[0] = 1; u[1] = 0; u[2] = 0; nn = 1; sg = 1; }
    for (int k = 0; k < 3; ++k) axis[k] = (float)(sg * u[k] / nn);
    if (evals) { evals[0] = ev[0]; evals[1] = ev[1]; evals[2] = ev[2]; }
}

/* Reference rule (PB/sapDetection/SapKernel.cu:62-97, axisVecMultiple):
 * axis = C00 >= C11 ? (C00 >= C22 ? 0 : 2) : (C11 >= C22 ? 1 : 2);
 * lamda = {axis, axis==0 ? 1 : 0, 3-axis-axis1}. */
int or_axis_coord_maxvar(const double mom[9], size_t n, float axis[3], int lamda[3]) {
    double C[3][3];
    cov_from_moments(mom, n, C);
    int a = C[0][0] >= C[1][1] ? 0 : 1;
    if (a) a = C[1][1] >= C[2][2] ? 1 : 2;
    else a = C[0][0] >= C[2][2] ? 0 : 2;
    int a1 = a == 0 ? 1 : 0;
    if (lamda) { lamda[0] = a; lamda[1] = a1; lamda[2] = 3 - a - a1; }
    axis[0] = axis[1] = axis[2] = 0.0f;
    axis[a] = 1.0f;
    return a;
}

/* Per-coordinate extent sums and maxima (fp64): hi_k - lo_k of each AABB,  */
/* 2r of each sphere's hull box.                                            */
void or_extents(const float* boxes, int kind, size_t n, double sum[3], double mx[3]) {
    for (int k = 0; k < 3; ++k) { sum[k] = 0.0; mx[k] = 0.0; }
    for (size_t i = 0; i < n; ++i) {
        double x[3];
        if (kind == OR_KIND_AABB) {
            const float* b = boxes + 6 * i;
            for (int k = 0; k < 3; ++k) x[k] = (double)b[k + 3] - (double)b[k];
        } else {
            const float* b = boxes + 4 * i;
            for (int k = 0; k < 3; ++k) x[k] = 2.0 * (double)b[3];
        }
        for (int k = 0; k < 3; ++k) {
            sum[k] += x[k];
            if (x[k] > mx[k]) mx[k] = x[k];
        }
    }
}

/* Slab axis of the slab sweep: the top eigenvector of P C P, P = I - u u^T  */
/* (the direction of largest spread orthogonal to u), any unit vector        */
/* orthogonal to u when there is no spread across u.  Generalises the        */
/* reference's choice of the two minor coordinate axes for its 4-cell split  */
/* (PB/sapDetection/SapKernel.cu:62-97, lamda[1], lamda[2]).                 */
void or_slab_axis(const double mom[9], size_t n, const float uf[3], float v[3]) {
    double C[3][3], P[3][3], PC[3][3], A[3][3], V[3][3], ev[3];
    const double u[3] = {uf[0], uf[1], uf[2]};
    cov_from_moments(mom, n, C);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = (i == j ? 1.0 : 0.0) - u[i] * u[j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) PC[i][j] = P[i][0] * C[0][j] + P[i][1] * C[1][j] + P[i][2] * C[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = PC[i][0] * P[0][j] + PC[i][1] * P[1][j] + PC[i][2] * P[2][j];
    jacobi3(A, V, ev);
    int best = 0;
    for (int k = 1; k < 3; ++k)
        if (ev[k] > ev[best]) best = k;
    double w[3] = {V[0][best], V[1][best], V[2][best]};
    const double d = w[0] * u[0] + w[1] * u[1] + w[2] * u[2];
    for (int k = 0; k < 3; ++k) w[k] -= d * u[k];
    double nn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (!(nn > 1e-6)) {
        int k = 0;
        for (int j = 1; j < 3; ++j)
            if (fabs(u[j]) < fabs(u[k])) k = j;
        double e[3] = {0.0, 0.0, 0.0};
        e[k] = 1.0;
        w[0] = u[1] * e[2] - u[2] * e[1];
        w[1] = u[2] * e[0] - u[0] * e[2];
        w[2] = u[0] * e[1] - u[1] * e[0];
        nn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (!(nn > 0)) { w[0] = 0; w[1] = 1; w[2] = 0; nn = 1; }
    }
    for (int k = 0; k < 3; ++k) v[k] = (float)(w[k] / nn);
}

/* Sweep-test model (no reference counterpart; the reference always sweeps  */
/* the coordinate axis of largest variance, SapKernel.cu:62-97).  A sweep   */
/* along unit u tests about n^2 E[ext_u] / (2 sqrt(pi var_u)) pairs for a    */
/* Gaussian spread; in slabs of width W_v on a second axis v about           */
/* n^2 E[ext_u] W_v / (4 pi sqrt(det C_uv)).  ext_u = sum |u_k| ext_k for an */
/* AABB, W_v <= sum |v_k| max ext_k.  Constants dropped; +inf without spread.*/
static double quad3(const double C[3][3], const double u[3], const double v[3]) {
    double r = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r += u[i] * C[i][j] * v[j];
    return r;
}
double or_axis_cost(const double mom[9], const double xsum[3], const double xmax[3], size_t n,
                    const float uf[3], const float vf[3]) {
    double C[3][3];
    cov_from_moments(mom, n, C);
    const double inv = n ? 1.0 / (double)n : 0.0;
    const double u[3] = {uf[0], uf[1], uf[2]};
    const double em = fabs(u[0]) * (xsum[0] * inv) + fabs(u[1]) * (xsum[1] * inv) +
                      fabs(u[2]) * (xsum[2] * inv);
    if (!vf) {
        const double var = quad3(C, u, u);
        return var > 0.0 ? em / sqrt(var) : INFINITY;
    }
    const double v[3] = {vf[0], vf[1], vf[2]};
    const double det = quad3(C, u, u) * quad3(C, v, v) - quad3(C, u, v) * quad3(C, u, v);
    const double wv = fabs(v[0]) * xmax[0] + fabs(v[1]) * xmax[1] + fabs(v[2]) * xmax[2];
    return det > 0.0 ? em * wv / sqrt(det) : INFINITY;
}

/* Automatic axis: the PCA axis (with slabs: the PCA axis and its slab axis) */
/* unless a coordinate axis (pair) has a strictly smaller model cost,        */
/* candidates in the order x, y, z (pairs (a,b), a != b, a-major).  Returns  */
/* 0 when PCA is kept, 1 for a coordinate choice.                            */
int or_axis_auto(const double mom[9], const double xsum[3], const double xmax[3], size_t n,
                 int slabs, float u[3], float v[3]) {
    or_pca_axis(mom, n, u, NULL);
    if (slabs) or_slab_axis(mom, n, u, v);
    if (n < 2) return 0;
    double best = or_axis_cost(mom, xsum, xmax, n, u, slabs ? v : NULL);
    int bu = -1, bv = -1;
    for (int a = 0; a < 3; ++a) {
        float e[3] = {0.0f, 0.0f, 0.0f};
        e[a] = 1.0f;
        if (!slabs) {
            const double c = or_axis_cost(mom, xsum, xmax, n, e, NULL);
            if (c < best) { best = c; bu = a; }
            continue;
        }
        for (int b = 0; b < 3; ++b) {
            if (a == b) continue;
            float f[3] = {0.0f, 0.0f, 0.0f};
            f[b] = 1.0f;
            const double c = or_axis_cost(mom, xsum, xmax, n, e, f);
            if (c < best) { best = c; bu = a; bv = b; }
        }
    }
    if (bu < 0) return 0;
    u[0] = u[1] = u[2] = 0.0f;
    u[bu] = 1.0f;
    if (slabs) {
        v[0] = v[1] = v[2] = 0.0f;
        v[bv] = 1.0f;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Projection to 32-bit order-preserving keys.  The reference projects       */
/* sphere centres, key = dot(axis,c) - r (+ a cell offset),                  */
/* PB/sapDetection/SapKernel.cu:150-213.  This restatement projects the box  */
/* hull conservatively: lower = sum_k u_k * (u_k >= 0 ? lo_k : hi_k), upper  */
/* symmetric, each widened by 2^-20 * sum|terms| + FLT_MIN unless the axis is */
/* a signed unit coordinate axis and the input is an AABB (then exact).      */
/* Evaluation order ((t0 + t1) + t2), no FMA -- the HIP kernel is identical. */
/* ------------------------------------------------------------------------ */
static int axis_is_unit_coord(const float u[3]) {
    int nz = 0, unit = 0;
    for (int k = 0; k < 3; ++k)
        if (u[k] != 0.0f) { ++nz; unit = (u[k] == 1.0f || u[k] == -1.0f); }
    return nz == 1 && unit;
}

static void project_one(const float* boxes, int kind, size_t i, const float u[3], int exact,
                        float* plo, float* phi) {
    float lo[3], hi[3];
    box_lohi(boxes, kind, i, lo, hi);
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
        if (kind == OR_KIND_AABB) {
            ml = (fabsf(a0) + fabsf(a1)) + fabsf(a2);
            mh = (fabsf(b0) + fabsf(b1)) + fabsf(b2);
        } else {
            const float* b = boxes + 4 * i;
            float r = b[3];
            float m0 = fabsf(u[0]) * (fabsf(b[0]) + r);
            float m1 = fabsf(u[1]) * (fabsf(b[1]) + r);
            float m2 = fabsf(u[2]) * (fabsf(b[2]) + r);
            ml = mh = (m0 + m1) + m2;
        }
        l = l - (ml * 0x1p-20f + FLT_MIN);
        h = h + (mh * 0x1p-20f + FLT_MIN);
    }
    *plo = l;
    *phi = h;
}

void or_project(const float* boxes, int kind, size_t n, const float axis[3], uint32_t* keys,
                uint32_t* hikeys) {
    int exact = (kind == OR_KIND_AABB) && axis_is_unit_coord(axis);
    for (size_t i = 0; i < n; ++i) {
        float l, h;
        project_one(boxes, kind, i, axis, exact, &l, &h);
        keys[i] = or_flip(l);
        if (hikeys) hikeys[i] = or_flip(h);
    }
}

/* ------------------------------------------------------------------------ */
/* Stable LSD radix sort.  Semantics of the reference sorter                 */
/* (PB/sapDetection/radixsort.h:180-198, radixsort.cu:1384-1480: stable key- */
/* value LSD, ceil(keyBits/digit) passes), restated as a counting sort.      */
/* ------------------------------------------------------------------------ */
void or_radix_sort_kv(uint32_t* keys, uint32_t* vals, size_t n, int digit_bits) {
    if (n < 2) return;
    if (digit_bits < 1) digit_bits = 8;
    if (digit_bits > 16) digit_bits = 16;
    size_t nb = (size_t)1 << digit_bits;
    uint32_t* k2 = (uint32_t*)malloc(n * 4);
    uint32_t* v2 = (uint32_t*)malloc(n * 4);
    size_t* cnt = (size_t*)malloc(nb * sizeof(size_t));
    for (int shift = 0; shift < 32; shift += digit_bits) {
        uint32_t mask = (uint32_t)(nb - 1);
        memset(cnt, 0, nb * sizeof(size_t));
        for (size_t i = 0; i < n; ++i) cnt[(keys[i] >> shift) & mask]++;
        size_t run = 0;
        for (size_t d = 0; d < nb; ++d) { size_t c = cnt[d]; cnt[d] = run; run += c; }
        for (size_t i = 0; i < n; ++i) {
            size_t d = (keys[i] >> shift) & mask;
            size_t o = cnt[d]++;
            k2[o] = keys[i];
            v2[o] = vals[i];
        }
        memcpy(keys, k2, n * 4);
        memcpy(vals, v2, n * 4);
    }
    free(k2); free(v2); free(cnt);
}

static inline uint64_t pair_logical_key(const uint32_t* p, int bbits) {
    return ((uint64_t)p[0] << bbits) | (uint64_t)p[1];
}

void or_radix_sort_pairs(uint32_t* pairs, size_t n, int bbits, int key_bits, int digit_bits) {
    if (n < 2) return;
    if (digit_bits < 1) digit_bits = 8;
    size_t nb = (size_t)1 << digit_bits;
    uint32_t* tmp = (uint32_t*)malloc(n * 8);
    size_t* cnt = (size_t*)malloc(nb * sizeof(size_t));
    for (int shift = 0; shift < key_bits; shift += digit_bits) {
        uint64_t mask = (uint64_t)(nb - 1);
        memset(cnt, 0, nb * sizeof(size_t));
        for (size_t i = 0; i < n; ++i) cnt[(pair_logical_key(pairs + 2 * i, bbits) >> shift) & mask]++;
        size_t run = 0;
        for (size_t d = 0; d < nb; ++d) { size_t c = cnt[d]; cnt[d] = run; run += c; }
        for (size_t i = 0; i < n; ++i) {
            size_t d = (pair_logical_key(pairs + 2 * i, bbits) >> shift) & mask;
            size_t o = cnt[d]++;
            tmp[2 * o] = pairs[2 * i];
            tmp[2 * o + 1] = pairs[2 * i + 1];
        }
        memcpy(pairs, tmp, n * 8);
    }
    free(tmp); free(cnt);
}

/* ------------------------------------------------------------------------ */
/* Pair predicates.                                                           */
/* AABB: PB/include/aabb.h:47-57 (Aabb::overlaps, inclusive <=).             */
/* Sphere: PB/sapDetection/SapKernel.cu:1897-1903, strict                    */
/*   (dx*dx + dy*dy + dz*dz) < (r0+r1)*(r0+r1), evaluated left to right.     */
/* ------------------------------------------------------------------------ */
int or_overlap_aabb(const float* a, const float* b) {
    int x = a[0] <= b[3] && b[0] <= a[3];
    int y = a[1] <= b[4] && b[1] <= a[4];
    int z = a[2] <= b[5] && b[2] <= a[5];
    return x && y && z;
}

int or_overlap_sphere(const float* a, const float* b) {
    float x0 = a[0] - b[0];
    float y0 = a[1] - b[1];
    float z0 = a[2] - b[2];
    float r0 = a[3] + b[3];
    float t1 = (x0 * x0 + y0 * y0) + z0 * z0;
    float e1 = r0 * r0;
    return t1 < e1;
}

static inline int overlap(const float* boxes, int kind, size_t i, size_t j) {
    if (kind == OR_KIND_AABB) return or_overlap_aabb(boxes + 6 * i, boxes + 6 * j);
    return or_overlap_sphere(boxes + 4 * i, boxes + 4 * j);
}

static int cmp_pair(const void* x, const void* y) {
    const uint32_t* a = (const uint32_t*)x;
    const uint32_t* b = (const uint32_t*)y;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
    return 0;
}

void or_canon_sort(uint32_t* pairs, uint64_t P) {
    qsort(pairs, (size_t)P, 8, cmp_pair);
}

/* Upper bound: first index t with keys[t] > v. */
static size_t upper_bound_u32(const uint32_t* keys, size_t n, uint32_t v) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (keys[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void choose_axis(const float* boxes, int kind, size_t n, int axis_mode,
                        const float* fixed_axis, float u[3]) {
    if (axis_mode == OR_AXIS_FIXED && fixed_axis) {
        u[0] = fixed_axis[0]; u[1] = fixed_axis[1]; u[2] = fixed_axis[2];
        return;
    }
    double mom[9];
    or_moments(boxes, kind, n, mom);
    if (axis_mode == OR_AXIS_COORD_MAXVAR) {
        or_axis_coord_maxvar(mom, n, u, NULL);
    } else if (axis_mode == OR_AXIS_AUTO) {
        double xs[3], xm[3];
        or_extents(boxes, kind, n, xs, xm);
        or_axis_auto(mom, xs, xm, n, 0, u, NULL);
    } else {
        or_pca_axis(mom, n, u, NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* Full SaP update.  Pipeline of PB/sapDetection/Sap.cpp:211-335 (axis,      */
/* project, sort, sweep, canonicalise) with one global sweep instead of the  */
/* reference's 4-cell + cross-cell passes: sorted object s tests every t > s */
/* with key_t <= hikey_s (the reference's early exit d_pos[cur] > key+2r,    */
/* SapKernel.cu:1882) and emits each unordered pair once as (min,max).       */
/* ------------------------------------------------------------------------ */
int64_t or_sap(const float* boxes, int kind, size_t n, int axis_mode, const float* fixed_axis,
               const uint32_t* group, uint32_t* pairs, uint64_t cap, uint64_t* tests_out,
               float axis_out[3], uint32_t sample_stride) {
    float u[3];
    if (n == 0) {
        if (tests_out) *tests_out = 0;
        return 0;
    }
    double t0 = now_s();
    choose_axis(boxes, kind, n, axis_mode, fixed_axis, u);
    if (axis_out) { axis_out[0] = u[0]; axis_out[1] = u[1]; axis_out[2] = u[2]; }
    uint32_t* keys = (uint32_t*)malloc(n * 4);
    uint32_t* hik = (uint32_t*)malloc(n * 4);
    uint32_t* perm = (uint32_t*)malloc(n * 4);
    or_project(boxes, kind, n, u, keys, hik);
    for (size_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    or_radix_sort_kv(keys, perm, n, 8);
    uint64_t P = 0, tests = 0;
    uint32_t stride = sample_stride ? sample_stride : 1;
    double t1 = now_s();
    for (size_t s = 0; s < n; s += stride) {
        uint32_t i = perm[s];
        size_t end = upper_bound_u32(keys, n, hik[i]);
        if (end > s + 1) tests += end - s - 1;
        for (size_t t = s + 1; t < end; ++t) {
            uint32_t j = perm[t];
            if (group && group[i] == group[j]) continue;
            if (overlap(boxes, kind, i, j)) {
                if (P < cap) {
                    pairs[2 * P] = i < j ? i : j;
                    pairs[2 * P + 1] = i < j ? j : i;
                }
                ++P;
            }
        }
    }
    double t2 = now_s();
    if (P <= cap) or_canon_sort(pairs, P);
    double t3 = now_s();
    g_times[0] = t1 - t0;
    g_times[1] = t2 - t1;
    g_times[2] = t3 - t2;
    if (tests_out) *tests_out = tests;
    free(keys); free(hik); free(perm);
    return (int64_t)P;
}

/* One sweep over a subset of boxes with given lower / upper keys (float,
 * compared through the order-preserving flip): per sorted box, candidates up
 * to the last lower key <= its upper key.  Adds the window sizes to *tests and,
 * unless count_only, the overlapping pairs (min, max original index). */
static void sweep_subset(const float* boxes, int kind, const uint32_t* ids, size_t m,
                         const float* lok, const float* hik, int count_only, uint64_t* tests,
                         uint32_t* pairs, uint64_t cap, uint64_t* P) {
    uint32_t* keys = (uint32_t*)malloc((m ? m : 1) * 4);
    uint32_t* perm = (uint32_t*)malloc((m ? m : 1) * 4);
    for (size_t k = 0; k < m; ++k) { keys[k] = or_flip(lok[k]); perm[k] = (uint32_t)k; }
    or_radix_sort_kv(keys, perm, m, 8);
    for (size_t s = 0; s < m; ++s) {
        const uint32_t a = perm[s];
        const size_t end = upper_bound_u32(keys, m, or_flip(hik[a]));
        if (end > s + 1) *tests += end - s - 1;
        if (count_only) continue;
        for (size_t t = s + 1; t < end; ++t) {
            const uint32_t i = ids[a], j = ids[perm[t]];
            if (overlap(boxes, kind, i, j)) {
                if (*P < cap) {
                    pairs[2 * *P] = i < j ? i : j;
                    pairs[2 * *P + 1] = i < j ? j : i;
                }
                ++*P;
            }
        }
    }
    free(keys);
    free(perm);
}

/* The reference's cell subdivision and cross pass, restated for work counts
 * (Sap.cpp:211-295; projectSubMultiple SapKernel.cu:150-213;
 * optimalAxisMultiple :2634-2678 -> cullCrossSphereMultiple :556-635,
 * cullRelevantSphereMultiple :967-1059, setRelevantSphere :475-491;
 * sweepPrunSphereSubElong / SubCross :2364-2517).
 *   axis: the coordinate axis of largest variance (axisVecMultiple); the two
 *   others a1 < a2 carry the split planes at 0 (mPoint1 = mPoint2 = 0).
 *   cell q of a box by the signs of its centre on a1, a2 (1: both >= 0, 2: a1
 *   < 0 <= a2, 3: a1 >= 0 > a2, 0: both < 0); cell pass keys lo + q*interval,
 *   hi + q*interval with interval larger than the world extent (the reference
 *   carries the previous frame's extent), so windows stay in their cell.
 *   maxR: the largest half-extent (radius) of the boxes straddling a plane;
 *   relevant boxes: within maxR of a plane (hi >= -maxR and lo <= maxR on a1
 *   or a2); cross pass: the relevant boxes swept again without cell offsets.
 *   The union of both passes, deduplicated, is the pair set.  For spheres
 *   lo/hi = c -/+ r, as in the reference; for AABBs the box itself.
 * Returns the number of unique pairs (canonical in `pairs` when <= cap);
 * stats: tests of the cell pass, of the cross pass, relevant boxes, pairs
 * before deduplication.  count_only: window sizes only (no pairs). */
int64_t or_sap_ref_cells(const float* boxes, int kind, size_t n, int count_only, uint32_t* pairs,
                         uint64_t cap, uint64_t stats[4]) {
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    if (n == 0) return 0;
    double mom[9];
    or_moments(boxes, kind, n, mom);
    float u[3];
    int lam[3];
    or_axis_coord_maxvar(mom, n, u, lam);
    const int a0 = lam[0];
    const int a1 = a0 == 0 ? 1 : 0, a2 = a0 == 2 ? 1 : 2;
    float wlo[3] = {INFINITY, INFINITY, INFINITY}, whi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i) {
        float lo[3], hi[3];
        box_lohi(boxes, kind, i, lo, hi);
        for (int k = 0; k < 3; ++k) {
            if (lo[k] < wlo[k]) wlo[k] = lo[k];
            if (hi[k] > whi[k]) whi[k] = hi[k];
        }
    }
    float interval = 0.0f;
    for (int k = 0; k < 3; ++k)
        if (whi[k] - wlo[k] > interval) interval = whi[k] - wlo[k];
    interval = interval * 2.0f + 1.0f;
    uint32_t* ids = (uint32_t*)malloc(n * 4);
    float* lok = (float*)malloc(n * 4);
    float* hik = (float*)malloc(n * 4);
    float maxr = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float lo[3], hi[3];
        box_lohi(boxes, kind, i, lo, hi);
        const float c1 = 0.5f * (lo[a1] + hi[a1]), c2 = 0.5f * (lo[a2] + hi[a2]);
        const int q = (c1 >= 0.0f && c2 >= 0.0f) ? 1 : (c1 < 0.0f && c2 >= 0.0f) ? 2
                    : (c1 >= 0.0f && c2 < 0.0f) ? 3 : 0;
        ids[i] = (uint32_t)i;
        lok[i] = lo[a0] + (float)q * interval;
        hik[i] = hi[a0] + (float)q * interval;
        const int st1 = hi[a1] >= 0.0f && lo[a1] <= 0.0f, st2 = hi[a2] >= 0.0f && lo[a2] <= 0.0f;
        if (st1 || st2) {
            const float h1 = 0.5f * (hi[a1] - lo[a1]), h2 = 0.5f * (hi[a2] - lo[a2]);
            const float h = h1 > h2 ? h1 : h2;
            if (h > maxr) maxr = h;
        }
    }
    uint64_t P = 0;
    sweep_subset(boxes, kind, ids, n, lok, hik, count_only, &stats[0], pairs, cap, &P);
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        float lo[3], hi[3];
        box_lohi(boxes, kind, i, lo, hi);
        if ((hi[a1] >= -maxr && lo[a1] <= maxr) || (hi[a2] >= -maxr && lo[a2] <= maxr)) {
            ids[m] = (uint32_t)i;
            lok[m] = lo[a0];
            hik[m] = hi[a0];
            ++m;
        }
    }
    stats[2] = m;
    sweep_subset(boxes, kind, ids, m, lok, hik, count_only, &stats[1], pairs, cap, &P);
    stats[3] = P;
    free(ids); free(lok); free(hik);
    if (count_only || P > cap) return (int64_t)P;
    or_canon_sort(pairs, P);
    uint64_t w = 0;  /* dedupe: pairs found by both passes */
    for (uint64_t k = 0; k < P; ++k)
        if (w == 0 || pairs[2 * k] != pairs[2 * (w - 1)] || pairs[2 * k + 1] != pairs[2 * (w - 1) + 1]) {
            pairs[2 * w] = pairs[2 * k];
            pairs[2 * w + 1] = pairs[2 * k + 1];
            ++w;
        }
    return (int64_t)w;
}

uint64_t or_sweep_tests(const float* boxes, int kind, size_t n, const float axis[3]) {
    if (n == 0) return 0;
    uint32_t* keys = (uint32_t*)malloc(n * 4);
    uint32_t* hik = (uint32_t*)malloc(n * 4);
    uint32_t* perm = (uint32_t*)malloc(n * 4);
    or_project(boxes, kind, n, axis, keys, hik);
    for (size_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    or_radix_sort_kv(keys, perm, n, 8);
    uint64_t tests = 0;
    for (size_t s = 0; s < n; ++s) {
        size_t end = upper_bound_u32(keys, n, hik[perm[s]]);
        if (end > s + 1) tests += end - s - 1;
    }
    free(keys); free(hik); free(perm);
    return tests;
}

/* Brute force: PB/cudaPhysics/demoParticles.h:1496-1568 (spheres) and
 * PB/cudaPhysics/demo.h:389-424 (AABBs), all pairs i<j, canonical output. */
int64_t or_brute(const float* boxes, int kind, size_t n, const uint32_t* group, uint32_t* pairs,
                 uint64_t cap) {
    uint64_t P = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) {
            if (group && group[i] == group[j]) continue;
            if (overlap(boxes, kind, i, j)) {
                if (P < cap) { pairs[2 * P] = (uint32_t)i; pairs[2 * P + 1] = (uint32_t)j; }
                ++P;
            }
        }
    return (int64_t)P;
}

/* Reference output format (PB/sapDetection/Sap.cpp:269-330, SapKernel.cu:
 * 2249-2360, PB/readme.txt): every unordered pair in both orientations, sorted
 * by (i,j), plus pairEntries = start offset of each distinct i. */
uint64_t or_sym_csr(const uint32_t* canon, uint64_t P, int32_t* pairs_out, int32_t* entries_out,
                    uint32_t* n_entries) {
    uint32_t* tmp = (uint32_t*)malloc((size_t)(2 * P + 1) * 8);
    for (uint64_t k = 0; k < P; ++k) {
        tmp[4 * k] = canon[2 * k];
        tmp[4 * k + 1] = canon[2 * k + 1];
        tmp[4 * k + 2] = canon[2 * k + 1];
        tmp[4 * k + 3] = canon[2 * k];
    }
    or_canon_sort(tmp, 2 * P);
    uint32_t ne = 0;
    for (uint64_t k = 0; k < 2 * P; ++k) {
        pairs_out[2 * k] = (int32_t)tmp[2 * k];
        pairs_out[2 * k + 1] = (int32_t)tmp[2 * k + 1];
        if (k == 0 || tmp[2 * k] != tmp[2 * k - 2]) {
            if (entries_out) entries_out[ne] = (int32_t)k;
            ++ne;
        }
    }
    if (n_entries) *n_entries = ne;
    free(tmp);
    return 2 * P;
}

/* ------------------------------------------------------------------------ */
/* Particle rigid body.  PB/ParticleRigidBody/ParticleRigidBodyKernel.h and  */
/* the math of PB/include/quaternion.h:90-148 and PB/src/cutil_math.h:       */
/* 400-557 (float4 ops component-wise, a /= s computed as a *= 1/s).         */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z, w; } f4;

static inline f4 ld4(const float* p) { f4 r = {p[0], p[1], p[2], p[3]}; return r; }
static inline void st4(float* p, f4 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w; }
static inline f4 add4(f4 a, f4 b) { f4 r = {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; return r; }
static inline f4 sub4(f4 a, f4 b) { f4 r = {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; return r; }
static inline f4 mul4s(f4 a, float s) { f4 r = {a.x * s, a.y * s, a.z * s, a.w * s}; return r; }
static inline f4 smul4(float s, f4 a) { f4 r = {s * a.x, s * a.y, s * a.z, s * a.w}; return r; }
static inline float dot4(f4 a, f4 b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
static inline float dot3(f4 a, f4 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* cross(float4,float4) -> float4 with w = 0 (quaternion.h:14-20, cutil_math.h:400-403) */
static inline f4 cross4(f4 a, f4 b) {
    f4 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f};
    return r;
}
/* Quaternion product (quaternion.h:130-139): cross + (a.w*b + b.w*a); w = a.w*b.w - dot3 */
static inline f4 qmul(f4 a, f4 b) {
    f4 ans = cross4(a, b);
    ans = add4(ans, add4(smul4(a.w, b), smul4(b.w, a)));
    ans.w = a.w * b.w - dot3(a, b);
    return ans;
}
/* rotateVector (quaternion.h:141-148): q * in * q^-1 with in.w carried */
static inline f4 qrot(f4 q, f4 in) {
    f4 qi = {-q.x, -q.y, -q.z, q.w};
    return qmul(qmul(q, in), qi);
}
/* normalizeQ (quaternion.h:116-128): a /= length(a) -> a *= 1/sqrtf(dot4) */
static inline f4 qnormalize(f4 q) {
    float l = sqrtf(dot4(q, q));
    float inv = 1.0f / l;
    return mul4s(q, inv);
}

/* integrateK, ParticleRigidBodyKernel.h:47-69 (gravity parameterised). */
void or_prb_integrate(float* pos, float* quat, float* linvel, const float* angvel, int nb,
                      float dt, const or_prb_params* p) {
    f4 g = {p->gravity[0], p->gravity[1], p->gravity[2], 0.0f};
    f4 gdt = mul4s(g, dt);
    for (int i = 0; i < nb; ++i) {
        f4 v = add4(ld4(linvel + 4 * i), gdt);
        st4(linvel + 4 * i, v);
        f4 q = ld4(quat + 4 * i);
        f4 w = ld4(angvel + 4 * i);
        w.w = 0.0f;
        q = add4(q, mul4s(qmul(w, q), dt * 0.5f));
        q = qnormalize(q);
        st4(quat + 4 * i, q);
        float* x = pos + 4 * i;
        x[0] = x[0] + v.x * dt;
        x[1] = x[1] + v.y * dt;
        x[2] = x[2] + v.z * dt;
        if (p->flags & OR_PRB_FLOOR_CLAMP) {
            if ((i / 6) % 2 == 0 && x[1] <= 0.25f) x[1] = 0.25f;
            else if ((i / 6) % 2 == 1 && x[1] <= 0.3f) x[1] = 0.3f;
        }
    }
}

/* updateParticleDataK, ParticleRigidBodyKernel.h:71-104. */
void or_prb_update_particles(const float* pos, const float* quat, const float* linvel,
                             const float* angvel, const int32_t* shape, const int32_t* pstart,
                             const int32_t* npart, const float* ref_pos, const float* ref_mass,
                             int max_per_shape, float* ppos, float* pvel, float* pmass, int nb) {
    for (int b = 0; b < nb; ++b) {
        f4 x = ld4(pos + 4 * b), v = ld4(linvel + 4 * b), q = ld4(quat + 4 * b);
        f4 w = ld4(angvel + 4 * b);
        for (int ip = 0; ip < npart[b]; ++ip) {
            int r = max_per_shape * shape[b] + ip;
            f4 ref = ld4(ref_pos + 4 * r);
            f4 cur = qrot(q, ref);
            int o = pstart[b] + ip;
            st4(pvel + 4 * o, add4(v, cross4(w, cur)));
            cur = add4(cur, x);
            f4 pp = {cur.x, cur.y, cur.z, ref.w};
            st4(ppos + 4 * o, pp);
            pmass[o] = ref_mass[r];
        }
    }
}

/* Rotation matrix from a quaternion (quaternion.h:90-106). */
static void qmat(f4 q, float R[3][3]) {
    float x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z;
    R[0][0] = 1 - 2 * y2 - 2 * z2;
    R[0][1] = 2 * q.x * q.y - 2 * q.w * q.z;
    R[0][2] = 2 * q.x * q.z + 2 * q.w * q.y;
    R[1][0] = 2 * q.x * q.y + 2 * q.w * q.z;
    R[1][1] = 1 - 2 * x2 - 2 * z2;
    R[1][2] = 2 * q.y * q.z - 2 * q.w * q.x;
    R[2][0] = 2 * q.x * q.z - 2 * q.w * q.y;
    R[2][1] = 2 * q.y * q.z + 2 * q.w * q.x;
    R[2][2] = 1 - 2 * x2 - 2 * y2;
}

/* prepareK, ParticleRigidBodyKernel.h:30-45: Iw^-1 = (R I0^-1) R^T. */
void or_prb_prepare(const float* quat, const int32_t* shape, const float* shape_inv_inertia,
                    float* out, int nb) {
    for (int b = 0; b < nb; ++b) {
        float R[3][3], I[3][3], T[3][3];
        qmat(ld4(quat + 4 * b), R);
        const float* s = shape_inv_inertia + 12 * shape[b];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) I[i][j] = s[4 * i + j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) T[i][j] = (R[i][0] * I[0][j] + R[i][1] * I[1][j]) + R[i][2] * I[2][j];
        float* o = out + 12 * b;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) o[4 * i + j] = (T[i][0] * R[j][0] + T[i][1] * R[j][1]) + T[i][2] * R[j][2];
            o[4 * i + 3] = 0.0f;
        }
    }
}

/* calcSpringCoeff, ParticleRigidBodyKernel.h:13-22. */
static inline void spring_coeff(float im, float jm, float dt2inv, float dtinv, const or_prb_params* p,
                                float* s, float* d) {
    if (p->flags & OR_PRB_ABS_SPRING) { *s = p->k_abs; *d = p->c_abs; return; }
    float m = im + jm;
    m = 1.0f / m;
    *s = m * dt2inv * p->sp;
    *d = m * dtinv * (1.0f - p->dp);
}

/* One boundary plane: penetration along unit normal nrm (floor: y = 0 with
 * normal +y), force -k*(dist-r)*n*scale + c*(-v): clearForceAndBoundary-
 * ConditionK, ParticleRigidBodyKernel.h:106-140.  Walls add the same law for
 * the x and z planes (the walled box of falling cubes), applied after the floor. */
static inline int plane_force(f4* out, float dist, f4 nrm, f4 pos, f4 vel, float mass,
                              float dt2inv, float dtinv, const or_prb_params* p) {
    if (!(dist < pos.w)) return 0;
    float s, d;
    spring_coeff(1.0f / mass, 0.5f / mass, dt2inv, dtinv, p, &s, &d);
    f4 nv = {-vel.x, -vel.y, -vel.z, -vel.w};
    f4 sf = mul4s(smul4(-s * (dist - pos.w), nrm), p->floor_scale);
    f4 df = smul4(d, nv);
    *out = add4(sf, df);
    return 1;
}

static inline f4 boundary_force(f4 pos, f4 vel, float mass, float dt2inv, float dtinv,
                                const or_prb_params* p) {
    f4 force = {0, 0, 0, 0}, c;
    f4 ny = {0, 1, 0, 0};
    if (plane_force(&c, pos.y, ny, pos, vel, mass, dt2inv, dtinv, p)) force = c;
    if (p->flags & OR_PRB_WALLS) {
        f4 px = {1, 0, 0, 0}, mx = {-1, 0, 0, 0}, pz = {0, 0, 1, 0}, mz = {0, 0, -1, 0};
        if (plane_force(&c, pos.x - p->wall_lo[0], px, pos, vel, mass, dt2inv, dtinv, p)) force = add4(force, c);
        if (plane_force(&c, p->wall_hi[0] - pos.x, mx, pos, vel, mass, dt2inv, dtinv, p)) force = add4(force, c);
        if (plane_force(&c, pos.z - p->wall_lo[1], pz, pos, vel, mass, dt2inv, dtinv, p)) force = add4(force, c);
        if (plane_force(&c, p->wall_hi[1] - pos.z, mz, pos, vel, mass, dt2inv, dtinv, p)) force = add4(force, c);
    }
    return force;
}

/* clearForceAndBoundaryConditionK + comupteFoceOnParticlesK
 * (ParticleRigidBodyKernel.h:106-208): one row of the symmetric sorted pair
 * list per particle, partners in list order; the loop is bounded by the pair
 * count (the reference bounds it by numPairEntries, :201). */
void or_prb_forces(const float* ppos, const float* pvel, const float* pmass,
                   const int32_t* sym, uint64_t n_sym, const int32_t* pbody, float* pforce, int np,
                   float dt, const or_prb_params* p) {
    float dt2inv = 1.0f / (dt * dt);
    float dtinv = 1.0f / dt;
    for (int i = 0; i < np; ++i)
        st4(pforce + 4 * i, boundary_force(ld4(ppos + 4 * i), ld4(pvel + 4 * i), pmass[i], dt2inv, dtinv, p));
    uint64_t k = 0;
    while (k < n_sym) {
        int i = sym[2 * k];
        f4 ip = ld4(ppos + 4 * i), iv = ld4(pvel + 4 * i);
        float im = pmass[i];
        f4 sF = {0, 0, 0, 0}, dF = {0, 0, 0, 0};
        for (; k < n_sym && sym[2 * k] == i; ++k) {
            int j = sym[2 * k + 1];
            if ((p->flags & OR_PRB_SKIP_SAME_BODY) && pbody && pbody[i] == pbody[j]) continue;
            f4 jp = ld4(ppos + 4 * j), jv = ld4(pvel + 4 * j);
            float jm = pmass[j];
            f4 rij = sub4(jp, ip);
            rij.w = 0.0f;
            f4 vij = sub4(jv, iv);
            float dist = sqrtf(dot4(rij, rij));
            float orig = ip.w + jp.w;
            float s, d;
            spring_coeff(1.0f / im, 1.0f / jm, dt2inv, dtinv, p, &s, &d);
            float pen = orig - dist;
            if (pen > 0.0f) {
                float inv = 1.0f / dist;
                sF = sub4(sF, mul4s(smul4(s * pen, rij), inv));
                dF = add4(dF, smul4(d, vij));
            }
        }
        f4 f = ld4(pforce + 4 * i);
        st4(pforce + 4 * i, add4(f, add4(sF, dF)));
    }
}

/* updateVelocityK, ParticleRigidBodyKernel.h:213-262. */
void or_prb_update_velocity(float* linvel, float* angvel, const float* quat, const float* inv_mass,
                            const float* inv_inertia, const int32_t* shape, const int32_t* pstart,
                            const int32_t* npart, const float* ref_pos, int max_per_shape,
                            const float* pforce, int nb, float dt, const or_prb_params* p) {
    for (int b = 0; b < nb; ++b) {
        f4 q = ld4(quat + 4 * b);
        f4 lf = {0, 0, 0, 0}, af = {0, 0, 0, 0};
        for (int ip = 0; ip < npart[b]; ++ip) {
            int r = max_per_shape * shape[b] + ip;
            f4 cur = qrot(q, ld4(ref_pos + 4 * r));
            f4 f = ld4(pforce + 4 * (pstart[b] + ip));
            lf = add4(lf, f);
            af = add4(af, cross4(cur, f));
        }
        f4 v = add4(ld4(linvel + 4 * b), mul4s(mul4s(lf, inv_mass[b]), dt));
        if (p->flags & OR_PRB_CAP_VEL) {
            f4 lv = v;
            lv.w = 0.0f;
            float len = sqrtf(dot4(lv, lv));
            if (len > p->max_lin_vel) {
                float inv = 1.0f / sqrtf(dot4(lv, lv));
                lv = mul4s(lv, inv);
                v = smul4(p->max_lin_vel, lv);
            }
        }
        st4(linvel + 4 * b, v);
        f4 w = ld4(angvel + 4 * b);
        if (p->flags & OR_PRB_INERTIA) {
            const float* I = inv_inertia + 12 * b;
            f4 dw = {(I[0] * af.x + I[1] * af.y) + I[2] * af.z, (I[4] * af.x + I[5] * af.y) + I[6] * af.z,
                     (I[8] * af.x + I[9] * af.y) + I[10] * af.z, 0.0f};
            w = add4(w, mul4s(dw, dt));
        } else {
            w = add4(w, mul4s(af, dt));
        }
        st4(angvel + 4 * b, w);
    }
}
