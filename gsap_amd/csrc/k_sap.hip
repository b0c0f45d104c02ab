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
    __shared__ double s_red[kWaves][15];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);

    // Per-update housekeeping for the later stages (which run after this kernel).
    {
        const uint32_t total = a.seg.nseg * a.npass_obj * (1u << a.bits_obj);
        const uint32_t chunk = (total + gridDim.x - 1) / gridDim.x;
        const uint32_t b0 = tile * chunk, b1 = min(b0 + chunk, total);
        for (uint32_t i = b0 + tid; i < b1; i += kWG) a.hist_obj[i] = 0u;
        if (a.slabs) {  // 256 slab counters per scene
            const uint32_t tot_s = a.seg.nseg * 256u;
            const uint32_t ch = (tot_s + gridDim.x - 1) / gridDim.x;
            const uint32_t c0 = tile * ch, c1 = min(c0 + ch, tot_s);
            for (uint32_t i = c0 + tid; i < c1; i += kWG) a.slab_cnt[i] = 0u;
        }
        if (tile == 0) {
            for (uint32_t i = tid; i < a.npass_pair * 256u; i += kWG) a.hist_pair[i] = 0u;
            if (tid < 2 * kMaxPasses) a.ctrl->tile_ctr[tid] = 0u;
            if (a.pair_claim)
                for (uint32_t i = tid; i < kPairClaimWords; i += kWG) a.pair_claim[i] = 0u;
            if (tid == 0) a.ctrl->epoch += 1u;
            if (tid == 0 && a.keep_key) {  // the coherent sort's side list and fallback flag
                a.ctrl->coh_far = 0u;
                a.ctrl->coh_fallback = 0u;
            }
        }
    }

    if (a.keep_key) {  // the key function is kept (no axis from these moments): the input check only
        bool bad = false;
        for (uint32_t i = ti.lo + tid; i < ti.hi; i += kWG) {
            if (a.kind == 0) {
                const float2* b = reinterpret_cast<const float2*>(a.in) + 3 * (size_t)i;
                const float2 p0 = b[0], p1 = b[1], p2 = b[2];
                bad |= !(isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y) &&
                         isfinite(p2.x) && isfinite(p2.y));
            } else {
                const float4 q = reinterpret_cast<const float4*>(a.in)[i];
                bad |= !(isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w));
            }
        }
        if (ballot(bad) && lane_id() == 0) atomicOr(&a.ctrl->flags, kFlagNonFinite);
        return;
    }
    // s[0..8]: sum c, sum c c^T; s[9..11]: sum of the extents; e[0..2]: largest extents
    double s[12], e[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 12; ++k) s[k] = 0.0;
    bool bad = false;
    for (uint32_t i = ti.lo + tid; i < ti.hi; i += kWG) {
        double c[3], x[3];
        if (a.kind == 0) {
            const float2* b = reinterpret_cast<const float2*>(a.in) + 3 * (size_t)i;
            float2 p0 = b[0], p1 = b[1], p2 = b[2];
            bad |= !(isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y) &&
                     isfinite(p2.x) && isfinite(p2.y));
            c[0] = ((double)p0.x + (double)p1.y) * 0.5;
            c[1] = ((double)p0.y + (double)p2.x) * 0.5;
            c[2] = ((double)p1.x + (double)p2.y) * 0.5;
            x[0] = (double)p1.y - (double)p0.x;
            x[1] = (double)p2.x - (double)p0.y;
            x[2] = (double)p2.y - (double)p1.x;
        } else {
            float4 q = reinterpret_cast<const float4*>(a.in)[i];
            bad |= !(isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && isfinite(q.w));
            c[0] = q.x; c[1] = q.y; c[2] = q.z;
            x[0] = x[1] = x[2] = 2.0 * (double)q.w;
        }
        s[0] += c[0]; s[1] += c[1]; s[2] += c[2];
        s[3] += c[0] * c[0]; s[4] += c[0] * c[1]; s[5] += c[0] * c[2];
        s[6] += c[1] * c[1]; s[7] += c[1] * c[2]; s[8] += c[2] * c[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s[9 + k] += x[k];
            e[k] = fmax(e[k], x[k]);
        }
    }
    const uint32_t w = tid >> 6, l = lane_id();
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        double v = wave_sum(s[k]);
        if (l == 0) s_red[w][k] = v;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = e[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if (l == 0) s_red[w][12 + k] = v;
    }
    if (ballot(bad) && l == 0) atomicOr(&a.ctrl->flags, kFlagNonFinite);
    __syncthreads();
    if (tid < 15) {
        double v = 0.0;
        for (int ww = 0; ww < kWaves; ++ww) v = tid < 12 ? v + s_red[ww][tid] : fmax(v, s_red[ww][tid]);
        a.partials[(size_t)tile * kMomStride + tid] = v;
    }
}

// Slab axis: the direction of largest spread orthogonal to the sweep axis u,
// the top eigenvector of P C P with P = I - u u^T (u is in its null space).
__device__ static void slab_axis(const double m[9], uint32_t n, const float uf[3], float v[3]) {
    const double inv = n ? 1.0 / (double)n : 0.0;
    const double mu[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
    double C[3][3];
    C[0][0] = m[3] * inv - mu[0] * mu[0];
    C[0][1] = C[1][0] = m[4] * inv - mu[0] * mu[1];
    C[0][2] = C[2][0] = m[5] * inv - mu[0] * mu[2];
    C[1][1] = m[6] * inv - mu[1] * mu[1];
    C[1][2] = C[2][1] = m[7] * inv - mu[1] * mu[2];
    C[2][2] = m[8] * inv - mu[2] * mu[2];
    const double u[3] = {uf[0], uf[1], uf[2]};
    double P[3][3], PC[3][3], A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = (i == j ? 1.0 : 0.0) - u[i] * u[j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) PC[i][j] = P[i][0] * C[0][j] + P[i][1] * C[1][j] + P[i][2] * C[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = PC[i][0] * P[0][j] + PC[i][1] * P[1][j] + PC[i][2] * P[2][j];
    double V[3][3], ev[3];
    jacobi3(A, V, ev);
    int best = 0;
    for (int k = 1; k < 3; ++k)
        if (ev[k] > ev[best]) best = k;
    double w[3] = {V[0][best], V[1][best], V[2][best]};
    // remove any component along u left by rounding
    const double d = w[0] * u[0] + w[1] * u[1] + w[2] * u[2];
    for (int k = 0; k < 3; ++k) w[k] -= d * u[k];
    double nn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (!(nn > 1e-6)) {  // no spread across u: any unit vector orthogonal to u
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

// One workgroup per scene: the scene's tile partials reduced in a fixed order
// (per-thread strided sums, then the wave/LDS tree), then the axis.  A kernel
// of its own rather than a last-tile ticket: the kernel boundary orders the
// partials, where a ticket needs a device-scope fence per tile (an L2
// write-back on a multi-XCD part).
// Sweep-test model of an axis (pair), from the covariance C, the mean extent
// per coordinate xm and the largest extent per coordinate xx.  A sweep along
// u tests each object against the objects whose lower ends fall inside its
// projected interval: about n^2 E[ext_u] f_u with f_u = 1 / (2 sqrt(pi var_u))
// for a Gaussian (0.98x that for a uniform spread of the same variance).  In
// slabs of width W_v (the largest extent on v) on a second axis v it is about
// n^2 E[ext_u] W_v / (4 pi sqrt(det C_uv)).  The extent of an AABB on a unit
// axis is sum |u_k| ext_k, so tilted axes pay for their longer projections.
// Costs below drop the constant factors; +inf when the spread vanishes.
__device__ static double proj_mean_extent(const double u[3], const double xm[3]) {
    return fabs(u[0]) * xm[0] + fabs(u[1]) * xm[1] + fabs(u[2]) * xm[2];
}
__device__ static double quad(const double C[3][3], const double u[3], const double v[3]) {
    double r = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r += u[i] * C[i][j] * v[j];
    return r;
}
__device__ static double axis_cost(const double C[3][3], const double xm[3], const double u[3]) {
    const double var = quad(C, u, u);
    return var > 0.0 ? proj_mean_extent(u, xm) / sqrt(var) : __builtin_inf();
}
__device__ static double pair_cost(const double C[3][3], const double xm[3], const double xx[3],
                                   const double u[3], const double v[3]) {
    const double det = quad(C, u, u) * quad(C, v, v) - quad(C, u, v) * quad(C, u, v);
    const double wv = fabs(v[0]) * xx[0] + fabs(v[1]) * xx[1] + fabs(v[2]) * xx[2];
    return det > 0.0 ? proj_mean_extent(u, xm) * wv / sqrt(det) : __builtin_inf();
}

// GSAP_AXIS_AUTO: the PCA axis (pair) unless a coordinate axis (pair) has a
// strictly smaller model cost; candidates in a fixed order, so ties keep the
// earlier one.  src = 0 (PCA) or 1 (coordinate).
__device__ static void choose_auto(const double m[15], uint32_t n, bool slabs, float u[3], float v[3],
                                   float* src) {
    choose_axis(m, n, 0, nullptr, nullptr, u, src);
    if (slabs) slab_axis(m, n, u, v);
    if (n < 2) return;
    const double inv = 1.0 / (double)n;
    const double mu[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
    double C[3][3];
    C[0][0] = m[3] * inv - mu[0] * mu[0];
    C[0][1] = C[1][0] = m[4] * inv - mu[0] * mu[1];
    C[0][2] = C[2][0] = m[5] * inv - mu[0] * mu[2];
    C[1][1] = m[6] * inv - mu[1] * mu[1];
    C[1][2] = C[2][1] = m[7] * inv - mu[1] * mu[2];
    C[2][2] = m[8] * inv - mu[2] * mu[2];
    const double xm[3] = {m[9] * inv, m[10] * inv, m[11] * inv};
    const double xx[3] = {m[12], m[13], m[14]};
    const double pu[3] = {u[0], u[1], u[2]};
    int bu = -1, bv = -1;
    if (!slabs) {
        double best = axis_cost(C, xm, pu);
        for (int a = 0; a < 3; ++a) {
            const double e[3] = {a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0};
            const double c = axis_cost(C, xm, e);
            if (c < best) { best = c; bu = a; }
        }
    } else {
        const double pv[3] = {v[0], v[1], v[2]};
        double best = pair_cost(C, xm, xx, pu, pv);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                if (a == b) continue;
                const double e[3] = {a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0};
                const double f[3] = {b == 0 ? 1.0 : 0.0, b == 1 ? 1.0 : 0.0, b == 2 ? 1.0 : 0.0};
                const double c = pair_cost(C, xm, xx, e, f);
                if (c < best) { best = c; bu = a; bv = b; }
            }
    }
    if (bu < 0) return;
    u[0] = u[1] = u[2] = 0.0f;
    u[bu] = 1.0f;
    if (slabs) {
        v[0] = v[1] = v[2] = 0.0f;
        v[bv] = 1.0f;
    }
    *src = 1.0f;
}

__global__ __launch_bounds__(kWG) void k_axis(ObjPipeArgs a) {
    __shared__ double s_red[kWaves][15];
    const uint32_t seg = blockIdx.x, tid = threadIdx.x, w = tid >> 6, l = lane_id();
    const uint32_t t0 = a.seg.tile_seg ? a.seg.seg_tile0[seg] : 0u;
    const uint32_t t1 = a.seg.tile_seg ? a.seg.seg_tile0[seg + 1] : a.ntiles;
    const uint32_t nseg_obj = a.seg.tile_seg ? a.seg.seg_n[seg] : a.n;
    // m[0..11] sums, m[12..14] maxima (exact in any order)
    double m[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) m[k] = 0.0;
    for (uint32_t t = t0 + tid; t < t1; t += kWG)
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            const double p = a.partials[(size_t)t * kMomStride + k];
            m[k] = k < 12 ? m[k] + p : fmax(m[k], p);
        }
#pragma unroll
    for (int k = 0; k < 15; ++k) {
        double v = m[k];
        if (k < 12) {
            v = wave_sum(v);
        } else {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        }
        if (l == 0) s_red[w][k] = v;
    }
    __syncthreads();
    if (tid == 0) {
        double mm[15];
        for (int k = 0; k < 15; ++k) {
            mm[k] = 0.0;
            for (int ww = 0; ww < kWaves; ++ww) mm[k] = k < 12 ? mm[k] + s_red[ww][k] : fmax(mm[k], s_red[ww][k]);
        }
        float u[3], v[3], src;
        const float* learned = a.learned_axis ? a.learned_axis + 4 * (size_t)seg : nullptr;
        if (a.axis_mode == 4) {
            choose_auto(mm, nseg_obj, a.slabs != 0, u, v, &src);
        } else {
            choose_axis(mm, nseg_obj, a.axis_mode, a.fixed_axis, learned, u, &src);
            if (a.slabs) slab_axis(mm, nseg_obj, u, v);
        }
        float* ax = a.axis + 4 * (size_t)seg;
        ax[0] = u[0]; ax[1] = u[1]; ax[2] = u[2]; ax[3] = src;
        if (a.seg_mom)
            for (int k = 0; k < 9; ++k) a.seg_mom[9 * (size_t)seg + k] = mm[k];
        if (a.slabs) {
            float* v2 = a.axis2 + 4 * (size_t)seg;
            v2[0] = v[0]; v2[1] = v[1]; v2[2] = v[2]; v2[3] = 0.0f;
        }
    }
}

// The axis pair of a sweep plane with normal w: a coordinate normal gives the
// two other coordinate axes, the sweep axis the one of larger variance (ties:
// the lower index); any other normal the direction of largest spread within
// the plane (top eigenvector of P C P, P = I - w w^T) and w x that.
__device__ static void plane_pair(const double m[9], uint32_t n, const float wf[3], float u[3],
                                  float v[3]) {
    const double inv = n ? 1.0 / (double)n : 0.0;
    const double mu[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
    double C[3][3];
    C[0][0] = m[3] * inv - mu[0] * mu[0];
    C[0][1] = C[1][0] = m[4] * inv - mu[0] * mu[1];
    C[0][2] = C[2][0] = m[5] * inv - mu[0] * mu[2];
    C[1][1] = m[6] * inv - mu[1] * mu[1];
    C[1][2] = C[2][1] = m[7] * inv - mu[1] * mu[2];
    C[2][2] = m[8] * inv - mu[2] * mu[2];
    int k = 0;
    for (int j = 1; j < 3; ++j)
        if (fabsf(wf[j]) > fabsf(wf[k])) k = j;
    for (int j = 0; j < 3; ++j) u[j] = v[j] = 0.0f;
    if (fabsf(wf[k]) >= 1.0f - 1e-6f) {  // coordinate normal
        const int a0 = k == 0 ? 1 : 0, a1 = k == 2 ? 1 : 2;
        const bool first = C[a0][a0] >= C[a1][a1];
        u[first ? a0 : a1] = 1.0f;
        v[first ? a1 : a0] = 1.0f;
        return;
    }
    double w[3] = {wf[0], wf[1], wf[2]};
    const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int j = 0; j < 3; ++j) w[j] /= wn;
    double P[3][3], PC[3][3], A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = (i == j ? 1.0 : 0.0) - w[i] * w[j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) PC[i][j] = P[i][0] * C[0][j] + P[i][1] * C[1][j] + P[i][2] * C[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = PC[i][0] * P[0][j] + PC[i][1] * P[1][j] + PC[i][2] * P[2][j];
    double V[3][3], ev[3];
    jacobi3(A, V, ev);
    // the eigenvector along w has eigenvalue 0; take the largest of the others
    int best = -1;
    for (int j = 0; j < 3; ++j) {
        const double d = fabs(V[0][j] * w[0] + V[1][j] * w[1] + V[2][j] * w[2]);
        if (d > 0.5) continue;
        if (best < 0 || ev[j] > ev[best]) best = j;
    }
    double x[3];
    if (best >= 0) {
        for (int j = 0; j < 3; ++j) x[j] = V[j][best];
    } else {  // no eigenvector clear of w (cannot happen for a symmetric A): any orthogonal one
        const int e = fabs(w[0]) < 0.9 ? 0 : 1;
        double t[3] = {0.0, 0.0, 0.0};
        t[e] = 1.0;
        x[0] = w[1] * t[2] - w[2] * t[1];
        x[1] = w[2] * t[0] - w[0] * t[2];
        x[2] = w[0] * t[1] - w[1] * t[0];
    }
    const double d = x[0] * w[0] + x[1] * w[1] + x[2] * w[2];
    for (int j = 0; j < 3; ++j) x[j] -= d * w[j];
    const double xn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (int j = 0; j < 3; ++j) x[j] /= xn;
    double y[3] = {w[1] * x[2] - w[2] * x[1], w[2] * x[0] - w[0] * x[2], w[0] * x[1] - w[1] * x[0]};
    const double yn = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    for (int j = 0; j < 3; ++j) {
        u[j] = (float)x[j];
        v[j] = (float)(y[j] / yn);
    }
}

__global__ void k_plane_pairs(const double* seg_mom, const uint32_t* seg_n, uint32_t n, uint32_t nseg,
                              const float* normals, float* axis, float* axis2) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const float* w = normals + 4 * (size_t)s;
    if (!(w[3] > 0.0f)) return;
    double m[9];
    for (int k = 0; k < 9; ++k) m[k] = seg_mom[9 * (size_t)s + k];
    float u[3], v[3];
    plane_pair(m, seg_n ? seg_n[s] : n, w, u, v);
    float* ax = axis + 4 * (size_t)s;
    float* a2 = axis2 + 4 * (size_t)s;
    ax[0] = u[0]; ax[1] = u[1]; ax[2] = u[2]; ax[3] = 3.0f;
    a2[0] = v[0]; a2[1] = v[1]; a2[2] = v[2]; a2[3] = 0.0f;
}

// ---------------------------------------------------------------- slabs
__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Per tile: the largest slab-axis extent and the ranges of both projections.
__global__ __launch_bounds__(kWG) void k_vrange(ObjPipeArgs a) {
    __shared__ float s_r[kWaves][kVr];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x, w = tid >> 6, l = lane_id();
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    const float* ax = a.axis + 4 * (size_t)ti.seg;
    const float* ax2 = a.axis2 + 4 * (size_t)ti.seg;
    const float u[3] = {ax[0], ax[1], ax[2]};
    const float v[3] = {ax2[0], ax2[1], ax2[2]};
    const bool exact = axis_exact(a.kind, u);
    float r[kVr] = {0.0f, __builtin_inff(), -__builtin_inff(), __builtin_inff(), -__builtin_inff()};
    for (uint32_t i = ti.lo + tid; i < ti.hi; i += kWG) {
        float lo[3], hi[3], c4[4], pl, ph, ql, qh;
        load_lohi(a.in, a.kind, i, lo, hi, c4);
        project_lohi(a.kind, lo, hi, c4, u, exact, &pl, &ph);
        project_lohi(a.kind, lo, hi, c4, v, false, &ql, &qh);
        r[0] = fmaxf(r[0], qh - ql);
        r[1] = fminf(r[1], ql);
        r[2] = fmaxf(r[2], qh);
        r[3] = fminf(r[3], pl);
        r[4] = fmaxf(r[4], ph);
    }
    r[0] = wave_maxf(r[0]);
    r[1] = wave_minf(r[1]);
    r[2] = wave_maxf(r[2]);
    r[3] = wave_minf(r[3]);
    r[4] = wave_maxf(r[4]);
    if (l == 0)
        for (int k = 0; k < kVr; ++k) s_r[w][k] = r[k];
    __syncthreads();
    if (tid == 0) {
        for (int ww = 1; ww < kWaves; ++ww) {
            r[0] = fmaxf(r[0], s_r[ww][0]);
            r[1] = fminf(r[1], s_r[ww][1]);
            r[2] = fmaxf(r[2], s_r[ww][2]);
            r[3] = fminf(r[3], s_r[ww][3]);
            r[4] = fmaxf(r[4], s_r[ww][4]);
        }
        for (int k = 0; k < kVr; ++k) a.vr_partials[(size_t)tile * kVr + k] = r[k];
    }
}

// Slab width W >= (1 + 2^-10) x the largest extent, so a box touches at most
// two slabs (the margin covers the rounding of the slab index), and >= the
// range / 254, so there are at most 255 slabs; keys quantise the sweep axis to
// 24 bits over its range (any monotone map keeps the sweep exact).
__global__ __launch_bounds__(kWG) void k_slab_params(ObjPipeArgs a) {
    __shared__ float s_r[kWaves][kVr];
    const uint32_t seg = blockIdx.x;  // one workgroup per scene, over the scene's tiles
    const uint32_t tid = threadIdx.x, w = tid >> 6, l = lane_id();
    const uint32_t t0 = a.seg.tile_seg ? a.seg.seg_tile0[seg] : 0u;
    const uint32_t t1 = a.seg.tile_seg ? a.seg.seg_tile0[seg + 1] : a.ntiles;
    float r[kVr] = {0.0f, __builtin_inff(), -__builtin_inff(), __builtin_inff(), -__builtin_inff()};
    for (uint32_t t = t0 + tid; t < t1; t += kWG) {
        const float* p = a.vr_partials + (size_t)t * kVr;
        r[0] = fmaxf(r[0], p[0]);
        r[1] = fminf(r[1], p[1]);
        r[2] = fmaxf(r[2], p[2]);
        r[3] = fminf(r[3], p[3]);
        r[4] = fmaxf(r[4], p[4]);
    }
    r[0] = wave_maxf(r[0]);
    r[1] = wave_minf(r[1]);
    r[2] = wave_maxf(r[2]);
    r[3] = wave_minf(r[3]);
    r[4] = wave_maxf(r[4]);
    if (l == 0)
        for (int k = 0; k < kVr; ++k) s_r[w][k] = r[k];
    __syncthreads();
    if (tid != 0) return;
    for (int ww = 1; ww < kWaves; ++ww) {
        r[0] = fmaxf(r[0], s_r[ww][0]);
        r[1] = fminf(r[1], s_r[ww][1]);
        r[2] = fmaxf(r[2], s_r[ww][2]);
        r[3] = fminf(r[3], s_r[ww][3]);
        r[4] = fmaxf(r[4], s_r[ww][4]);
    }
    float* sp = a.slab + kSlabParams * (size_t)seg;
    // A coherent update keeps the last refresh's parameters while the slab
    // width still exceeds (1 + 2^-10) x every box's extent: slab and key
    // indices of values outside the old ranges clamp, a monotone map, so the
    // sweep stays exact and no box touches more than two slabs.
    if (a.keep_key && sp[5] > 0.0f && isfinite(r[0]) && r[0] * (1.0f + 0x1p-10f) <= sp[5]) return;
    float v0 = 0.0f, invW = 0.0f, S = 1.0f, W = 0.0f;
    const float span = r[2] - r[1];
    if (isfinite(r[0]) && isfinite(span) && span > 0.0f && r[0] > 0.0f) {
        W = fmaxf(r[0] * (1.0f + 0x1p-10f), span * (1.0f / 254.0f) * (1.0f + 0x1p-10f));
        invW = 1.0f / W;
        if (isfinite(invW) && invW > 0.0f) {
            v0 = r[1];
            S = fminf(floorf(span * invW) + 1.0f, 255.0f);
        } else {
            invW = 0.0f;
        }
    }
    float u0 = 0.0f, idu = 0.0f;
    const float uspan = r[4] - r[3];
    if (isfinite(uspan) && uspan > 0.0f) {
        idu = kSlabQMax / uspan;
        if (isfinite(idu)) u0 = r[3]; else idu = 0.0f;
    }
    sp[0] = v0; sp[1] = invW; sp[2] = S; sp[3] = u0; sp[4] = idu; sp[5] = W; sp[6] = 0.0f; sp[7] = 0.0f;
}


// ---------------------------------------------------------------- projection
// Digit histograms of one key for every pass of the object sort.
__device__ __forceinline__ void hist_key(uint32_t (*s_h)[256], uint32_t key, uint32_t npass,
                                         uint32_t bits) {
    const uint32_t mask = (1u << bits) - 1u;
    for (uint32_t p = 0; p < npass; ++p) atomicAdd(&s_h[p][(key >> (p * bits)) & mask], 1u);
}

__global__ __launch_bounds__(kWG) void k_project(ObjPipeArgs a) {
    if (a.run_if && *a.run_if == 0u) return;  // the coherent sort needed no fallback
    __shared__ uint32_t s_h[kMaxPasses][256];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    const uint32_t npass = a.npass_obj, bits = a.bits_obj, radix = 1u << bits;
    for (uint32_t i = tid; i < kMaxPasses * 256; i += kWG) (&s_h[0][0])[i] = 0u;
    const float* ax = a.axis + 4 * (size_t)ti.seg;
    const float u[3] = {ax[0], ax[1], ax[2]};
    const bool exact = axis_exact(a.kind, u);
    __syncthreads();
    if (!a.slabs) {
        for (uint32_t i = ti.lo + tid; i < ti.hi; i += kWG) {
            float lo[3], hi[3], c4[4];
            load_lohi(a.in, a.kind, i, lo, hi, c4);
            float pl, ph;
            project_lohi(a.kind, lo, hi, c4, u, exact, &pl, &ph);
            const uint32_t key = flip_f32(pl);
            a.keys[i] = key;
            a.vals[i] = i;
            hist_key(s_h, key, npass, bits);
        }
    } else {
        // Box i -> entries 2i (its first slab, flagged) and 2i+1 (the next
        // slab when it straddles a boundary, else an unused entry that sorts
        // last).  Key = slab << kSlabQBits | quantised lower bound on u.
        const float* ax2 = a.axis2 + 4 * (size_t)ti.seg;
        const float v[3] = {ax2[0], ax2[1], ax2[2]};
        const float* sp = a.slab + kSlabParams * (size_t)ti.seg;
        __shared__ uint32_t s_sc[256];  // entries per slab in this tile
        s_sc[tid] = 0u;
        __syncthreads();
        for (uint32_t i = ti.lo + tid; i < ti.hi; i += kWG) {
            float lo[3], hi[3], c4[4], pl, ph, ql, qh;
            load_lohi(a.in, a.kind, i, lo, hi, c4);
            project_lohi(a.kind, lo, hi, c4, u, exact, &pl, &ph);
            project_lohi(a.kind, lo, hi, c4, v, false, &ql, &qh);
            const uint32_t s0 = slab_of(ql, sp);
            uint32_t s1 = slab_of(qh, sp);
            if (s1 > s0 + 1u) {  // cannot happen with W > max extent; fail loudly
                atomicOr(&a.ctrl->flags, kFlagSlab);
                s1 = s0 + 1u;
            }
            const uint32_t q = qkey(pl, sp);
            const uint32_t k0 = (s0 << kSlabQBits) | q;
            const uint32_t k1 = s1 > s0 ? ((s1 << kSlabQBits) | q) : kSlabNull;
            reinterpret_cast<uint2*>(a.keys)[i] = make_uint2(k0, k1);
            reinterpret_cast<uint2*>(a.vals)[i] = make_uint2(i | kFirstBit, i);
            hist_key(s_h, k0, npass, bits);
            hist_key(s_h, k1, npass, bits);
            atomicAdd(&s_sc[s0], 1u);
            if (s1 > s0) atomicAdd(&s_sc[s1], 1u);
        }
        __syncthreads();
        if (s_sc[tid] && !a.no_slab_count) atomicAdd(&a.slab_cnt[256 * (size_t)ti.seg + tid], s_sc[tid]);
    }
    __syncthreads();
    uint32_t* h = a.hist_obj + (size_t)ti.seg * npass * radix;
    for (uint32_t i = tid; i < npass * radix; i += kWG) {
        const uint32_t v = s_h[i / radix][i % radix];
        if (v) atomicAdd(&h[i], v);
    }
}

// ---------------------------------------------------------------- gather
// Sorted position s -> candidate record rec[s]; sweep end (first position
// whose key exceeds s's upper key) by binary search; the sweep-test count.
// grp_end[g]: the sweep ends of each wave's 64 positions minus one, re-ordered
// by sweep end where a window is longer than kShortWindow (then the records
// too, into grp with aux = s, and the top bit of the wave's grp_end words is
// set); the object side of the sweep.
template <int KIND>
__global__ __launch_bounds__(kWG) void k_gather(ObjPipeArgs a, const uint32_t* __restrict__ K,
                                                const uint32_t* __restrict__ V) {
    using Rec = typename std::conditional<KIND == 0, RecAabb, RecSphere>::type;
    __shared__ uint32_t s_sb[257];  // slab mode: first sorted position of each slab (+ end)
    // each wave's next kGatherWin sorted keys: most sweep ends are found
    // there, by a binary search in LDS, instead of a chain of dependent
    // global loads
    __shared__ uint32_t s_kw[kWaves][kGatherWin];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    uint32_t* const kw = s_kw[tid >> 6];
    const float* ax = a.axis + 4 * (size_t)ti.seg;
    const float u[3] = {ax[0], ax[1], ax[2]};
    const bool exact = axis_exact(KIND, u);
    unsigned long long tests = 0;
    const float* sp = a.slab + kSlabParams * (size_t)ti.seg;
    if (a.slabs) {  // a scene's entries are sorted slab-major: slab starts are the prefix of its counts
        __shared__ uint32_t s_tmp[kWaves];
        uint32_t tot;
        const uint32_t c = a.slab_cnt[256 * (size_t)ti.seg + tid];
        const uint32_t ex = block_excl_scan(c, s_tmp, &tot);
        s_sb[tid] = ti.seg_start + ex;
        if (tid == 0) s_sb[256] = ti.seg_start + tot;
        __syncthreads();
    }
    // the next position's value and key are loaded one iteration ahead, so
    // that the box load (which needs the value) is issued at once
    uint32_t vs_n = 0u, key_n = 0u;
    if (ti.lo + tid < ti.hi) {
        vs_n = V[ti.lo + tid];
        key_n = a.slabs ? K[ti.lo + tid] : 0u;
    }
    for (uint32_t sub = ti.lo; sub < ti.hi; sub += kWG) {
        const uint32_t s = sub + tid;
        const bool valid = s < ti.hi;
        const uint32_t vs = vs_n, key = key_n;
        const uint32_t orig = a.slabs ? (vs & ~kFirstBit) : vs;
        // an unused second entry (slab mode) overlaps nothing: its box is
        // not read
        const bool unused = a.slabs && key == kSlabNull;
        float lo[3], hi[3], c4[4] = {0.0f, 0.0f, 0.0f, 0.0f}, pl = 0.0f, ph = 0.0f;
        if (valid && !unused) load_lohi(a.in, KIND, orig, lo, hi, c4);
        if (s + kWG < ti.hi) {
            vs_n = V[s + kWG];
            key_n = a.slabs ? K[s + kWG] : 0u;
        }
        // keys at positions [wb, we): after the wave's first position
        const uint32_t wb = sub + (tid & ~63u) + 1u;
        const uint32_t we = min(wb + kGatherWin, ti.seg_end);
#pragma unroll
        for (uint32_t j = lane_id(); j < kGatherWin; j += 64u)
            if (wb + j < we) kw[j] = K[wb + j];
        __builtin_amdgcn_wave_barrier();
        Rec r;
        uint32_t end = 0xFFFFFFFFu;  // invalid lanes rank last
        if (valid) {
            if (!unused) {
                project_lohi(KIND, lo, hi, c4, u, exact, &pl, &ph);
            } else {
                lo[0] = lo[1] = lo[2] = hi[0] = hi[1] = hi[2] = 0.0f;
            }
            // first t in [s+1, seg_end) with K[t] > hk
            const uint32_t hk = a.slabs ? (((key >> kSlabQBits) << kSlabQBits) | qkey(ph, sp))
                                        : flip_f32(ph);
            uint32_t b = s + 1, e = ti.seg_end;
            if (a.slabs) e = unused ? b : s_sb[(key >> kSlabQBits) + 1];  // within the slab
            const uint32_t le = min(e, we);  // the part of [b, e) held in LDS
            uint32_t lh = le;
            while (b < lh) {
                const uint32_t mid = b + ((lh - b) >> 1);
                if (kw[mid - wb] <= hk) b = mid + 1; else lh = mid;
            }
            if (b == le) {  // every held key is <= hk: the end lies past them
                while (b < e) {
                    const uint32_t mid = b + ((e - b) >> 1);
                    if (K[mid] <= hk) b = mid + 1; else e = mid;
                }
            }
            end = b;
            tests += (unsigned long long)(b - s - 1);
            if constexpr (KIND == 0) {
                r.lo[0] = lo[0]; r.lo[1] = lo[1]; r.lo[2] = lo[2];
                r.hi[0] = hi[0]; r.hi[1] = hi[1]; r.hi[2] = hi[2];
            } else {
                r.c[0] = c4[0]; r.c[1] = c4[1]; r.c[2] = c4[2]; r.c[3] = c4[3];
                r.pad[0] = r.pad[1] = 0u;
            }
            r.orig = vs;  // slab mode: with kFirstBit
            r.aux = 0u;
            if (a.slabs) {
                if (unused) {  // a box that overlaps nothing
                    if constexpr (KIND == 0) {
                        r.lo[0] = r.lo[1] = r.lo[2] = __builtin_inff();
                        r.hi[0] = r.hi[1] = r.hi[2] = -__builtin_inff();
                    } else {
                        r.c[0] = r.c[1] = r.c[2] = __builtin_inff();
                        r.c[3] = 0.0f;
                    }
                    r.orig = orig;
                    r.aux = s;
                } else {
                    // first position of this entry's slab: candidates of another
                    // slab are told apart by it in the sweep
                    r.aux = s_sb[key >> kSlabQBits];
                }
            }
            reinterpret_cast<Rec*>(a.rec)[s] = r;
        }
        // place among the wave's 64 positions ordered by (sweep end,
        // position): the number of smaller keys (keys are unique), one
        // scalar-broadcast compare per lane of the wave.  Only in waves with
        // a window longer than kShortWindow: the sweep takes the other waves
        // one object per lane, in any order, so they keep sorted order.
        const uint32_t lane = lane_id();
        const unsigned long long mk = ((unsigned long long)end << 6) | lane;
        const uint32_t wlen = (valid && end > s + 1u) ? end - s - 1u : 0u;
        uint32_t rank = lane;
        const bool perm = ballot(wlen > kShortWindow) != 0ull;  // uniform
        if (perm) {
            rank = 0;
            const uint32_t mlo = (uint32_t)mk, mhi = (uint32_t)(mk >> 32);
#pragma unroll 8
            for (int j = 0; j < 64; ++j) {
                const uint32_t jlo = (uint32_t)__builtin_amdgcn_readlane((int)mlo, j);
                const uint32_t jhi = (uint32_t)__builtin_amdgcn_readlane((int)mhi, j);
                rank += (((unsigned long long)jhi << 32) | jlo) < mk ? 1u : 0u;
            }
        }
        if (valid) {
            const uint32_t g = s - lane + rank;  // the wave's first position + rank
            // the grouped record only where the order differs from rec's; the
            // top bit of every grp_end word of the wave says which
            if (perm) {
                r.aux = s;
                reinterpret_cast<Rec*>(a.grp)[g] = r;
            }
            a.grp_end[g] = (end - 1u) | (perm ? 0x80000000u : 0u);
        }
    }
    // one atomic per workgroup (the counter is shared by the whole grid)
    __shared__ unsigned long long s_t[kWaves];
    tests = wave_sum(tests);
    if (lane_id() == 0) s_t[tid >> 6] = tests;
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0;
        for (int v = 0; v < kWaves; ++v) t += s_t[v];
        if (t) {
            atomicAdd(&a.ctrl->tests, t);
            if (a.seg_tests) atomicAdd(&a.seg_tests[(size_t)ti.seg * a.seg_tests_stride], t);
        }
    }
}

// ---------------------------------------------------------------- sweep
// One wave per 64 consecutive sorted objects, taken in sweep-end order (the
// gather writes that copy).  Lanes are candidates: the wave walks its
// objects' candidate range from its first position 64 sorted positions at a
// time (coalesced vector loads of the candidate stream), and for each
// still-active object (scalar loads; the wave's 64 objects, 2 KB, stay in the
// scalar cache) every lane runs the exact predicate against its candidate.
// Objects whose sweep end precedes the chunk are retired, so the work is the
// sweep-test count plus under one chunk pair and the object's offset in the
// wave per object.  Candidates past an object's sweep end can never overlap
// it (the projected intervals are conservative); only t > s needs a mask,
// and only in the wave's first chunks.
// Hits are staged per wave in LDS (ballot + mbcnt compaction) and flushed to
// the pair buffer with one atomic claim of exactly the staged count when the
// stage fills; the last wave of the workgroup to finish flushes what every
// wave has left, so the pair buffer holds exactly the pairs, densely.
constexpr uint32_t kStage = kSweepStage;

// SLAB: records are slab entries; a hit also needs the candidate in the
// object's slab (its slab starts at or before the object) and one of the two
// entries to be its box's first slab (the pair's slab: that of the larger
// lower bound), so each pair is found exactly once.
template <int KIND, bool GROUP, int MODE, bool SLAB, bool S4>  // MODE 0 canonical, 1 symmetric, 2 count
__global__ __launch_bounds__(kWG) void k_sweep(SweepArgs a) {
    using Rec = typename std::conditional<KIND == 0, RecAabb, RecSphere>::type;
    // The object records are read-only here; the constant address space lets
    // their wave-uniform loads go through the scalar cache.
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) Rec CRec;
#else
    typedef const Rec CRec;
#endif
    // Staged hits.  An 8-byte entry holds (object, candidate); symmetric mode
    // stages both orientations, two entries per hit.  With S4 (fewer than
    // 2^26 entries) an entry is one word -- the object's slot among the
    // wave's 64 objects above the candidate's index -- and the flush expands
    // it (symmetric mode: into both orientations) through the wave's table of
    // object indices.  Each flush makes one contended slot claim, and the
    // flushes are most of what emitting pairs costs, so the same LDS holding
    // two to three times the hits per flush makes the sweep faster.
    using SE = typename std::conditional<S4, uint32_t, uint2>::type;
    constexpr uint32_t kCap = S4 ? kSweepStage4 : (SLAB ? kSweepStageSlab : kStage);
    constexpr uint32_t kSt = MODE == 2 ? 1u : kCap;
    constexpr uint32_t kPer = (MODE == 1 && !S4) ? 2u : 1u;  // entries per hit
    constexpr uint32_t kOut = (MODE == 1 && S4) ? 2u : 1u;   // pairs written per entry
    __shared__ SE s_stage[kWaves][kSt];
    __shared__ uint32_t s_wobj[kWaves][S4 ? 64 : 1];  // S4: the wave's object indices
    // digit histograms of the pairs written, for the pair sort's first
    // passes; with a.joint: the first digit's histogram, then the per-part
    // second-digit counts as 16-bit halves of kJointWords words
    constexpr uint32_t kJointWords = kMaxRegions * 256u / 2u;
    constexpr uint32_t kPhWords = 256u + kJointWords;
    static_assert(kPhWords >= kSweepHistPasses * 256u, "histogram words");
    __shared__ uint32_t s_ph[MODE == 2 ? 1 : kPhWords];
    __shared__ uint32_t s_fill[kWaves];
    __shared__ uint32_t s_np[kWaves];
    __shared__ unsigned long long s_ex[kWaves];
    __shared__ uint32_t s_done;
    const Rec* __restrict__ R = reinterpret_cast<const Rec*>(a.rec);

    const uint32_t subs = a.tile / kSweepTile;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x, a.xcd_chunk);
    const uint32_t tile = blk / subs, sub = blk % subs;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    const uint32_t lo = ti.lo + sub * kSweepTile;
    if (lo >= ti.hi) return;  // the whole workgroup
    if (threadIdx.x == 0) s_done = 0u;
    if constexpr (MODE != 2) {
        if (a.hist)
            for (uint32_t i = threadIdx.x; i < kPhWords; i += kWG) s_ph[i] = 0u;
    }
    __syncthreads();  // the only block-level barrier
    const uint32_t hi = min(lo + kSweepTile, ti.hi);
    const uint32_t lane = lane_id(), w = uniform(threadIdx.x >> 6);
    const uint32_t g0 = lo + 64u * w;
    // directed pairs when symmetric (also in the pre-count)
    const uint32_t mult = (MODE == 1 || (MODE == 2 && a.sym)) ? 2u : 1u;
    const unsigned long long cap = a.slot_cap;
    SE* stage = s_stage[w];

    uint32_t npairs = 0;  // pairs (or directed pairs) found by this wave (< 64 * 2n)
    uint32_t fill = 0;    // staged entries of this wave
    unsigned long long ex = 0;

    // Pair region of this workgroup: with several regions, the workgroups of
    // each XCD (dealt round-robin) share one allocation counter instead of the
    // whole grid contending for a single one.
    const uint32_t region = a.nregions > 1 ? blockIdx.x % a.nregions : 0u;
    uint2* const rout = a.out + (size_t)region * a.region_stride;
    auto claim = [&](uint32_t cnt) {  // issue the slot claim; the base is read later
        unsigned long long braw = 0;
        if (lane == 0) braw = atomicAdd(&a.ctrl->region_fill[region], (unsigned long long)cnt);
        return braw;
    };
    // Per-part counts: each wave adds at most kJointWave + kCap * kOut to a
    // 16-bit half before it flushes them all (atomic exchanges, so the other
    // waves' concurrent adds are kept), so no half exceeds kWaves *
    // (kJointWave + kCap * kOut) < 2^16 between flushes.
    constexpr uint32_t kJointWave = 65535u / kWaves - kCap * kOut - 1u;
    static_assert(kWaves * (kJointWave + kCap * kOut) < 65536u, "16-bit joint counts");
    uint32_t jcount = 0;  // pairs this wave counted since it last flushed them
    auto flush_joint = [&]() {
        if constexpr (MODE != 2) {
            for (uint32_t i = lane; i < kJointWords; i += 64u) {
                const uint32_t v = atomicExch(&s_ph[256u + i], 0u);
                if (v & 0xFFFFu) atomicAdd(&a.joint[2u * i], v & 0xFFFFu);
                if (v >> 16) atomicAdd(&a.joint[2u * i + 1u], v >> 16);
            }
        }
    };
    // entry i of a stage -> (object, candidate); tab: that wave's object table
    auto entry = [&](const SE* st, const uint32_t* tab, uint32_t i) {
        if constexpr (S4) {
            const uint32_t e = st[i];
            return make_uint2(tab[e >> kStage4Shift], e & kStage4Mask);
        } else {
            (void)tab;
            return st[i];
        }
    };
    auto count_key = [&](uint2 pr) {
        const unsigned long long k = ((unsigned long long)pr.x << a.bbits) | (unsigned long long)pr.y;
        if (a.joint) {
            const uint32_t d1 = (uint32_t)(k >> a.hist_shift0) & 255u;
            const uint32_t j = ((d1 >> 5) << 8) | ((uint32_t)(k >> (a.hist_shift0 + 8)) & 255u);
            atomicAdd(&s_ph[d1], 1u);
            atomicAdd(&s_ph[256u + (j >> 1)], 1u << ((j & 1u) << 4));
        } else {
            for (int q = 0; q < a.hist_npass; ++q)
                atomicAdd(&s_ph[q * 256 + (uint32_t)((k >> (a.hist_shift0 + 8 * q)) & 255u)], 1u);
        }
    };
    auto count_digits = [&](const SE* st, const uint32_t* tab, uint32_t cnt) {
        if constexpr (MODE != 2) {
            if (a.hist) {
                for (uint32_t i = lane; i < cnt; i += 64u) {
                    uint2 pr = entry(st, tab, i);
                    if constexpr (MODE == 0) pr = make_uint2(min(pr.x, pr.y), max(pr.x, pr.y));
                    count_key(pr);
                    if constexpr (kOut == 2) count_key(make_uint2(pr.y, pr.x));
                }
                if (a.joint) {
                    jcount += cnt * kOut;
                    if (jcount > kJointWave) {
                        flush_joint();
                        jcount = 0;
                    }
                }
            }
        }
    };
    auto store = [&](const SE* st, const uint32_t* tab, uint32_t cnt, unsigned long long base) {
        for (uint32_t i = lane; i < cnt; i += 64u) {
            uint2 pr = entry(st, tab, i);
            if constexpr (MODE == 0) pr = make_uint2(min(pr.x, pr.y), max(pr.x, pr.y));
            if constexpr (kOut == 2) {
                const unsigned long long s = base + 2ull * i;
                if (s < cap) rout[s] = pr;
                if (s + 1u < cap) rout[s + 1u] = make_uint2(pr.y, pr.x);
            } else {
                if (base + i < cap) rout[base + i] = pr;
            }
        }
    };
    auto first_lane = [](unsigned long long v) {
        return (unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)v) |
               ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32);
    };
    // Write the staged entries st[0, cnt) of a wave (cnt > 0, wave-uniform):
    // claim the slots of their pairs in the workgroup's region, count the
    // pairs' sort digits in LDS while the claim is in flight, then store them
    // (within the capacity).  Canonical mode stages (object, candidate) and
    // orders the pair here, once per pair rather than once per tested chunk.
    // (After a capacity overflow the digit counts include unwritten pairs;
    // the update then reports the overflow and skips the pair sort.)
    auto write_out = [&](const SE* st, const uint32_t* tab, uint32_t cnt) {
        const unsigned long long braw = claim(cnt * kOut);
        count_digits(st, tab, cnt);
        store(st, tab, cnt, first_lane(braw));
    };
    const uint32_t* const wtab = s_wobj[w];
    auto flush = [&]() {  // write this wave's stage out (fill > 0)
        write_out(stage, wtab, fill);
        npairs += fill * kOut;
        fill = 0;
        __builtin_amdgcn_wave_barrier();
    };
    // stage this lane's hit (object in wave slot q, candidate) as entry r
    auto put = [&](uint32_t r, uint32_t q, uint32_t oorig, uint32_t corig) {
        if constexpr (S4) {
            (void)oorig;
            stage[r] = (q << kStage4Shift) | corig;
        } else {
            (void)q;
            stage[r] = make_uint2(oorig, corig);
            if constexpr (MODE == 1) stage[r + 1] = make_uint2(corig, oorig);
        }
    };

    if (g0 < hi) {
        const uint32_t nobj = uniform(min(64u, hi - g0));
        // sweep ends (stored minus one, top bit: the wave was regrouped, and
        // its records are in grp; ascending in lane order then)
        const uint32_t gw = lane < nobj ? a.grp_end[g0 + lane] : 0u;
        const bool perm = ballot(lane < nobj && (gw >> 31) != 0u) != 0ull;
        const uint32_t ge = lane < nobj ? (gw & 0x7FFFFFFFu) + 1u : 0u;
        const uint32_t wend = uniform(wave_max_u32(ge));
        // Short windows (every object of the wave has at most kShortWindow
        // candidates: sparse scenes, slabs of small scenes): each lane takes
        // its own object and walks its window, one candidate per step (the
        // lanes' candidates are close together in sorted order), instead of
        // testing every object against whole 128-candidate chunk pairs.  Hits
        // are staged by ballot exactly as below; in slab mode a window never
        // leaves its slab, so only the first-slab rule applies.
        Rec me{};
        uint32_t my_s = 0u, len = 0u;
        if (lane < nobj) {
            if (perm) {
                me = reinterpret_cast<const Rec*>(a.grp)[g0 + lane];
                my_s = me.aux;
            } else {  // not regrouped: every window is short, and rec has the records
                me = R[g0 + lane];
                my_s = g0 + lane;
            }
            len = ge > my_s + 1u ? ge - my_s - 1u : 0u;
        }
        if constexpr (S4) s_wobj[w][lane] = SLAB ? (me.orig & ~kFirstBit) : me.orig;
        const uint32_t maxlen = uniform(wave_max_u32(len));
        if (maxlen <= kShortWindow) {
            const uint32_t myorig = SLAB ? (me.orig & ~kFirstBit) : me.orig;
            const uint32_t mygrp = (GROUP && lane < nobj) ? a.group[myorig] : 0u;
            // candidates four at a time: their loads are issued together, so
            // the wave waits on one load latency per four candidates
            for (uint32_t k0 = 1u; k0 <= maxlen; k0 += 4u) {
            Rec cb[4];
#pragma unroll
            for (uint32_t j = 0; j < 4u; ++j)
                if (k0 + j <= len) cb[j] = R[my_s + k0 + j];
#pragma unroll
            for (uint32_t j = 0; j < 4u; ++j) {
                const uint32_t k = k0 + j;
                if (k > maxlen) break;  // uniform
                const bool act = k <= len;
                bool hit = false;
                uint32_t corig = 0u;
                if (act) {
                    const Rec& c = cb[j];
                    if constexpr (KIND == 0) {
                        const float d0 = c.lo[0] - me.hi[0], d1 = me.lo[0] - c.hi[0];
                        const float d2 = c.lo[1] - me.hi[1], d3 = me.lo[1] - c.hi[1];
                        const float d4 = c.lo[2] - me.hi[2], d5 = me.lo[2] - c.hi[2];
                        const float mx = fmaxf(fmaxf(fmaxf(d0, d1), d2), fmaxf(fmaxf(d3, d4), d5));
                        hit = mx <= 0.0f;
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
                    corig = SLAB ? (c.orig & ~kFirstBit) : c.orig;
                    if constexpr (SLAB) hit = hit && ((me.orig | c.orig) & kFirstBit) != 0u;
                    if constexpr (GROUP) hit = hit && a.group[corig] != mygrp;
                }
                const unsigned long long m = ballot(hit);
                if (m == 0ull) continue;
                if constexpr (MODE == 2) {
                    npairs += (uint32_t)__popcll(m) * mult;
                } else {
                    const uint32_t cnt = (uint32_t)__popcll(m) * kPer;  // entries
                    if (fill + cnt > kCap) flush();  // stage full: write it out
                    if (hit) put(fill + kPer * mbcnt(m), lane, myorig, corig);
                    fill += cnt;
                    __builtin_amdgcn_wave_barrier();
                }
            }
            }
            ex = wave_sum((unsigned long long)(lane < nobj ? maxlen : 0u));
        } else {
        CRec* __restrict__ O = (CRec*)(a.grp) + g0;
        // Slab mode: a candidate past an object's sweep end may belong to the
        // next slab (and overlap it).  That only happens in a wave whose
        // objects are not all in one slab (all windows then end inside that
        // slab); such a "mixed" wave also tests candidate slab <= object slab.
        bool mixed = false;
        if constexpr (SLAB) {
            uint32_t sb = 0u, sbn = 0u;
            if (lane < nobj) {
                sb = R[reinterpret_cast<const Rec*>(a.grp)[g0 + lane].aux].aux;
                sbn = ~sb;
            } else {
                sb = 0u;
                sbn = 0u;
            }
            // max of sb and of ~sb over the valid lanes: all equal iff max(sb) == ~max(~sb)
            mixed = uniform(wave_max_u32(sb)) != ~uniform(wave_max_u32(sbn));
        }

        uint32_t kfirst = 0;
        // Candidate chunks are double-buffered: the next chunk's records are
        // requested before this chunk's objects are tested.
        // Candidates are taken two 64-record chunks at a time, so that each
        // of the wave's object records is read (through the scalar cache)
        // once per 128 candidates; the next pair of chunks is requested
        // before this pair's objects are tested.
        struct Chunk {
            Rec c;
            uint32_t t, corig, gc;
            unsigned long long cfm;  // lanes whose candidate entry is its box's first slab
        };
        auto chunk = [&](Rec c, uint32_t t) {
            if (t >= wend) {  // a candidate that overlaps nothing
                if constexpr (KIND == 0) {
                    c.lo[0] = c.lo[1] = c.lo[2] = __builtin_inff();
                    c.hi[0] = c.hi[1] = c.hi[2] = -__builtin_inff();
                } else {
                    c.c[0] = __builtin_inff();
                }
            }
            Chunk k;
            k.c = c;
            k.t = t;
            k.corig = SLAB ? (c.orig & ~kFirstBit) : c.orig;
            k.gc = GROUP ? a.group[k.corig] : 0u;
            k.cfm = SLAB ? ballot((c.orig & kFirstBit) != 0u) : 0ull;
            return k;
        };
        Rec cn{}, cn2{};
        // the wave's candidates start after its first position
        if (g0 + 1u < wend) {
            cn = R[min(g0 + 1u + lane, wend - 1u)];
            cn2 = R[min(g0 + 65u + lane, wend - 1u)];
        }
        for (uint32_t t0 = g0 + 1; t0 < wend; t0 += 128u) {
            // retire objects whose sweep end is at or before this chunk pair
            kfirst = (uint32_t)__popcll(ballot(lane < nobj && ge <= t0));
            const uint32_t t = t0 + lane;
            const Rec c1r = cn, c2r = cn2;
            if (t0 + 128u < wend) {
                cn = R[min(t + 128u, wend - 1u)];
                cn2 = R[min(t + 192u, wend - 1u)];
            }
            const bool two = t0 + 64u < wend;  // uniform
            const Chunk k1 = chunk(c1r, t);
            const Chunk k2 = chunk(c2r, t + 64u);

            // The exact test is the only per-lane arithmetic on the common path;
            // the other conditions are lane masks (scalar ANDs) or uniform branches.
            // oorig, fsel (all lanes when the object's entry is its box's first
            // slab, else none) and q (the object's slot in the wave) are per object.
            auto one = [&](const Rec& o, uint32_t q, uint32_t oorig, unsigned long long fsel,
                           const Chunk& ck, auto masked_tag, auto mixed_tag) {
                constexpr bool MASKED = decltype(masked_tag)::value;
                constexpr bool MIXED = decltype(mixed_tag)::value;
                const Rec& c = ck.c;
                const uint32_t t = ck.t, corig = ck.corig, gc = ck.gc;
                const unsigned long long cfm = ck.cfm;
                bool hit;  // this lane's test; the wave works with masks below
                if constexpr (KIND == 0) {
                    // Inclusive AABB test (aabb.h:47-57) on the vector ALU only:
                    // for finite floats with denormals preserved (the kernel's f32
                    // denorm mode), fl(x - y) <= 0 exactly when x <= y, so the six
                    // compares become six subtractions, a max tree and one compare
                    // -- no scalar-ALU mask arithmetic, which is the scarcer unit
                    // (one scalar ALU per CU vs four SIMDs).
                    const float d0 = c.lo[0] - o.hi[0], d1 = o.lo[0] - c.hi[0];
                    const float d2 = c.lo[1] - o.hi[1], d3 = o.lo[1] - c.hi[1];
                    const float d4 = c.lo[2] - o.hi[2], d5 = o.lo[2] - c.hi[2];
                    const float mx = fmaxf(fmaxf(fmaxf(d0, d1), d2), fmaxf(fmaxf(d3, d4), d5));
                    hit = mx <= 0.0f;
                } else {
#pragma clang fp contract(off)
                    const float x0 = o.c[0] - c.c[0];
                    const float y0 = o.c[1] - c.c[1];
                    const float z0 = o.c[2] - c.c[2];
                    const float r0 = o.c[3] + c.c[3];
                    const float t1 = (x0 * x0 + y0 * y0) + z0 * z0;
                    const float e1 = r0 * r0;
                    hit = t1 < e1;
                }
                // The wave's hit mask is the AND of the ballots of the compares
                // (scalar ANDs; a ballot of a combined bool would be rebuilt
                // through VGPRs), and the lanes that stage a pair are selected
                // by the mask itself (inverse ballot: the mask becomes exec).
                unsigned long long m = ballot(hit);
                if (MASKED) m &= ballot(t > o.aux);  // candidates after the object
                if constexpr (SLAB) {
                    if (MIXED) m &= ballot(c.aux <= o.aux);  // candidate in the object's slab
                    m &= fsel | cfm;  // one of them is a first slab
                }
                if (GROUP) m &= ballot(a.group[oorig] != gc);
                if (m == 0ull) return;
                if constexpr (MODE == 2) {
                    npairs += (uint32_t)__popcll(m) * mult;
                } else {
                    const uint32_t cnt = (uint32_t)__popcll(m) * kPer;  // entries
                    if (fill + cnt > kCap) flush();  // stage full: write it out
                    if (__builtin_amdgcn_inverse_ballot_w64(m)) put(fill + kPer * mbcnt(m), q, oorig, corig);
                    fill += cnt;
                    __builtin_amdgcn_wave_barrier();
                }
            };
            auto process = [&](const Rec& o, uint32_t q, auto masked_tag, auto mixed_tag) {
                const uint32_t oorig = SLAB ? (o.orig & ~kFirstBit) : o.orig;
                const unsigned long long fsel = (SLAB && !(o.orig & kFirstBit)) ? 0ull : ~0ull;
                one(o, q, oorig, fsel, k1, masked_tag, mixed_tag);
                if (two) one(o, q, oorig, fsel, k2, masked_tag, mixed_tag);
            };
            auto objects = [&](auto masked_tag, auto mixed_tag) {
                uint32_t q = kfirst;
                for (; q + 4u <= nobj; q += 4u) {  // four scalar loads in flight
                    const Rec o0 = O[q], o1 = O[q + 1u], o2 = O[q + 2u], o3 = O[q + 3u];
                    process(o0, q, masked_tag, mixed_tag);
                    process(o1, q + 1u, masked_tag, mixed_tag);
                    process(o2, q + 2u, masked_tag, mixed_tag);
                    process(o3, q + 3u, masked_tag, mixed_tag);
                }
                for (; q < nobj; ++q) {
                    const Rec o = O[q];
                    process(o, q, masked_tag, mixed_tag);
                }
            };
            // an object may sit at or after the chunk start only in the first
            // chunks of the wave (its objects lie in [g0, g0 + nobj))
            const bool early = t0 < g0 + nobj;  // uniform
            if (SLAB && mixed) {
                if (early) objects(std::true_type{}, std::true_type{});
                else objects(std::false_type{}, std::true_type{});
            } else {
                if (early) objects(std::true_type{}, std::false_type{});
                else objects(std::false_type{}, std::false_type{});
            }
        }
        // candidate tests executed: every active object takes whole pairs of
        // 64-lane chunks (an upper bound: the wave's last pair may be single)
        if (lane < nobj && ge > g0 + 1) ex = (unsigned long long)((ge - g0 - 1 + 127u) / 128u) * 128u;
        ex = wave_sum(ex);
        }  // long windows
    }

    // Hand this wave's remainder and counters to the workgroup; the last wave
    // to arrive writes them all out with one atomic per counter.
    if constexpr (MODE != 2) npairs += fill * kOut;  // the staged remainder
    if (lane == 0) {
        s_fill[w] = fill;
        s_np[w] = npairs;
        s_ex[w] = ex;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    uint32_t old = 0;
    if (lane == 0) old = atomicAdd(&s_done, 1u);
    old = __builtin_amdgcn_readfirstlane(old);
    if (old != kWaves - 1u) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    unsigned long long np = 0, ext = 0;
    uint32_t tot = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
        np += s_np[v];
        ext += s_ex[v];
        tot += s_fill[v];
    }
    if (lane == 0) {
        // pairs of a writing sweep are the sum of its slot claims (k_pair_regions)
        if (MODE == 2 && np) {
            atomicAdd(&a.ctrl->part[0][blockIdx.x % kMaxRegions], np);
            if (a.seg_pairs) atomicAdd(&a.seg_pairs[ti.seg], np);
        }
        if (ext) atomicAdd(&a.ctrl->part[1][blockIdx.x % kMaxRegions], ext);
    }
    if constexpr (MODE != 2) {
        if (tot) {
            const unsigned long long braw = claim(tot * kOut);
#pragma unroll
            for (int v = 0; v < kWaves; ++v) count_digits(s_stage[v], s_wobj[v], s_fill[v]);
            unsigned long long base = first_lane(braw);
#pragma unroll
            for (int v = 0; v < kWaves; ++v) {
                const uint32_t f = s_fill[v];
                store(s_stage[v], s_wobj[v], f, base);
                base += (unsigned long long)f * kOut;
            }
        }
        if (a.hist) {  // every wave's digits are counted (they arrived after counting)
            const uint32_t nh = a.joint ? 256u : (uint32_t)a.hist_npass * 256u;
            for (uint32_t i = lane; i < nh; i += 64u) {
                const uint32_t c = s_ph[i];
                if (c) {
                    atomicAdd(&a.hist[i], c);
                    // the first digit also per region, for the sort's first pass
                    if (a.hist_reg && i < 256u) atomicAdd(&a.hist_reg[region * 256u + i], c);
                }
            }
            if (a.joint) flush_joint();
        }
    }
}

// ---------------------------------------------------------------- CSR entries
// Row starts of the sorted directed pair list (reference pairEntries).
__global__ void k_rowflag(const uint2* __restrict__ P, const unsigned long long* n_dev,
                          unsigned long long cap, uint32_t nobj, uint32_t* rowstart, uint32_t* has) {
    unsigned long long m = *n_dev;
    if (m > cap) m = 0;  // capacity overflow: the pair sort wrote nothing
    for (unsigned long long k = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; k < m;
         k += (unsigned long long)gridDim.x * blockDim.x) {
        const uint32_t x = P[k].x;
        if (x < nobj && (k == 0 || P[k - 1].x != x)) {
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
    const uint32_t b0 = blockIdx.x * 2048u, tid = threadIdx.x;
    // this block's base: the earlier blocks' totals, summed by all threads
    uint32_t base = 0;
    for (uint32_t b = tid; b < blockIdx.x; b += kWG) base += block_sums[b];
    base = wave_sum(base);
    if (lane_id() == 0) s_tmp[tid >> 6] = base;
    __syncthreads();
    base = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) base += s_tmp[w];
    __syncthreads();  // s_tmp is the scans' scratch below
    if (tid == 0 && blockIdx.x == gridDim.x - 1) ctrl->num_entries = base + block_sums[blockIdx.x];
    uint32_t run = base;
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

__global__ void k_finish(Ctrl* ctrl, unsigned long long cap, unsigned long long* d_counts,
                         const float* axis, unsigned long long* axis_out) {
    if (threadIdx.x == 0) {
        unsigned long long np = ctrl->num_pairs, ex = 0;
        for (int r = 0; r < kMaxRegions; ++r) {
            np += ctrl->part[0][r];
            ex += ctrl->part[1][r];
            ctrl->part[0][r] = 0;
            ctrl->part[1][r] = 0;
        }
        uint32_t fl = ctrl->flags;
        if (np > cap) fl |= kFlagCapacity;
        d_counts[0] = np;
        d_counts[1] = ctrl->num_entries;
        d_counts[2] = ctrl->tests;
        d_counts[3] = fl;
        d_counts[4] = ex;
        if (axis_out) {  // the (first) scene's axis and its source, as float bits
            axis_out[0] = (unsigned long long)__float_as_uint(axis[0]) |
                          ((unsigned long long)__float_as_uint(axis[1]) << 32);
            axis_out[1] = (unsigned long long)__float_as_uint(axis[2]) |
                          ((unsigned long long)__float_as_uint(axis[3]) << 32);
        }
        ctrl->num_pairs = 0;
        ctrl->num_entries = 0;
        ctrl->tests = 0;
        for (int r = 0; r < kMaxRegions; ++r) ctrl->region_fill[r] = 0;
        ctrl->flags = 0;
    }
}

// One workgroup: region layout for the pair sort (see launch_pair_regions).
__global__ void k_pair_regions(Ctrl* ctrl, unsigned long long region_cap, uint32_t nregions,
                               uint32_t tile) {
    if (threadIdx.x != 0) return;
    // A region is rounded up to whole tiles, so a tile of the first sort pass
    // lies in one region, and its slots past region_valid are padding that no
    // pass reads.  After a capacity overflow (more pairs found than the
    // regions hold, region_cap together) the sort is skipped: every region is
    // empty.  The digit counts then include pairs that were not written, so a
    // pass would place the written ones by counts that do not match them and
    // leave an unsorted range, while the local stage's window search relies on
    // sorted buckets.
    unsigned long long found = 0;
    for (uint32_t r = 0; r < nregions; ++r) found += ctrl->region_fill[r];
    ctrl->num_pairs = found;  // every pair the sweep found was claimed a slot
    const bool skip = found > region_cap;
    unsigned long long run = 0, real = 0;
    for (uint32_t r = 0; r < nregions; ++r) {
        const unsigned long long f = skip ? 0ull : ctrl->region_fill[r];
        const unsigned long long p = nregions > 1 ? (f + tile - 1) / tile * tile : f;
        ctrl->region_start[r] = run;
        ctrl->region_valid[r] = run + f;
        run += p;
        real += f;
    }
    for (uint32_t r = nregions; r <= (uint32_t)kMaxRegions; ++r) ctrl->region_start[r] = run;
    for (uint32_t r = nregions; r < (uint32_t)kMaxRegions; ++r) ctrl->region_valid[r] = run;
    ctrl->sort_n = run;
    ctrl->sort_real = real;
}

// Per-scene ranges of the sorted (scene-major) canonical pair list.
__global__ void k_scene_offsets(const uint2* __restrict__ P, const Ctrl* ctrl, unsigned long long cap,
                                const uint32_t* __restrict__ seg_start, uint32_t nseg,
                                unsigned long long* out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nseg) return;
    unsigned long long m = ctrl->num_pairs;
    if (m > cap) m = 0;  // capacity overflow: the pair sort wrote nothing
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
hipError_t launch_slab_params(const ObjPipeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_vrange, dim3(a.ntiles), dim3(kWG), 0, st, a);
    hipLaunchKernelGGL(k_slab_params, dim3(a.seg.tile_seg ? a.seg.nseg : 1u), dim3(kWG), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_moments(const ObjPipeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_moments, dim3(a.ntiles), dim3(kWG), 0, st, a);
    const uint32_t nseg = a.seg.tile_seg ? a.seg.nseg : 1u;
    if (!a.keep_key) hipLaunchKernelGGL(k_axis, dim3(nseg), dim3(kWG), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_plane_pairs(const double* seg_mom, const uint32_t* seg_n, uint32_t n, uint32_t nseg,
                              const float* normals, float* axis, float* axis2, hipStream_t st) {
    if (nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_plane_pairs, dim3((nseg + 255) / 256), dim3(256), 0, st, seg_mom, seg_n, n,
                       nseg, normals, axis, axis2);
    return hipGetLastError();
}
hipError_t launch_project(const ObjPipeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_project, dim3(a.ntiles), dim3(kWG), 0, st, a);
    return hipGetLastError();
}
hipError_t launch_gather(const ObjPipeArgs& a, const uint32_t* K, const uint32_t* V, hipStream_t st) {
    if (a.kind == 0)
        hipLaunchKernelGGL(k_gather<0>, dim3(a.ntiles), dim3(kWG), 0, st, a, K, V);
    else
        hipLaunchKernelGGL(k_gather<1>, dim3(a.ntiles), dim3(kWG), 0, st, a, K, V);
    return hipGetLastError();
}

template <int KIND, bool GROUP, int MODE>
static void sweep_launch(const SweepArgs& a, hipStream_t st) {
    const uint32_t blocks = a.ntiles * (a.tile / kSweepTile);
    // one-word stage entries while every candidate index fits below the
    // object slot (entry indices bound the box indices)
    const bool s4 = MODE != 2 && a.n <= kStage4Mask;
    if (a.slabs) {
        if (s4)
            hipLaunchKernelGGL((k_sweep<KIND, GROUP, MODE, true, true>), dim3(blocks), dim3(kWG), 0, st, a);
        else
            hipLaunchKernelGGL((k_sweep<KIND, GROUP, MODE, true, false>), dim3(blocks), dim3(kWG), 0, st, a);
    } else {
        if (s4)
            hipLaunchKernelGGL((k_sweep<KIND, GROUP, MODE, false, true>), dim3(blocks), dim3(kWG), 0, st, a);
        else
            hipLaunchKernelGGL((k_sweep<KIND, GROUP, MODE, false, false>), dim3(blocks), dim3(kWG), 0, st, a);
    }
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
    hipLaunchKernelGGL(k_rowflag, dim3(grid), dim3(256), 0, st, P, n_dev, cap, nobj, rowstart, has);
    const uint32_t nb = (nobj + 2047) / 2048;
    if (nb == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_entries_count, dim3(nb), dim3(kWG), 0, st, has, nobj, block_sums);
    hipLaunchKernelGGL(k_entries_write, dim3(nb), dim3(kWG), 0, st, has, rowstart, nobj, block_sums,
                       entries, ctrl);
    return hipGetLastError();
}

hipError_t launch_finish(Ctrl* ctrl, unsigned long long cap, unsigned long long* d_counts, int sym,
                         const float* axis, unsigned long long* axis_out, hipStream_t st) {
    (void)sym;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, st, ctrl, cap, d_counts, axis, axis_out);
    return hipGetLastError();
}

hipError_t launch_pair_regions(Ctrl* ctrl, unsigned long long region_cap, uint32_t nregions,
                               uint32_t tile, hipStream_t st) {
    hipLaunchKernelGGL(k_pair_regions, dim3(1), dim3(64), 0, st, ctrl, region_cap, nregions, tile);
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
