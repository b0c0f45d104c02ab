// k_mlp.hip -- learned sweep-axis selector on gfx950.
//
// No reference counterpart: the reference picks the coordinate axis of largest
// variance (physicboxSap2.16_Fermi/sapDetection/SapKernel.cu:62-97) and counts
// per-object sweep work in CounterThreadSphPCATexSub (:1429-1487).  Here:
//   k_features     64 scale-free features per scene from the fp64 moments and
//                  per-tile sums of the boxes' half-extents
//   k_axis_tests   the judge: exact sweep-test count of a scene for each of
//                  the candidate axes (in-LDS bitonic sort of its keys), plus
//                  the scene's pair count -- training labels
//   k_mlp_fwd/bwd  a 64-128-128-4 MLP, bf16 weights and activations, fp32
//                  accumulation, trained with Adam by these kernels
//   k_axis_infer   inference per scene, snapped to the nearest candidate axis
//                  (slab mode: the nearest coordinate axis, the normal of the
//                  sweep plane); low-confidence scenes keep their PCA axis (pair)
// The axis never changes the pair set, only the sweep work.
#include "gsap_device.h"

namespace gsap {

// parameter layout of the fp32 master copy (row-major [in][out])
constexpr uint32_t kOffW1 = 0, kOffB1 = kOffW1 + kFeat * kHid, kOffW2 = kOffB1 + kHid,
                   kOffB2 = kOffW2 + kHid * kHid, kOffW3 = kOffB2 + kHid,
                   kOffB3 = kOffW3 + kHid * kOut, kParams = kOffB3 + kOut;
static_assert(kParams == 25348, "layout");

__device__ __forceinline__ uint16_t to_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------------ features
// Per-tile sums (fp32 within a tile, fp64 across tiles):
//  0-2 h_k, 3-5 h_k^2, 6-8 |c_k - mu_k|, 9-11 (c_k - mu_k)^4, 12-35 3x8 histogram
//  of log2(h_k / l) in [-4, 4) (l = mean spacing), 36-38 h_x h_y, h_x h_z, h_y h_z.
__device__ static void jacobi3f(double A[3][3], double V[3][3]) {
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
}

struct SceneStats {
    double mu[3], C[3][3], s2, ell;
};
__device__ static SceneStats scene_stats(const double* m, uint32_t n) {
    SceneStats S;
    const double inv = n ? 1.0 / (double)n : 0.0;
    for (int k = 0; k < 3; ++k) S.mu[k] = m[k] * inv;
    S.C[0][0] = m[3] * inv - S.mu[0] * S.mu[0];
    S.C[0][1] = S.C[1][0] = m[4] * inv - S.mu[0] * S.mu[1];
    S.C[0][2] = S.C[2][0] = m[5] * inv - S.mu[0] * S.mu[2];
    S.C[1][1] = m[6] * inv - S.mu[1] * S.mu[1];
    S.C[1][2] = S.C[2][1] = m[7] * inv - S.mu[1] * S.mu[2];
    S.C[2][2] = m[8] * inv - S.mu[2] * S.mu[2];
    S.s2 = S.C[0][0] + S.C[1][1] + S.C[2][2];
    if (!(S.s2 > 1e-30)) S.s2 = 1e-30;
    S.ell = sqrt(S.s2) * pow((double)(n ? n : 1), -1.0 / 3.0);  // mean spacing
    return S;
}

constexpr uint32_t kFeatGroup = 8;  // sort tiles per feature workgroup

__global__ __launch_bounds__(kWG) void k_features(ObjPipeArgs a, float* tile_sums, float*) {
    __shared__ float s_red[kWaves][kFeatSums];
    // 16 copies of the density histograms, lane l adding to copy l mod 16:
    // most of a wave's centres fall in a few central bins, and LDS atomics on
    // one address serialise.  A copy is kDensWords + 1 words long, so the
    // copies of a bin lie in different LDS banks (kDensWords is a multiple of
    // the 32 banks).
    constexpr uint32_t kCopies = 16;
    __shared__ uint32_t s_pl[kCopies][kDensWords + 1];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x, l = lane_id(), w = tid >> 6;
    const TileInfo ti = tile_info(a.seg, a.n, tile, a.tile);
    // one workgroup per kFeatGroup sort tiles of a scene (the first one's):
    // fewer partial histograms to add up
    const uint32_t t0 = a.seg.tile_seg ? a.seg.seg_tile0[ti.seg] : 0u;
    if ((tile - t0) % kFeatGroup != 0u) return;
    const uint32_t i_end = min(ti.lo + kFeatGroup * a.tile, ti.seg_end);
    const uint32_t n = ti.seg_end - ti.seg_start;
    const SceneStats S = scene_stats(a.seg_mom + 9 * (size_t)ti.seg, n);
    for (uint32_t i = tid; i < kCopies * (kDensWords + 1); i += kWG) (&s_pl[0][0])[i] = 0u;
    uint32_t* const pl = s_pl[l & (kCopies - 1u)];
    float acc[kFeatSums];
#pragma unroll
    for (int k = 0; k < kFeatSums; ++k) acc[k] = 0.0f;
    const float inv_ell = (float)(1.0 / S.ell);
    unsigned long long ext8[3] = {0ull, 0ull, 0ull};  // extent classes, 8 x 8-bit counts per axis
    // density grids: bins over mean +- 2 standard deviations per axis
    float gs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double sd = sqrt(S.C[k][k] > 1e-30 ? S.C[k][k] : 1e-30);
        gs[k] = (float)(1.0 / (4.0 * sd));
    }
    __syncthreads();
    for (uint32_t i = ti.lo + tid; i < i_end; i += kWG) {
        float lo[3], hi[3], c4[4];
        load_lohi(a.in, a.kind, i, lo, hi, c4);
        float h[3], d[3];
        int g[3], g1[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            h[k] = 0.5f * (hi[k] - lo[k]);
            d[k] = 0.5f * (lo[k] + hi[k]) - (float)S.mu[k];
            acc[k] += h[k];
            acc[3 + k] += h[k] * h[k];
            acc[6 + k] += fabsf(d[k]);
            const float d2 = d[k] * d[k];
            acc[9 + k] += d2 * d2;
            // extent class: floor(log2(h / ell)) + 4 in [0, 7], from the
            // exponent field; counted in 8-bit fields (a thread takes at most
            // kFeatGroup * tile / kWG <= 128 boxes)
            const float r = h[k] * inv_ell;
            int bin = r > 0.0f ? (int)((__float_as_uint(r) >> 23) & 0xFFu) - 127 + 4 : 0;
            bin = bin < 0 ? 0 : (bin > 7 ? 7 : bin);
            ext8[k] += 1ull << (8 * bin);
        }
        acc[36] += h[0] * h[1];
        acc[37] += h[0] * h[2];
        acc[38] += h[1] * h[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float z = d[k] * gs[k];  // in units of 4 standard deviations
            const float q = floorf(z * (float)kPlaneGrid) + (float)(kPlaneGrid / 2);
            g[k] = (int)fminf(fmaxf(q, 0.0f), (float)(kPlaneGrid - 1));
            const float q1 = floorf(z * (float)kAxisBins) + (float)(kAxisBins / 2);
            g1[k] = (int)fminf(fmaxf(q1, 0.0f), (float)(kAxisBins - 1));
        }
        atomicAdd(&pl[g[0] * kPlaneGrid + g[1]], 1u);
        atomicAdd(&pl[kPlaneBins + g[0] * kPlaneGrid + g[2]], 1u);
        atomicAdd(&pl[2 * kPlaneBins + g[1] * kPlaneGrid + g[2]], 1u);
#pragma unroll
        for (int k = 0; k < 3; ++k) atomicAdd(&pl[3 * kPlaneBins + k * kAxisBins + g1[k]], 1u);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[12 + 8 * k + b] = (float)((ext8[k] >> (8 * b)) & 0xFFu);
#pragma unroll
    for (int k = 0; k < kFeatSums; ++k) {
        const float v = wave_sum(acc[k]);
        if (l == 0) s_red[w][k] = v;
    }
    __syncthreads();
    if (tid < kFeatSums) {
        float v = 0.0f;
        for (int ww = 0; ww < kWaves; ++ww) v += s_red[ww][tid];
        tile_sums[(size_t)tile * kFeatSums + tid] = v;
    }
    for (uint32_t i = tid; i < (uint32_t)kDensWords; i += kWG) {
        uint32_t c = 0;
#pragma unroll
        for (uint32_t k = 0; k < kCopies; ++k) c += s_pl[k][i];
        if (c) atomicAdd(&a.plane_hist[(size_t)ti.seg * kDensWords + i], c);
    }
}

// One wave per scene: the tile sums in tile order (fp64), then the features.
// A kernel of its own so that no per-tile device-scope fence is needed.
__global__ __launch_bounds__(64) void k_features_fin(ObjPipeArgs a, const float* tile_sums,
                                                     float* feats) {
    __shared__ double s_fin[kFeatSums];
    const uint32_t seg = blockIdx.x, tid = threadIdx.x;
    const uint32_t n = a.seg.tile_seg ? a.seg.seg_n[seg] : a.n;
    const SceneStats S = scene_stats(a.seg_mom + 9 * (size_t)seg, n);
    const uint32_t t0 = a.seg.tile_seg ? a.seg.seg_tile0[seg] : 0u;
    const uint32_t t1 = a.seg.tile_seg ? a.seg.seg_tile0[seg + 1] : a.ntiles;
    if (tid < kFeatSums) {
        double v = 0.0;
        for (uint32_t t = t0; t < t1; t += kFeatGroup) v += (double)tile_sums[(size_t)t * kFeatSums + tid];
        s_fin[tid] = v;
    }
    // collision rates of the density histograms: sum of squared counts per
    // histogram (3 planes, then 3 axes), lane-parallel
    __shared__ double s_rate[6];
    {
        const uint32_t* dh = a.plane_hist + (size_t)seg * kDensWords;
        double q[6];
#pragma unroll
        for (int h = 0; h < 6; ++h) {
            const int base = h < 3 ? h * kPlaneBins : 3 * kPlaneBins + (h - 3) * kAxisBins;
            const int nb = h < 3 ? kPlaneBins : kAxisBins;
            double c = 0.0;
            for (int b = (int)tid; b < nb; b += 64) {
                const double x = (double)dh[base + b];
                c += x * x;
            }
            q[h] = c;
        }
#pragma unroll
        for (int h = 0; h < 6; ++h) {
            const double v = wave_sum(q[h]);
            if (tid == 0) s_rate[h] = v;
        }
    }
    __syncthreads();
    if (tid != 0) return;
    const double nn = n ? (double)n : 1.0, s2 = S.s2, e = S.ell;
    float f[kFeat];
    f[0] = (float)(log2(nn) / 16.0);
    f[1] = (float)(S.C[0][0] / s2); f[2] = (float)(S.C[0][1] / s2); f[3] = (float)(S.C[0][2] / s2);
    f[4] = (float)(S.C[1][1] / s2); f[5] = (float)(S.C[1][2] / s2); f[6] = (float)(S.C[2][2] / s2);
    double A[3][3], V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = S.C[i][j];
    jacobi3f(A, V);
    int ord[3] = {0, 1, 2};  // eigenvalues descending
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (A[ord[j]][ord[j]] > A[ord[i]][ord[i]]) { int t = ord[i]; ord[i] = ord[j]; ord[j] = t; }
    for (int r = 0; r < 3; ++r) {
        const int c = ord[r];
        f[7 + r] = (float)(A[c][c] / s2);
        int big = 0;
        for (int k = 1; k < 3; ++k)
            if (fabs(V[k][c]) > fabs(V[big][c])) big = k;
        const double sg = V[big][c] < 0 ? -1.0 : 1.0;
        for (int k = 0; k < 3; ++k) f[10 + 3 * r + k] = (float)(sg * V[k][c]);
    }
    for (int k = 0; k < 3; ++k) {
        const double hk = s_fin[k] / nn;
        f[19 + k] = (float)(hk / e);
        f[22 + k] = (float)(s_fin[3 + k] / nn / (e * e));
        f[25 + k] = (float)(s_fin[6 + k] / nn / sqrt(s2));
        const double ckk = S.C[k][k] > 1e-30 ? S.C[k][k] : 1e-30;
        for (int b = 0; b < 8; ++b) f[31 + 8 * k + b] = (float)(s_fin[12 + 8 * k + b] / nn);
        const double ratio = 2.0 * hk / sqrt(ckk);
        f[58 + k] = (float)(ratio < 4.0 ? ratio / 4.0 : 1.0);
    }
    // Density cost estimates.  A sweep tests each box against the boxes
    // whose projected centres fall near its own: about n^2 (extent) x (the
    // collision rate of the projected centres, sum of squared bin fractions
    // / bin size).  Axis k (plain sweep): mean extent on k x that rate on k;
    // plane normal to k (slab sweep): the product of the two mean extents x
    // that rate on the plane.  The features are each cost's log2 relative to
    // the mean of the three (only the ranking and ratios matter), halved and
    // clamped to [-1, 1].
    {
        double sd[3], hm[3], lc1[3], lc2[3];
        for (int k = 0; k < 3; ++k) {
            sd[k] = sqrt(S.C[k][k] > 1e-30 ? S.C[k][k] : 1e-30);
            hm[k] = s_fin[k] / nn + 1e-30;
        }
        // sum over bins of (count / n)^2
        auto rate = [&](int h) { return s_rate[h] / (nn * nn) + 1e-30; };
        const int pa[3][2] = {{1, 2}, {0, 2}, {0, 1}};  // the plane normal to axis k
        const int pbase[3] = {2, 1, 0};                 // its histogram (xy, xz, yz order)
        for (int k = 0; k < 3; ++k) {
            const double bin1 = 4.0 * sd[k] / kAxisBins;
            lc1[k] = log2(hm[k] * rate(3 + k) / bin1);
            const int a0 = pa[k][0], a1 = pa[k][1];
            const double area = (4.0 * sd[a0] / kPlaneGrid) * (4.0 * sd[a1] / kPlaneGrid);
            lc2[k] = log2(hm[a0] * hm[a1] * rate(pbase[k]) / area);
        }
        const double m1 = (lc1[0] + lc1[1] + lc1[2]) / 3.0, m2 = (lc2[0] + lc2[1] + lc2[2]) / 3.0;
        for (int k = 0; k < 3; ++k) {
            f[28 + k] = (float)fmin(fmax(0.5 * (lc1[k] - m1), -1.0), 1.0);
            f[61 + k] = (float)fmin(fmax(0.5 * (lc2[k] - m2), -1.0), 1.0);
        }
    }
    f[55] = (float)(s_fin[36] / nn / (e * e));
    f[56] = (float)(s_fin[37] / nn / (e * e));
    f[57] = (float)(s_fin[38] / nn / (e * e));
    float* out = feats + (size_t)kFeat * seg;
    for (int k = 0; k < kFeat; ++k) out[k] = isfinite(f[k]) ? f[k] : 0.0f;
}

// ------------------------------------------------------------------ MLP
// One wave per sample.  Layer 1: lane j holds hidden units j and j+64.
template <bool STORE>
__device__ __forceinline__ void mlp_row(const float* xrow, MlpWeights w, float* s_x, float* s_h,
                                        uint16_t* xb, uint16_t* h1, uint16_t* h2, float y[kOut]) {
    const uint32_t lane = lane_id();
    const uint16_t xv = to_bf16(xrow[lane]);
    if (STORE) xb[lane] = xv;
    s_x[lane] = from_bf16(xv);
    __builtin_amdgcn_wave_barrier();
    float a0 = w.b1[lane], a1 = w.b1[lane + 64];
    for (int i = 0; i < kFeat; ++i) {
        const float xi = s_x[i];
        a0 += xi * from_bf16(w.w1[i * kHid + lane]);
        a1 += xi * from_bf16(w.w1[i * kHid + lane + 64]);
    }
    const uint16_t q0 = to_bf16(a0 > 0.0f ? a0 : 0.0f), q1 = to_bf16(a1 > 0.0f ? a1 : 0.0f);
    if (STORE) { h1[lane] = q0; h1[lane + 64] = q1; }
    s_h[lane] = from_bf16(q0);
    s_h[lane + 64] = from_bf16(q1);
    __builtin_amdgcn_wave_barrier();
    a0 = w.b2[lane];
    a1 = w.b2[lane + 64];
    for (int i = 0; i < kHid; ++i) {
        const float hi = s_h[i];
        a0 += hi * from_bf16(w.w2[i * kHid + lane]);
        a1 += hi * from_bf16(w.w2[i * kHid + lane + 64]);
    }
    const uint16_t r0 = to_bf16(a0 > 0.0f ? a0 : 0.0f), r1 = to_bf16(a1 > 0.0f ? a1 : 0.0f);
    if (STORE) { h2[lane] = r0; h2[lane + 64] = r1; }
    const float g0 = from_bf16(r0), g1 = from_bf16(r1);
#pragma unroll
    for (int k = 0; k < kOut; ++k) {
        float p = g0 * from_bf16(w.w3[lane * kOut + k]) + g1 * from_bf16(w.w3[(lane + 64) * kOut + k]);
        y[k] = wave_sum(p) + w.b3[k];
    }
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(kWG) void k_mlp_fwd(const float* x, uint32_t B, MlpWeights w,
                                                 uint16_t* xb, uint16_t* h1, uint16_t* h2, float* y) {
    __shared__ float s_x[kWaves][kFeat];
    __shared__ float s_h[kWaves][kHid];
    const uint32_t wv = threadIdx.x >> 6, row = blockIdx.x * kWaves + wv;
    if (row >= B) return;
    float out[kOut];
    mlp_row<true>(x + (size_t)row * kFeat, w, s_x[wv], s_h[wv], xb + (size_t)row * kFeat,
                  h1 + (size_t)row * kHid, h2 + (size_t)row * kHid, out);
    if (lane_id() < kOut) {
        float v = out[0];
        for (int k = 1; k < kOut; ++k) v = lane_id() == (uint32_t)k ? out[k] : v;
        y[(size_t)row * kOut + lane_id()] = v;
    }
}

// Loss per sample (target row: 64 candidate regrets, then log(1 + pairs)):
//   L = sum_c p_c r_c + lambda (y3 - t64)^2,  p = softmax_c(kappa (u.a_c)^2),
// u = y[0:3] / |y[0:3]|, r_c = tests_c / min tests - 1.  The axis term is the
// expected regret of the candidates the prediction points at; near-ties
// between good axes cost nothing.  Mean over the batch.  Lane c handles
// candidate c; per-row backward to dy, dh2, dh1 (fp32).
__global__ __launch_bounds__(kWG) void k_mlp_bwd_rows(const float* y, const float* tgt, uint32_t B,
                                                      MlpWeights w, const uint16_t* h1,
                                                      const uint16_t* h2, float* dh1, float* dh2,
                                                      float* dy, float lambda, float kappa,
                                                      const float* cand, float* loss) {
    __shared__ float s_d[kWaves][kHid];
    const uint32_t wv = threadIdx.x >> 6, row = blockIdx.x * kWaves + wv, lane = lane_id();
    if (row >= B) return;
    const float* yr = y + (size_t)row * kOut;
    const float* t = tgt + (size_t)row * kTgt;
    const float u0 = yr[0], u1 = yr[1], u2 = yr[2];
    const float nu = sqrtf(u0 * u0 + u1 * u1 + u2 * u2) + 1e-6f;
    const float h0 = u0 / nu, h1v = u1 / nu, h2v = u2 / nu;
    const float a0 = cand[3 * lane], a1 = cand[3 * lane + 1], a2 = cand[3 * lane + 2];
    const float dc = h0 * a0 + h1v * a1 + h2v * a2;
    const float sc = kappa * dc * dc;
    float mx = sc;
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    const float e = expf(sc - mx);
    const float z = wave_sum(e);
    const float pc = e / z, rc = t[lane];
    const float la = wave_sum(pc * rc);  // expected regret
    const float gs = kappa * pc * (rc - la) * 2.0f * dc;  // dL/d(u_hat) contribution of c
    const float g0 = wave_sum(gs * a0), g1 = wave_sum(gs * a1), g2 = wave_sum(gs * a2);
    const float gr = g0 * h0 + g1 * h1v + g2 * h2v;
    const float invB = 1.0f / (float)B;
    float dyr[kOut];
    dyr[0] = (g0 - gr * h0) / nu * invB;
    dyr[1] = (g1 - gr * h1v) / nu * invB;
    dyr[2] = (g2 - gr * h2v) / nu * invB;
    const float ec = yr[3] - t[kCand];
    dyr[3] = 2.0f * lambda * ec * invB;
    if (lane < kOut) {
        float v = dyr[0];
        for (int k = 1; k < kOut; ++k) v = lane == (uint32_t)k ? dyr[k] : v;
        dy[(size_t)row * kOut + lane] = v;
    }
    if (lane == 0) loss[row] = la + lambda * ec * ec;
    // dh2 = W3 dy masked by relu
    for (int j = lane; j < kHid; j += 64) {
        float v = 0.0f;
        for (int k = 0; k < kOut; ++k) v += from_bf16(w.w3[j * kOut + k]) * dyr[k];
        v = from_bf16(h2[(size_t)row * kHid + j]) > 0.0f ? v : 0.0f;
        dh2[(size_t)row * kHid + j] = v;
        s_d[wv][j] = v;
    }
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < kHid; i += 64) {
        float v = 0.0f;
        for (int j = 0; j < kHid; ++j) v += from_bf16(w.w2[i * kHid + j]) * s_d[wv][j];
        v = from_bf16(h1[(size_t)row * kHid + i]) > 0.0f ? v : 0.0f;
        dh1[(size_t)row * kHid + i] = v;
    }
}

// Weight gradients: one thread per parameter, summed over the batch in row order.
__global__ void k_mlp_wgrad(uint32_t B, const uint16_t* xb, const uint16_t* h1, const uint16_t* h2,
                            const float* dh1, const float* dh2, const float* dy, float* g) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) return;
    float s = 0.0f;
    if (p < kOffB1) {
        const uint32_t i = (p - kOffW1) / kHid, j = (p - kOffW1) % kHid;
        for (uint32_t r = 0; r < B; ++r) s += from_bf16(xb[r * kFeat + i]) * dh1[r * kHid + j];
    } else if (p < kOffW2) {
        const uint32_t j = p - kOffB1;
        for (uint32_t r = 0; r < B; ++r) s += dh1[r * kHid + j];
    } else if (p < kOffB2) {
        const uint32_t i = (p - kOffW2) / kHid, j = (p - kOffW2) % kHid;
        for (uint32_t r = 0; r < B; ++r) s += from_bf16(h1[r * kHid + i]) * dh2[r * kHid + j];
    } else if (p < kOffW3) {
        const uint32_t j = p - kOffB2;
        for (uint32_t r = 0; r < B; ++r) s += dh2[r * kHid + j];
    } else if (p < kOffB3) {
        const uint32_t j = (p - kOffW3) / kOut, k = (p - kOffW3) % kOut;
        for (uint32_t r = 0; r < B; ++r) s += from_bf16(h2[r * kHid + j]) * dy[r * kOut + k];
    } else {
        const uint32_t k = p - kOffB3;
        for (uint32_t r = 0; r < B; ++r) s += dy[r * kOut + k];
    }
    g[p] = s;
}

__global__ void k_adam(float* master, const float* g, float* m, float* v, uint32_t np, float lr,
                       float b1, float b2, float eps, float c1, float c2, uint16_t* wb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float gg = g[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gg;
    const float vv = b2 * v[p] + (1.0f - b2) * gg * gg;
    m[p] = mm;
    v[p] = vv;
    const float upd = lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    const float x = master[p] - upd;
    master[p] = x;
    // bf16 copies of the weight matrices (biases stay fp32)
    if (p < kOffB1) wb[p - kOffW1] = to_bf16(x);
    else if (p >= kOffW2 && p < kOffB2) wb[kFeat * kHid + (p - kOffW2)] = to_bf16(x);
    else if (p >= kOffW3 && p < kOffB3) wb[kFeat * kHid + kHid * kHid + (p - kOffW3)] = to_bf16(x);
}

// Inference per scene; lane c scores candidate c.  A confident prediction
// replaces the scene's PCA axis by the nearest candidate.
__global__ __launch_bounds__(kWG) void k_axis_infer(const float* feats, uint32_t nseg, MlpWeights w,
                                                    const float* cand, float conf_min, float* axis,
                                                    float* normals) {
    __shared__ float s_x[kWaves][kFeat];
    __shared__ float s_h[kWaves][kHid];
    const uint32_t wv = threadIdx.x >> 6, seg = blockIdx.x * kWaves + wv, lane = lane_id();
    if (seg >= nseg) return;
    float y[kOut];
    mlp_row<false>(feats + (size_t)seg * kFeat, w, s_x[wv], s_h[wv], nullptr, nullptr, nullptr, y);
    const float nu = sqrtf(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    const float c0 = cand[3 * lane], c1 = cand[3 * lane + 1], c2 = cand[3 * lane + 2];
    float sc = nu > 0.0f ? fabsf(y[0] * c0 + y[1] * c1 + y[2] * c2) / nu : 0.0f;
    // slab mode: the sweep plane's normal is one of the coordinate axes (the
    // first three candidates); a tilted plane lengthens every box's projection
    if (normals && lane >= 3) sc = -1.0f;
    // wave argmax, lowest index on ties
    float best = sc;
    uint32_t bi = lane;
    for (int o = 32; o > 0; o >>= 1) {
        const float ob = __shfl_xor(best, o, 64);
        const uint32_t oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    // plain mode: the prediction must lie within acos(conf_min) of its
    // candidate; slab mode names one of three planes (the nearest coordinate
    // axis is always within 55 degrees), so only a degenerate output falls back
    const bool confident = nu > 0.1f && (normals != nullptr || best >= conf_min);
    if (lane == 0 && normals) {  // slab mode: the candidate is the sweep plane's normal
        float* nr = normals + 4 * (size_t)seg;
        nr[0] = cand[3 * bi];
        nr[1] = cand[3 * bi + 1];
        nr[2] = cand[3 * bi + 2];
        nr[3] = confident ? 1.0f : 0.0f;
    } else if (lane == 0 && confident) {
        float* ax = axis + 4 * (size_t)seg;
        ax[0] = cand[3 * bi];
        ax[1] = cand[3 * bi + 1];
        ax[2] = cand[3 * bi + 2];
        ax[3] = 3.0f;
    }
}

// ------------------------------------------------------------------ labels
// Exact sweep-test count of scene s on axis a (conservative keys identical to
// the pipeline): tests = sum_i #{j : key_j <= hikey_i} - n(n+1)/2.  Scenes of
// at most 2048 objects; keys sorted in LDS (bitonic, 64-bit key|index).  For
// axis 0 the scene's pair count is also produced by sweeping the sorted keys.
constexpr uint32_t kLabelMax = 2048;

__global__ __launch_bounds__(kWG) void k_axis_tests(const void* in, int kind, const uint32_t* seg_start,
                                                    const uint32_t* seg_n, const float* axes,
                                                    uint32_t naxes, unsigned long long* tests,
                                                    unsigned long long* pairs) {
    __shared__ unsigned long long s_k[kLabelMax];
    __shared__ unsigned long long s_part[kWaves];
    const uint32_t seg = blockIdx.x / naxes, ax = blockIdx.x % naxes, tid = threadIdx.x;
    const uint32_t base = seg_start[seg], n = seg_n[seg];
    if (n > kLabelMax) return;  // host rejects such scenes
    const float u[3] = {axes[3 * ax], axes[3 * ax + 1], axes[3 * ax + 2]};
    const bool exact = axis_exact(kind, u);
    uint32_t m = 1;
    while (m < n) m <<= 1;
    for (uint32_t i = tid; i < m; i += kWG) {
        unsigned long long k = ~0ull;
        if (i < n) {
            float lo[3], hi[3], c4[4], pl, ph;
            load_lohi(in, kind, base + i, lo, hi, c4);
            project_lohi(kind, lo, hi, c4, u, exact, &pl, &ph);
            k = ((unsigned long long)flip_f32(pl) << 32) | i;
        }
        s_k[i] = k;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < m; i += kWG) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const unsigned long long x = s_k[i], y = s_k[p];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { s_k[i] = y; s_k[p] = x; }
                }
            }
            __syncthreads();
        }
    unsigned long long cnt = 0, pc = 0;
    for (uint32_t s = tid; s < n; s += kWG) {
        const uint32_t i = (uint32_t)s_k[s];
        float lo[3], hi[3], c4[4], pl, ph;
        load_lohi(in, kind, base + i, lo, hi, c4);
        project_lohi(kind, lo, hi, c4, u, exact, &pl, &ph);
        const unsigned long long hk = ((unsigned long long)flip_f32(ph) << 32) | 0xFFFFFFFFull;
        uint32_t b = 0, e = n;
        while (b < e) {
            const uint32_t mid = (b + e) >> 1;
            if (s_k[mid] <= hk) b = mid + 1; else e = mid;
        }
        cnt += b;
        if (pairs && ax == 0) {
            for (uint32_t t = s + 1; t < b; ++t) {
                const uint32_t j = (uint32_t)s_k[t];
                float lo2[3], hi2[3], d4[4];
                load_lohi(in, kind, base + j, lo2, hi2, d4);
                bool hit;
                if (kind == 0) {
                    hit = lo[0] <= hi2[0] && lo2[0] <= hi[0] && lo[1] <= hi2[1] && lo2[1] <= hi[1] &&
                          lo[2] <= hi2[2] && lo2[2] <= hi[2];
                } else {
#pragma clang fp contract(off)
                    const float x0 = c4[0] - d4[0], y0 = c4[1] - d4[1], z0 = c4[2] - d4[2];
                    const float r0 = c4[3] + d4[3];
                    hit = ((x0 * x0 + y0 * y0) + z0 * z0) < r0 * r0;
                }
                pc += hit ? 1u : 0u;
            }
        }
    }
    cnt = wave_sum(cnt);
    pc = wave_sum(pc);
    if (lane_id() == 0) s_part[tid >> 6] = cnt;
    __syncthreads();
    if (tid == 0) {
        unsigned long long tot = 0;
        for (int i = 0; i < kWaves; ++i) tot += s_part[i];
        tests[(size_t)seg * naxes + ax] = tot - (unsigned long long)n * (n + 1) / 2;
    }
    __syncthreads();
    if (pairs && ax == 0) {
        if (lane_id() == 0) s_part[tid >> 6] = pc;
        __syncthreads();
        if (tid == 0) {
            unsigned long long tot = 0;
            for (int i = 0; i < kWaves; ++i) tot += s_part[i];
            pairs[seg] = tot;
        }
    }
}

// ------------------------------------------------------------------ launchers
hipError_t launch_features(const ObjPipeArgs& a, float* tile_sums, float* feats, hipStream_t st) {
    const uint32_t nseg = a.seg.tile_seg ? a.seg.nseg : 1u;
    static_assert(kDensWords % 2 == 0, "density histograms cleared as 64-bit words");
    hipError_t e = launch_zero_u64((unsigned long long*)a.plane_hist, nseg * (kDensWords / 2), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_features, dim3(a.ntiles), dim3(kWG), 0, st, a, tile_sums, feats);
    hipLaunchKernelGGL(k_features_fin, dim3(nseg), dim3(64), 0, st, a, tile_sums, feats);
    return hipGetLastError();
}
hipError_t launch_axis_infer(const float* feats, uint32_t nseg, MlpWeights w, const float* cand,
                             float conf_min, float* axis, float* normals, hipStream_t st) {
    if (nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_axis_infer, dim3((nseg + kWaves - 1) / kWaves), dim3(kWG), 0, st, feats,
                       nseg, w, cand, conf_min, axis, normals);
    return hipGetLastError();
}
hipError_t launch_axis_tests(const void* in, int kind, const uint32_t* seg_start, const uint32_t* seg_n,
                             uint32_t nseg, const float* axes, uint32_t naxes, unsigned long long* tests,
                             unsigned long long* pairs, hipStream_t st) {
    if (nseg == 0 || naxes == 0) return hipSuccess;
    hipLaunchKernelGGL(k_axis_tests, dim3(nseg * naxes), dim3(kWG), 0, st, in, kind, seg_start, seg_n,
                       axes, naxes, tests, pairs);
    return hipGetLastError();
}
hipError_t mlp_forward(const float* x, uint32_t B, MlpWeights w, uint16_t* xb, uint16_t* h1,
                       uint16_t* h2, float* y, hipStream_t st) {
    if (B == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((B + kWaves - 1) / kWaves), dim3(kWG), 0, st, x, B, w, xb, h1,
                       h2, y);
    return hipGetLastError();
}
hipError_t mlp_backward(const float* y, const float* target, uint32_t B, MlpWeights w,
                        const uint16_t* xb, const uint16_t* h1, const uint16_t* h2, float* dh1,
                        float* dh2, float* dy, float* grads, float lambda_count, float kappa,
                        const float* cand, float* loss, hipStream_t st) {
    if (B == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_bwd_rows, dim3((B + kWaves - 1) / kWaves), dim3(kWG), 0, st, y, target, B,
                       w, h1, h2, dh1, dh2, dy, lambda_count, kappa, cand, loss);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((kParams + 255) / 256), dim3(256), 0, st, B, xb, h1, h2, dh1,
                       dh2, dy, grads);
    return hipGetLastError();
}
hipError_t mlp_adam(float* master, const float* grads, float* m, float* v, uint32_t nparam, float lr,
                    float b1, float b2, float eps, uint32_t step, uint16_t* w_bf16,
                    const uint32_t* bf16_map, hipStream_t st) {
    (void)bf16_map;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((nparam + 255) / 256), dim3(256), 0, st, master, grads, m, v, nparam,
                       lr, b1, b2, eps, c1, c2, w_bf16);
    return hipGetLastError();
}

}  // namespace gsap
