// gsap_mlp_api.cpp -- host side of the learned sweep-axis selector
// (gsap_mlp_* and gsap_axis_* in include/gsap.h).  No reference counterpart.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/gsap.h"
#include "gsap_internal.h"

using namespace gsap;

namespace gsap_err_bridge {
void set(const std::string& m);
}

namespace {
constexpr uint32_t kParams = 25348;
constexpr uint32_t kW2 = kFeat * kHid, kW3 = kW2 + kHid * kHid, kWB = kW3 + kHid * kOut;
constexpr uint32_t kOffB1 = kFeat * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
                   kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid * kOut;

constexpr float kKappa = 30.0f;  // softmax sharpness of the expected-regret loss

int32_t mfail(int32_t code, const std::string& m) {
    gsap_err_bridge::set(m);
    return code;
}

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
}  // namespace

#define MCHK(expr)                                                                    \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return mfail(e_ == hipErrorOutOfMemory ? GSAP_ERR_OOM : GSAP_ERR_HIP,      \
                         std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

struct gsap_mlp {
    float* master = nullptr;  // fp32 parameters (layout above)
    float* m = nullptr;
    float* v = nullptr;
    float* grads = nullptr;
    uint16_t* wb = nullptr;   // bf16 weight matrices W1 | W2 | W3
    float* cand = nullptr;    // 64 candidate axes (device)
    uint32_t step = 0;
    // activations for a training batch
    uint32_t cap_b = 0;
    uint16_t *xb = nullptr, *h1 = nullptr, *h2 = nullptr;
    float *y = nullptr, *dy = nullptr, *dh1 = nullptr, *dh2 = nullptr, *loss = nullptr;
};

static MlpWeights weights_of(const gsap_mlp* h) {
    MlpWeights w;
    w.w1 = h->wb;
    w.w2 = h->wb + kW2;
    w.w3 = h->wb + kW3;
    w.b1 = h->master + kOffB1;
    w.b2 = h->master + kOffB2;
    w.b3 = h->master + kOffB3;
    return w;
}

static int32_t upload_params(gsap_mlp* h, const std::vector<float>& p) {
    std::vector<uint16_t> wb(kWB);
    for (uint32_t i = 0; i < kOffB1; ++i) wb[i] = host_bf16(p[i]);
    for (uint32_t i = 0; i < kHid * kHid; ++i) wb[kW2 + i] = host_bf16(p[kOffW2 + i]);
    for (uint32_t i = 0; i < kHid * kOut; ++i) wb[kW3 + i] = host_bf16(p[kOffW3 + i]);
    MCHK(hipMemcpy(h->master, p.data(), kParams * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(h->wb, wb.data(), kWB * 2, hipMemcpyHostToDevice));
    MCHK(hipMemset(h->m, 0, kParams * 4));
    MCHK(hipMemset(h->v, 0, kParams * 4));
    h->step = 0;
    return GSAP_OK;
}

extern "C" int32_t gsap_mlp_create(uint64_t seed, gsap_mlp** out) {
    if (!out) return mfail(GSAP_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    if (gsap_device_count() <= 0) return mfail(GSAP_ERR_HIP, "no HIP device");
    gsap_mlp* h = new gsap_mlp();
    hipError_t e = hipMalloc((void**)&h->master, kParams * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h->m, kParams * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h->v, kParams * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h->grads, kParams * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h->wb, kWB * 2);
    if (e == hipSuccess) e = hipMalloc((void**)&h->cand, kCand * 3 * 4);
    if (e == hipSuccess) {
        float c[kCand * 3];
        gsap_axis_candidates(c);
        e = hipMemcpy(h->cand, c, sizeof c, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        gsap_mlp_destroy(h);
        return mfail(GSAP_ERR_OOM, "gsap_mlp_create: allocation failed");
    }
    // He-uniform initialisation, zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> p(kParams, 0.0f);
    auto fill = [&](uint32_t off, uint32_t fan_in, uint32_t count) {
        const double b = std::sqrt(6.0 / fan_in);
        std::uniform_real_distribution<double> U(-b, b);
        for (uint32_t i = 0; i < count; ++i) p[off + i] = (float)U(rng);
    };
    fill(0, kFeat, kFeat * kHid);
    fill(kOffW2, kHid, kHid * kHid);
    fill(kOffW3, kHid, kHid * kOut);
    int32_t rc = upload_params(h, p);
    if (rc) {
        gsap_mlp_destroy(h);
        return rc;
    }
    *out = h;
    return GSAP_OK;
}

static void free_batch(gsap_mlp* h) {
    void* ps[] = {h->xb, h->h1, h->h2, h->y, h->dy, h->dh1, h->dh2, h->loss};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    h->xb = h->h1 = h->h2 = nullptr;
    h->y = h->dy = h->dh1 = h->dh2 = h->loss = nullptr;
    h->cap_b = 0;
}

extern "C" int32_t gsap_mlp_destroy(gsap_mlp* h) {
    if (!h) return GSAP_OK;
    (void)hipDeviceSynchronize();
    free_batch(h);
    void* ps[] = {h->master, h->m, h->v, h->grads, h->wb, h->cand};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete h;
    return GSAP_OK;
}

extern "C" int32_t gsap_mlp_num_params(void) { return (int32_t)kParams; }

extern "C" int32_t gsap_mlp_set_params(gsap_mlp* h, const float* p, uint32_t n) {
    if (!h || !p || n != kParams) return mfail(GSAP_ERR_INVALID_ARG, "gsap_mlp_set_params: bad size");
    return upload_params(h, std::vector<float>(p, p + n));
}

extern "C" int32_t gsap_mlp_get_params(gsap_mlp* h, float* p, uint32_t n) {
    if (!h || !p || n != kParams) return mfail(GSAP_ERR_INVALID_ARG, "gsap_mlp_get_params: bad size");
    MCHK(hipDeviceSynchronize());
    MCHK(hipMemcpy(p, h->master, kParams * 4, hipMemcpyDeviceToHost));
    return GSAP_OK;
}

static int32_t ensure_batch(gsap_mlp* h, uint32_t B) {
    if (B <= h->cap_b) return GSAP_OK;
    free_batch(h);
    MCHK(hipMalloc((void**)&h->xb, (size_t)B * kFeat * 2));
    MCHK(hipMalloc((void**)&h->h1, (size_t)B * kHid * 2));
    MCHK(hipMalloc((void**)&h->h2, (size_t)B * kHid * 2));
    MCHK(hipMalloc((void**)&h->y, (size_t)B * kOut * 4));
    MCHK(hipMalloc((void**)&h->dy, (size_t)B * kOut * 4));
    MCHK(hipMalloc((void**)&h->dh1, (size_t)B * kHid * 4));
    MCHK(hipMalloc((void**)&h->dh2, (size_t)B * kHid * 4));
    MCHK(hipMalloc((void**)&h->loss, (size_t)B * 4));
    h->cap_b = B;
    return GSAP_OK;
}

extern "C" int32_t gsap_mlp_forward(gsap_mlp* h, void* stream, const float* d_x, uint32_t n, float* d_y) {
    if (!h || (!d_x && n) || (!d_y && n)) return mfail(GSAP_ERR_INVALID_ARG, "null argument");
    if (int32_t rc = ensure_batch(h, n ? n : 1)) return rc;
    MCHK(mlp_forward(d_x, n, weights_of(h), h->xb, h->h1, h->h2, d_y, (hipStream_t)stream));
    return GSAP_OK;
}

// Loss and parameter gradient of one batch (no update), for gradient checks.
extern "C" int32_t gsap_mlp_loss_grad(gsap_mlp* h, void* stream, const float* d_x, const float* d_t,
                                      uint32_t n, float lambda_count, float* h_loss, float* h_grads) {
    if (!h || !d_x || !d_t || n == 0) return mfail(GSAP_ERR_INVALID_ARG, "bad argument");
    if (int32_t rc = ensure_batch(h, n)) return rc;
    hipStream_t st = (hipStream_t)stream;
    const MlpWeights w = weights_of(h);
    MCHK(mlp_forward(d_x, n, w, h->xb, h->h1, h->h2, h->y, st));
    MCHK(mlp_backward(h->y, d_t, n, w, h->xb, h->h1, h->h2, h->dh1, h->dh2, h->dy, h->grads,
                      lambda_count, kKappa, h->cand, h->loss, st));
    MCHK(hipStreamSynchronize(st));
    if (h_loss) {
        std::vector<float> l(n);
        MCHK(hipMemcpy(l.data(), h->loss, n * 4, hipMemcpyDeviceToHost));
        double s = 0;
        for (float x : l) s += x;
        *h_loss = (float)(s / n);
    }
    if (h_grads) MCHK(hipMemcpy(h_grads, h->grads, kParams * 4, hipMemcpyDeviceToHost));
    return GSAP_OK;
}

// Minibatch Adam over rows [0, n) in order, `epochs` times.
extern "C" int32_t gsap_mlp_train(gsap_mlp* h, void* stream, const float* d_x, const float* d_t,
                                  uint32_t n, uint32_t epochs, uint32_t batch, float lr,
                                  float lambda_count, float* h_epoch_loss) {
    if (!h || !d_x || !d_t || n == 0 || batch == 0) return mfail(GSAP_ERR_INVALID_ARG, "bad argument");
    if (int32_t rc = ensure_batch(h, batch)) return rc;
    hipStream_t st = (hipStream_t)stream;
    const MlpWeights w = weights_of(h);
    std::vector<float> l(batch);
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        double acc = 0;
        uint32_t cnt = 0;
        for (uint32_t b0 = 0; b0 + batch <= n; b0 += batch) {
            const float* x = d_x + (size_t)b0 * kFeat;
            const float* t = d_t + (size_t)b0 * kTgt;
            MCHK(mlp_forward(x, batch, w, h->xb, h->h1, h->h2, h->y, st));
            MCHK(mlp_backward(h->y, t, batch, w, h->xb, h->h1, h->h2, h->dh1, h->dh2, h->dy,
                              h->grads, lambda_count, kKappa, h->cand, h->loss, st));
            MCHK(mlp_adam(h->master, h->grads, h->m, h->v, kParams, lr, 0.9f, 0.999f, 1e-8f,
                          ++h->step, h->wb, nullptr, st));
            if (h_epoch_loss) {
                MCHK(hipMemcpyAsync(l.data(), h->loss, batch * 4, hipMemcpyDeviceToHost, st));
                MCHK(hipStreamSynchronize(st));
                for (float v : l) acc += v;
                cnt += batch;
            }
        }
        if (h_epoch_loss) h_epoch_loss[ep] = cnt ? (float)(acc / cnt) : 0.0f;
    }
    MCHK(hipStreamSynchronize(st));
    return GSAP_OK;
}

// Fixed candidate axes: the 3 coordinate axes, 6 face diagonals, 4 body
// diagonals and 51 Fibonacci points of the upper hemisphere; each with its
// largest-magnitude component positive.
extern "C" int32_t gsap_axis_candidates(float* h_axes) {
    if (!h_axes) return mfail(GSAP_ERR_INVALID_ARG, "null argument");
    std::vector<double> a;
    auto push = [&](double x, double y, double z) {
        const double nn = std::sqrt(x * x + y * y + z * z);
        x /= nn; y /= nn; z /= nn;
        const double ax[3] = {x, y, z};
        int big = 0;
        for (int k = 1; k < 3; ++k)
            if (std::fabs(ax[k]) > std::fabs(ax[big]) + 1e-12) big = k;
        const double sg = ax[big] < 0 ? -1.0 : 1.0;
        a.push_back(sg * x); a.push_back(sg * y); a.push_back(sg * z);
    };
    push(1, 0, 0); push(0, 1, 0); push(0, 0, 1);
    push(1, 1, 0); push(1, -1, 0); push(1, 0, 1); push(1, 0, -1); push(0, 1, 1); push(0, 1, -1);
    push(1, 1, 1); push(1, 1, -1); push(1, -1, 1); push(-1, 1, 1);
    const double ga = M_PI * (3.0 - std::sqrt(5.0));
    for (int i = 0; i < 51; ++i) {
        const double z = 1.0 - (i + 0.5) / 51.0, r = std::sqrt(1.0 - z * z), ph = ga * i;
        push(r * std::cos(ph), r * std::sin(ph), z);
    }
    for (int i = 0; i < kCand * 3; ++i) h_axes[i] = (float)a[i];
    // exact unit coordinate axes
    for (int k = 0; k < 9; ++k) h_axes[k] = (k % 4 == 0) ? 1.0f : 0.0f;
    return GSAP_OK;
}

namespace gsap_internal_api {
int32_t attach_model(gsap_sap* h, const uint16_t* d_wb, size_t wb_bytes, const float* d_master,
                     size_t master_bytes, float conf_min);
}

extern "C" int32_t gsap_sap_set_learned_model(gsap_sap* h, const gsap_mlp* m, float conf_min) {
    if (!h || !m) return mfail(GSAP_ERR_INVALID_ARG, "null argument");
    return gsap_internal_api::attach_model(h, m->wb, kWB * 2, m->master, kParams * 4, conf_min);
}
