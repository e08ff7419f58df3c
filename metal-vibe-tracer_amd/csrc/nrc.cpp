// nrc.cpp -- the learned radiance cache object behind vt_nrc_* (include/vt.h).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "nrc.h"
#include "scene.h"
#include "vt.h"

using vt::check;
using vt::DeviceBuffer;
using vt::fail;

struct vt_nrc {
    vt_context *ctx = nullptr;
    DeviceBuffer params, grads, m, v, ema, wbf16, wema, loss;
    uint32_t step = 0;
    vt::NrcDevice view() const {
        return vt::NrcDevice{params.as<float>(), grads.as<float>(), m.as<float>(), v.as<float>(), ema.as<float>(),
                             wbf16.as<uint16_t>(), wema.as<uint16_t>()};
    }
};

namespace vt {
NrcDevice nrc_device(vt_nrc *n) { return n->view(); }
uint32_t nrc_next_step(vt_nrc *n) { return ++n->step; }
}  // namespace vt

namespace {

uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// He-uniform initialisation: U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)); the first layer's fan-in is its
// 265 live features (264 encodings + the constant), padded columns and unused output rows are 0.
std::vector<float> initial_params(uint32_t seed) {
    std::vector<float> p(vt::kNrcParams, 0.0f);
    size_t o = 0;
    auto fill = [&](int rows, int cols, int live_cols, int live_rows, int fan_in) {
        const float bound = std::sqrt(6.0f / (float)fan_in);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c, ++o) {
                if (c >= live_cols || r >= live_rows) continue;
                uint32_t h = hash32(hash32((uint32_t)o * 2654435761u) ^ seed);
                float u = (float)(h >> 8) * (1.0f / 16777216.0f);
                p[o] = (2.0f * u - 1.0f) * bound;
            }
    };
    fill(vt::kNrcHidden, vt::kNrcFeatures, 265, vt::kNrcHidden, 265);
    for (int l = 0; l < 3; ++l) fill(vt::kNrcHidden, vt::kNrcHidden, vt::kNrcHidden, vt::kNrcHidden, vt::kNrcHidden);
    fill(vt::kNrcOutPad, vt::kNrcHidden, vt::kNrcHidden, vt::kNrcOut, vt::kNrcHidden);
    // The output layer starts at 1/20 of that scale, so the untrained cache answers near zero instead
    // of at radiances far above a typical scene's, which it would first have to unlearn.
    for (size_t i = vt::kNrcParams - vt::kNrcW5; i < vt::kNrcParams; ++i) p[i] *= 0.05f;
    return p;
}

void upload_params(vt_nrc *n, const float *p) {
    check(hipMemcpy(n->params.ptr, p, (size_t)vt::kNrcParams * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(vt::launch_nrc_sync(n->view(), nullptr), "nrc sync");
    check(vt::launch_nrc_zero(n->grads.as<float>(), vt::kNrcParams, nullptr), "nrc zero");
    check(vt::launch_nrc_zero(n->m.as<float>(), vt::kNrcParams, nullptr), "nrc zero");
    check(vt::launch_nrc_zero(n->v.as<float>(), vt::kNrcParams, nullptr), "nrc zero");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    n->step = 0;
}

}  // namespace

extern "C" {

vt_status vt_nrc_create(vt_context *ctx, uint32_t seed, vt_nrc **out) {
    return vt::guarded_call([&]() -> vt_status {
        if (!ctx || !out) return fail(VT_ERROR_INVALID_ARGUMENT, "context and out must not be NULL");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        auto n = std::make_unique<vt_nrc>();
        n->ctx = ctx;
        const size_t bytes = (size_t)vt::kNrcParams * 4;
        n->params.alloc(bytes);
        n->grads.alloc(bytes);
        n->m.alloc(bytes);
        n->v.alloc(bytes);
        n->ema.alloc((size_t)vt::kNrcParams * 4);
        n->wbf16.alloc((size_t)vt::kNrcParams * 2);
        n->wema.alloc((size_t)vt::kNrcParams * 2);
        n->loss.alloc(4);
        std::vector<float> p = initial_params(seed);
        upload_params(n.get(), p.data());
        *out = n.release();
        return VT_OK;
    });
}

void vt_nrc_destroy(vt_nrc *n) { delete n; }

vt_status vt_nrc_get_params(vt_nrc *n, float *params) {
    return vt::guarded_call([&]() -> vt_status {
        if (!n || !params) return fail(VT_ERROR_INVALID_ARGUMENT, "cache and params must not be NULL");
        check(hipSetDevice(n->ctx->device), "hipSetDevice");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check(hipMemcpy(params, n->params.ptr, (size_t)vt::kNrcParams * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        return VT_OK;
    });
}

vt_status vt_nrc_set_params(vt_nrc *n, const float *params) {
    return vt::guarded_call([&]() -> vt_status {
        if (!n || !params) return fail(VT_ERROR_INVALID_ARGUMENT, "cache and params must not be NULL");
        for (int i = 0; i < vt::kNrcParams; ++i)
            if (!std::isfinite(params[i])) return fail(VT_ERROR_INVALID_ARGUMENT, "parameters must be finite");
        check(hipSetDevice(n->ctx->device), "hipSetDevice");
        upload_params(n, params);
        return VT_OK;
    });
}

vt_status vt_nrc_infer(vt_nrc *n, const float *inputs, uint32_t count, float *outputs, void *stream) {
    return vt::guarded_call([&]() -> vt_status {
        if (!n || (count && (!inputs || !outputs))) return fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        check(hipSetDevice(n->ctx->device), "hipSetDevice");
        check(vt::launch_nrc_infer(n->view(), inputs, count, outputs, (hipStream_t)stream), "nrc infer");
        return VT_OK;
    });
}

vt_status vt_nrc_train(vt_nrc *n, const float *inputs, const float *targets, uint32_t count, float learning_rate,
                       void *stream) {
    return vt::guarded_call([&]() -> vt_status {
        if (!n || (count && (!inputs || !targets))) return fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        if (!(learning_rate > 0.0f) || !std::isfinite(learning_rate)) return fail(VT_ERROR_INVALID_ARGUMENT, "learning rate must be > 0");
        if (count == 0) return VT_OK;
        check(hipSetDevice(n->ctx->device), "hipSetDevice");
        hipStream_t s = (hipStream_t)stream;
        check(vt::launch_nrc_backward(n->view(), inputs, targets, count, (float)count * vt::kNrcOut, nullptr, s), "nrc backward");
        check(vt::launch_nrc_adam(n->view(), ++n->step, learning_rate, s), "nrc adam");
        return VT_OK;
    });
}

vt_status vt_nrc_gradients(vt_nrc *n, const float *inputs, const float *targets, uint32_t count, float *grads,
                           float *loss) {
    return vt::guarded_call([&]() -> vt_status {
        if (!n || (count && (!inputs || !targets)) || !grads) return fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        check(hipSetDevice(n->ctx->device), "hipSetDevice");
        check(vt::launch_nrc_zero(n->grads.as<float>(), vt::kNrcParams, nullptr), "nrc zero");
        check(vt::launch_nrc_zero(n->loss.as<float>(), 1, nullptr), "nrc zero");
        check(vt::launch_nrc_backward(n->view(), inputs, targets, count, (float)count * vt::kNrcOut, n->loss.as<float>(), nullptr),
              "nrc backward");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check(hipMemcpy(grads, n->grads.ptr, (size_t)vt::kNrcParams * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        float l = 0.0f;
        check(hipMemcpy(&l, n->loss.ptr, 4, hipMemcpyDeviceToHost), "hipMemcpy");
        if (loss) *loss = count ? l / ((float)count * vt::kNrcOut) : 0.0f;
        check(vt::launch_nrc_zero(n->grads.as<float>(), vt::kNrcParams, nullptr), "nrc zero");
        return VT_OK;
    });
}

}  // extern "C"
