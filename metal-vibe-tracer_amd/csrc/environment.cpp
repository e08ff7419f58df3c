// Environment map: host sampling tables and the scene's device copy (vt_scene_set_environment,
// vt_environment_tables).
//
// The tables follow MaterialLibrary.environmentImportance (Sources/StudioModel.swift:1300-1380) step
// for step, in its arithmetic, because the renderer's light samples must land in the same cells with
// the same densities as the reference's: float luminance (floored at 1e-6); a separable 3 x 3 maximum
// (the shader filters bilinearly while the density is constant per cell, so each bright texel lends its
// weight to its neighbours; wrapping in longitude, clamped in latitude); the mean luminance over solid
// angle, summed in double; per table a row of weights times sin(theta) summed in float -- table 0 the
// MIS-compensated max(0, dilated - mean) (KARLIK2019), table 1 the dilated luminance, table 0 falling
// back to table 1 when no texel exceeds the mean -- and CDFs normalised by their float total with the
// last entry set to exactly 1.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "scene.h"
#include "vt.h"

using vt::fail;

namespace vt {
namespace {

// max(a, b, c) as Swift evaluates it: max(max(a, b), c), where max(x, y) = y >= x ? y : x.
inline float swift_max(float a, float b) { return b >= a ? b : a; }
inline float swift_max3(float a, float b, float c) { return swift_max(swift_max(a, b), c); }

// In-place CDFs of `rows` rows of `count` weights (entry i of row r at r * stride + i * step).
void to_cdf(float *w, uint32_t count, uint32_t rows, size_t stride, size_t step) {
    for (uint32_t r = 0; r < rows; ++r) {
        float *row = w + (size_t)r * stride;
        float total = 0.0f;
        for (uint32_t i = 0; i < count; ++i) total += row[i * step];
        total = swift_max(1e-12f, total);
        float sum = 0.0f;
        for (uint32_t i = 0; i < count; ++i) {
            sum += row[i * step] / total;
            row[i * step] = i == count - 1 ? 1.0f : sum;
        }
    }
}

}  // namespace

void environment_tables(const float *rgba, uint32_t w, uint32_t h, float *rows, float *cols) {
    const size_t n = (size_t)w * h;
    std::vector<float> lum(n), wide(n);
    for (size_t i = 0; i < n; ++i) {
        const float *p = rgba + 4 * i;
        lum[i] = swift_max(1e-6f, 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2]);
    }
    for (uint32_t y = 0; y < h; ++y) {
        const float *l = lum.data() + (size_t)y * w;
        float *o = wide.data() + (size_t)y * w;
        for (uint32_t x = 0; x < w; ++x) o[x] = swift_max3(l[x], l[(x + w - 1) % w], l[(x + 1) % w]);
    }
    const float pi = 3.14159265358979323846f;
    double weighted = 0.0, solid = 0.0;
    for (uint32_t y = 0; y < h; ++y) {
        const double st = (double)swift_max(1e-4f, std::sin((pi * ((float)y + 0.5f)) / (float)h));
        double row = 0.0;
        for (uint32_t x = 0; x < w; ++x) row += (double)lum[(size_t)y * w + x];
        weighted += row * st;
        solid += st * (double)w;
    }
    const float mean = (float)(weighted / solid);
    std::fill(rows, rows + 2 * (size_t)h, 0.0f);
    float compensated_total = 0.0f;
    for (uint32_t y = 0; y < h; ++y) {
        const float st = swift_max(1e-4f, std::sin((pi * ((float)y + 0.5f)) / (float)h));
        const float *up = wide.data() + (size_t)(y == 0 ? 0 : y - 1) * w;
        const float *mid = wide.data() + (size_t)y * w;
        const float *down = wide.data() + (size_t)std::min(h - 1, y + 1) * w;
        float *c0 = cols + (size_t)y * w, *c1 = cols + n + (size_t)y * w;
        for (uint32_t x = 0; x < w; ++x) {
            const float dilated = swift_max3(up[x], mid[x], down[x]);
            const float comp = swift_max(0.0f, dilated - mean) * st, orig = dilated * st;
            rows[2 * y] += comp;
            rows[2 * y + 1] += orig;
            c0[x] = comp;
            c1[x] = orig;
            compensated_total += comp;
        }
    }
    if (!(compensated_total > 0.0f)) {
        for (uint32_t y = 0; y < h; ++y) rows[2 * y] = rows[2 * y + 1];
        std::memcpy(cols, cols + n, n * sizeof(float));
    }
    to_cdf(rows, h, 2, 1, 2);
    to_cdf(cols, w, 2 * h, w, 1);
}

}  // namespace vt

namespace {

vt_status check_size(uint32_t width, uint32_t height) {
    if (width == 0 || height == 0 || width > 16384 || height > 8192 || (uint64_t)width * height > 33554432ull)
        return fail(VT_ERROR_INVALID_ARGUMENT, "Could not decode the environment image (maximum 16384×8192).");
    return VT_OK;
}

}  // namespace

vt_status vt_environment_tables(const float *rgba, uint32_t width, uint32_t height, float *rows, float *cols) {
    return vt::guarded_call([&]() -> vt_status {
        if (!rgba || !rows || !cols) return fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        if (vt_status st = check_size(width, height)) return st;
        vt::environment_tables(rgba, width, height, rows, cols);
        return VT_OK;
    });
}

vt_status vt_scene_set_environment(vt_scene *scene, const float *rgba, uint32_t width, uint32_t height) {
    return vt::guarded_call([&]() -> vt_status {
        if (!scene) return fail(VT_ERROR_INVALID_ARGUMENT, "scene must not be NULL");
        vt::check(hipSetDevice(scene->ctx->device), "hipSetDevice");
        if (!rgba) {
            scene->env_image.reset();
            scene->env_rows.reset();
            scene->env_cols.reset();
            scene->env_w = scene->env_h = 0;
            scene->ctx->resident -= scene->env_accounted;
            scene->env_accounted = 0;
            scene->generation = vt_scene::next_generation();
            return VT_OK;
        }
        if (vt_status st = check_size(width, height)) return st;
        const size_t n = (size_t)width * height;
        std::vector<float> pixels(rgba, rgba + 4 * n);
        for (float &v : pixels) {
            if (!std::isfinite(v)) return fail(VT_ERROR_INVALID_ARGUMENT, "Environment pixels must be finite.");
            if (v < 0.0f) v = 0.0f;  // colour conversions can leave negative radiance
        }
        const uint64_t need = (uint64_t)n * 24 + (uint64_t)height * 8;
        if (scene->ctx->resident - scene->env_accounted + need > scene->ctx->budget)
            return fail(VT_ERROR_OUT_OF_MEMORY, "Environment image exceeds the safe GPU memory budget.");
        std::vector<float> rows(2 * (size_t)height), cols(2 * n);
        vt::environment_tables(pixels.data(), width, height, rows.data(), cols.data());
        scene->env_image.upload(pixels.data(), pixels.size() * sizeof(float));
        scene->env_rows.upload(rows.data(), rows.size() * sizeof(float));
        scene->env_cols.upload(cols.data(), cols.size() * sizeof(float));
        scene->env_w = width;
        scene->env_h = height;
        scene->ctx->resident += need - scene->env_accounted;
        scene->env_accounted = need;
        scene->generation = vt_scene::next_generation();
        return VT_OK;
    });
}
