/*
 * environment.c -- oracle (test infrastructure only): the environment map's sampling tables,
 * restated from MaterialLibrary.environmentImportance (Sources/StudioModel.swift:1300-1380) in the
 * reference's own arithmetic: float luminance, a separable 3 x 3 maximum (wrapping in longitude,
 * clamped in latitude), the solid-angle mean luminance accumulated in double, MIS compensation
 * (KARLIK2019: the compensated table samples max(0, L - mean), the original table L, both times
 * sin theta), float row sums, and per-row CDFs normalised by a float total with the last entry set to 1.
 *
 * rows: H x 2 floats, rows[2 y + table] (the reference's 2 x H texture: column = table);
 * cols: 2H x W floats, cols[(table H + y) W + x] (its W x 2H texture: the compensated table first).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static float max3f(float a, float b, float c) {
    /* Swift's max(a, b, c): the larger of a and b, then of that and c. */
    float m = b > a ? b : a;
    return c > m ? c : m;
}

static void cdf_rows(float *w, uint32_t count, uint32_t rows, size_t stride, size_t step) {
    for (uint32_t row = 0; row < rows; ++row) {
        const size_t start = (size_t)row * stride;
        float total = 0.0f;
        for (uint32_t i = 0; i < count; ++i) total += w[start + i * step];
        total = total > 1e-12f ? total : 1e-12f;
        float sum = 0.0f;
        for (uint32_t i = 0; i < count; ++i) {
            sum += w[start + i * step] / total;
            w[start + i * step] = i == count - 1 ? 1.0f : sum;
        }
    }
}

int vo_env_tables(const float *rgba, uint32_t width, uint32_t height, float *rows, float *cols) {
    if (!rgba || width == 0 || height == 0) return -1;
    const size_t n = (size_t)width * height;
    float *lum = malloc(n * sizeof(float)), *hor = malloc(n * sizeof(float));
    if (!lum || !hor) {
        free(lum);
        free(hor);
        return -1;
    }
    for (size_t i = 0; i < n; ++i) {
        const float *p = rgba + 4 * i;
        float l = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
        lum[i] = l > 1e-6f ? l : 1e-6f;
    }
    for (uint32_t y = 0; y < height; ++y) {
        const size_t base = (size_t)y * width;
        for (uint32_t x = 0; x < width; ++x)
            hor[base + x] = max3f(lum[base + x], lum[base + (x + width - 1) % width], lum[base + (x + 1) % width]);
    }
    const float pi = 3.14159265358979323846f;
    double mean_sum = 0.0, angle_sum = 0.0;
    for (uint32_t y = 0; y < height; ++y) {
        float st = sinf((pi * ((float)y + 0.5f)) / (float)height);
        const double sin_theta = (double)(st > 1e-4f ? st : 1e-4f);
        double row = 0.0;
        for (uint32_t x = 0; x < width; ++x) row += (double)lum[(size_t)y * width + x];
        mean_sum += row * sin_theta;
        angle_sum += sin_theta * (double)width;
    }
    const float mean = (float)(mean_sum / angle_sum);
    memset(rows, 0, 2 * (size_t)height * sizeof(float));
    float total = 0.0f;
    for (uint32_t y = 0; y < height; ++y) {
        const float theta = (pi * ((float)y + 0.5f)) / (float)height;
        float st = sinf(theta);
        st = st > 1e-4f ? st : 1e-4f;
        const size_t above = (size_t)(y > 0 ? y - 1 : 0) * width, here = (size_t)y * width,
                     below = (size_t)(y + 1 < height ? y + 1 : height - 1) * width;
        for (uint32_t x = 0; x < width; ++x) {
            const float dilated = max3f(hor[above + x], hor[here + x], hor[below + x]);
            const float d = dilated - mean;
            const float compensated = (d > 0.0f ? d : 0.0f) * st, original = dilated * st;
            rows[2 * y] += compensated;
            rows[2 * y + 1] += original;
            cols[here + x] = compensated;
            cols[n + here + x] = original;
            total += compensated;
        }
    }
    if (!(total > 0.0f)) {
        for (uint32_t y = 0; y < height; ++y) rows[2 * y] = rows[2 * y + 1];
        memcpy(cols, cols + n, n * sizeof(float));
    }
    cdf_rows(rows, height, 2, 1, 2);
    cdf_rows(cols, width, 2 * height, width, 1);
    free(lum);
    free(hor);
    return 0;
}
