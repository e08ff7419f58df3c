/*
 * energy.c -- oracle (test infrastructure only): the energy-compensation tables the layered BSDF
 * reads (openpbr_impl.h).
 *
 * Ideal metal (129 x 129 + 129), baked on the CPU: the recipe of
 * scripts/generate_metal_energy.py:12-33, 55-63 -- roughness i/128, cosine max(1e-5, (j/128)^2),
 * 16384 Hammersley samples of the GGX VNDF (openpbr_sample_aniso_ggx_smith_vndf), mean of
 * 1 - G1(l) for the mirrored direction l (0 below the horizon); the average complement integrates
 * the table, piecewise linear in sqrt(cosine), exactly.  Float arithmetic, as the reference's bake;
 * the tests pin the result to the reference's committed output (Shaders/MetalEnergy.metal).
 *
 * Dielectric tables (32 nodes per axis, the axes of
 * Vendor/OpenPBR/impl/openpbr_microfacet_multiple_scattering_data.h:27-58: relative IOR
 * 1/2.5..1 | 1..2.5, sqrt(alpha) j/31, cosine k/31): OpenPBR's constant data
 * (Vendor/OpenPBR/impl/data/openpbr_*dielectric*_data.h), which has no published generator.  The
 * caller hands them over once (vo_energy_set_dielectric; the Python wrapper loads the committed
 * fixture made from the vendored arrays); reading a table before that aborts.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct { float x, y, z; } f3;
static f3 F3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static f3 nrm(f3 a) { float l = sqrtf(dot3(a, a)); return F3(a.x / l, a.y / l, a.z / l); }

static uint32_t reverse_bits(uint32_t v) {
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
    v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
    return (v >> 16) | (v << 16);
}
/* openpbr_sample_aniso_ggx_smith_vndf, isotropic. */
static f3 vndf(float alpha, f3 wi, float r0, float r1) {
    f3 ih = nrm(F3(alpha * wi.x, alpha * wi.y, wi.z));
    float c = fminf(ih.z, 1.0f);
    float ang = 6.28318530717958647692f * r0;
    float opc = 1.0f + c, hz = (1.0f - r1) * opc, omz = r1 * opc, omc = 1.0f - c, opz = hz + omc;
    float tc = sqrtf(opz * omz);
    f3 m = F3(tc * cosf(ang) + ih.x, tc * sinf(ang) + ih.y, hz);
    return nrm(F3(alpha * m.x, alpha * m.y, m.z));
}
static float g1(f3 v, float alpha) {
    float z2 = v.z * v.z;
    if (z2 == 0.0f) return 0.0f;
    return 2.0f / (1.0f + sqrtf(1.0f + ((alpha * v.x) * (alpha * v.x) + (alpha * v.y) * (alpha * v.y)) / z2));
}

static float g_metal[129 * 129 + 129];
static uint16_t g_luts[VO_LUT_COUNT][32 * 32 * 32];
static int g_ready = 0;
static int g_dielectric_set = 0;

/* Relative IOR of node i of the 32-node IOR axis. */
float vo_energy_ior_node(int i) {
    if (i >= 16) return 1.0f + (float)(i - 16) / 15.0f * 1.5f;
    return 1.0f / (1.0f + (float)(15 - i) / 15.0f * 1.5f);
}
static void bake(void) {
    /* Ideal metal. */
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < 129; ++i) {
        for (int j = 0; j < 129; ++j) {
            if (i == 0) { g_metal[i * 129 + j] = 0.0f; continue; }
            float roughness = (float)i / 128.0f;
            float cosine = fmaxf(1e-5f, powf((float)j / 128.0f, 2.0f));
            f3 v = F3(sqrtf(fmaxf(0.0f, 1.0f - cosine * cosine)), 0.0f, cosine);
            float alpha = roughness * roughness, sum = 0.0f;
            for (uint32_t s = 0; s < 16384u; ++s) {
                float r0 = ((float)s + 0.5f) / 16384.0f, r1 = (float)reverse_bits(s) * 2.3283064365386963e-10f;
                f3 h = vndf(alpha, v, r0, r1);
                float d = dot3(v, h);
                f3 l = F3(-v.x + 2.0f * d * h.x, -v.y + 2.0f * d * h.y, -v.z + 2.0f * d * h.z);
                float e = l.z > 0.0f ? g1(l, alpha) : 0.0f;
                sum += 1.0f - e;
            }
            g_metal[i * 129 + j] = sum / 16384.0f;
        }
    }
    for (int r = 0; r < 129; ++r) {
        double integral = 0.0;
        for (int c = 0; c < 128; ++c) {
            double lo = c / 128.0, hi = (c + 1) / 128.0;
            double slope = ((double)g_metal[r * 129 + c + 1] - (double)g_metal[r * 129 + c]) * 128.0;
            double offset = (double)g_metal[r * 129 + c] - slope * lo;
            integral += offset * (pow(hi, 4) - pow(lo, 4)) + 0.8 * slope * (pow(hi, 5) - pow(lo, 5));
        }
        g_metal[129 * 129 + r] = (float)integral;
    }
}

static void ensure(void) {
#pragma omp critical(vo_energy_bake)
    {
        if (!g_ready) {
            bake();
            g_ready = 1;
        }
    }
}
void vo_energy_set_dielectric(const uint16_t *tables) {
    static const int sizes[VO_LUT_COUNT] = {32768, 1024, 1024, 32768, 1024};
    for (int id = 0; id < VO_LUT_COUNT; ++id) {
        memcpy(g_luts[id], tables, (size_t)sizes[id] * sizeof(uint16_t));
        tables += sizes[id];
    }
    g_dielectric_set = 1;
}
static float g_ltc[3072];
static int g_ltc_set;
void vo_ltc_set(const float *table) {
    memcpy(g_ltc, table, sizeof g_ltc);
    g_ltc_set = 1;
}
const float *vo_ltc_table(void) {
    if (!g_ltc_set) {
        fprintf(stderr, "oracle: fuzz LTC table not set (vo_ltc_set)\n");
        abort();
    }
    return g_ltc;
}
const uint16_t *vo_energy_table(int id) {
    if (!g_dielectric_set) {
        fprintf(stderr, "oracle: dielectric energy tables not set (vo_energy_set_dielectric)\n");
        abort();
    }
    if (!g_ready) ensure();
    return g_luts[id];
}
const float *vo_metal_energy_table(void) {
    if (!g_ready) ensure();
    return g_metal;
}
