/*
 * energy.c -- oracle (test infrastructure only): the energy-compensation tables the layered BSDF
 * reads (openpbr_impl.h), baked on the CPU.
 *
 * Ideal metal (129 x 129 + 129): the recipe of scripts/generate_metal_energy.py:12-33, 55-63 --
 * roughness i/128, cosine max(1e-5, (j/128)^2), 16384 Hammersley samples of the GGX VNDF
 * (openpbr_sample_aniso_ggx_smith_vndf), mean of 1 - G1(l) for the mirrored direction l (0 below
 * the horizon); the average complement integrates the table, piecewise linear in sqrt(cosine),
 * exactly.  Float arithmetic, as the reference's bake.
 *
 * Dielectric tables (32 nodes per axis, the axes of
 * Vendor/OpenPBR/impl/openpbr_microfacet_multiple_scattering_data.h:27-58: relative IOR
 * 1/2.5..1 | 1..2.5, roughness j/31, cosine k/31): the upstream data files carry no generator, so
 * they are defined here by the same single-scattering integrals -- 1024 Hammersley VNDF samples per
 * node, each scattered both ways with its Fresnel weight: E_R = <F G1(l_r)>, E_T = <(1 - F) G1(l_t)>.
 *   ideal dielectric energy complement   1 - (E_R + E_T)
 *   opaque dielectric energy complement  1 - E_R
 *   averages                             1 - 2 int E(mu) mu dmu (piecewise linear in mu)
 *   reflection ratio                     avg E_R / (avg E_R + avg E_T)
 * stored as unorm16 like the upstream tables.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
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
static float fresnel_f(float eta, float ci) {
    if (eta == 1.0f) return 0.0f;
    float st2 = (1.0f - ci * ci) / (eta * eta);
    if (st2 >= 1.0f) return 1.0f;
    float ct = sqrtf(1.0f - st2), eci = eta * ci, ect = eta * ct;
    float rs = (ci - ect) / (ci + ect), rp = (ct - eci) / (ct + eci);
    return 0.5f * (rs * rs + rp * rp);
}

static float g_metal[129 * 129 + 129];
static uint16_t g_luts[VO_LUT_COUNT][32 * 32 * 32];
static int g_ready = 0;

/* Relative IOR of node i of the 32-node IOR axis. */
float vo_energy_ior_node(int i) {
    if (i >= 16) return 1.0f + (float)(i - 16) / 15.0f * 1.5f;
    return 1.0f / (1.0f + (float)(15 - i) / 15.0f * 1.5f);
}
static uint16_t unorm16(double v) {
    if (!(v > 0)) return 0;
    if (v >= 1) return 65535;
    return (uint16_t)lrint(v * 65535.0);
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
    /* Dielectric single scattering per node. */
    static float er[32][32][32], et[32][32][32];
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int i = 0; i < 32; ++i) {
        for (int j = 0; j < 32; ++j) {
            float eta = vo_energy_ior_node(i), roughness = (float)j / 31.0f, alpha = roughness * roughness;
            for (int k = 0; k < 32; ++k) {
                float c = fmaxf(1e-5f, (float)k / 31.0f);
                f3 v = F3(sqrtf(fmaxf(0.0f, 1.0f - c * c)), 0.0f, c);
                float R = 0.0f, T = 0.0f;
                for (uint32_t s = 0; s < 1024u; ++s) {
                    float r0 = ((float)s + 0.5f) / 1024.0f, r1 = (float)reverse_bits(s) * 2.3283064365386963e-10f;
                    f3 h = vndf(alpha, v, r0, r1);
                    float d = dot3(v, h);
                    if (!(d > 0.0f)) continue;
                    float F = fresnel_f(eta, d);
                    f3 l = F3(-v.x + 2.0f * d * h.x, -v.y + 2.0f * d * h.y, -v.z + 2.0f * d * h.z);
                    if (l.z > 0.0f) R += F * g1(l, alpha);
                    float inv = 1.0f / eta, s2t = inv * inv * fmaxf(0.0f, 1.0f - d * d);
                    if (s2t < 1.0f) {
                        float ct = sqrtf(1.0f - s2t);
                        f3 t = nrm(F3(-v.x * inv + h.x * (inv * d - ct), -v.y * inv + h.y * (inv * d - ct),
                                      -v.z * inv + h.z * (inv * d - ct)));
                        if (t.z < 0.0f) T += (1.0f - F) * g1(t, alpha);
                    }
                }
                er[i][j][k] = R / 1024.0f;
                et[i][j][k] = T / 1024.0f;
            }
        }
    }
    for (int i = 0; i < 32; ++i) {
        for (int j = 0; j < 32; ++j) {
            double aR = 0, aT = 0;
            for (int k = 0; k < 31; ++k) {
                double lo = k / 31.0, hi = (k + 1) / 31.0;
                double sR = (er[i][j][k + 1] - er[i][j][k]) * 31.0, oR = er[i][j][k] - sR * lo;
                double sT = (et[i][j][k + 1] - et[i][j][k]) * 31.0, oT = et[i][j][k] - sT * lo;
                aR += 2.0 * (oR * (hi * hi - lo * lo) / 2.0 + sR * (hi * hi * hi - lo * lo * lo) / 3.0);
                aT += 2.0 * (oT * (hi * hi - lo * lo) / 2.0 + sT * (hi * hi * hi - lo * lo * lo) / 3.0);
            }
            for (int k = 0; k < 32; ++k) {
                g_luts[VO_LUT_DIELECTRIC][i * 1024 + j * 32 + k] = unorm16(1.0 - ((double)er[i][j][k] + et[i][j][k]));
                g_luts[VO_LUT_OPAQUE][i * 1024 + j * 32 + k] = unorm16(1.0 - (double)er[i][j][k]);
            }
            g_luts[VO_LUT_DIELECTRIC_AVG][i * 32 + j] = unorm16(1.0 - (aR + aT));
            g_luts[VO_LUT_OPAQUE_AVG][i * 32 + j] = unorm16(1.0 - aR);
            g_luts[VO_LUT_DIELECTRIC_RATIO][i * 32 + j] = unorm16(aR + aT > 0 ? aR / (aR + aT) : 1.0);
        }
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
const uint16_t *vo_energy_table(int id) {
    if (!g_ready) ensure();
    return g_luts[id];
}
const float *vo_metal_energy_table(void) {
    if (!g_ready) ensure();
    return g_metal;
}
