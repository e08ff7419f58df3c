/*
 * integrator.c -- oracle (test infrastructure only): the reference's pass-2 path loop in MIS mode,
 * restated per pixel for Lambertian surfaces and triangle emitters.  Compiled twice: REAL = float
 * (the reference's arithmetic: "the reference integrator") and REAL = double (the FP64 oracle the
 * relative MSE is measured against; traversal stays on the float geometry).
 *
 * Follows main.swift:
 *   pixel_sampler :508-515, wavelength_numbers :2218-2230, decorrelate_shading_seed :106-108,
 *   camera ray :7420-7436, frameJitter :8773-8786 (host),
 *   spectral_lambda / spectral_bin / spectrum_rgb :795-833, spectral_from_lagrange :898-903,
 *   spectral_emission :918-944,
 *   sample_direct_light (imported branch, power sampling) :2888-2917, imported_geometry :2720-2724,
 *   imported_emitter_area_pdf :2853-2864, eval_light_pdf :3070-3075, light_visible :3254-3264,
 *   ray tolerances :1343-1369 (scene-graph mode), mesh_hit_attributes :1587-1602,
 *   eval_bsdf / eval_bsdf_pdf (DIFFUSE) :2659-2684, sample_bsdf (DIFFUSE) :3352-3358,
 *   sample_cosine_hemisphere / make_basis :2448-2476, power_heuristic :2454-2460,
 *   emission_weight :3462-3466, scattering_limit :3471-3473,
 *   shading_kernel primary NEE :7655-7672, path loop :7712-7828, accumulation :7862-7876.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifndef REAL
#define VO_F32_UNIT 1
#define REAL float
#define SUFFIX(name) name##_f32
#define RSQRT sqrtf
#define RATAN atanf
#define RCOS cosf
#define RSIN sinf
#define RTAN tanf
#define RFABS fabsf
#define RFMAX fmaxf
#define RFMIN fminf
#define RFLOOR floorf
#define RACOS acosf
#define RATAN2 atan2f
#define RNEXTAFTER nextafterf
#endif

typedef REAL real;
typedef struct { real x, y, z; } v3;

static inline v3 V(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mulv(v3 a, real s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline real dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline real len(v3 a) { return RSQRT(dot(a, a)); }
static inline v3 norm(v3 a) { return mulv(a, (real)1 / RSQRT(dot(a, a))); }
static inline real maxabs(v3 p) { return RFMAX(RFABS(p.x), RFMAX(RFABS(p.y), RFABS(p.z))); }
static inline v3 ld(const float *p) { return V((real)p[0], (real)p[1], (real)p[2]); }

static const real PI_R = (real)3.14159265358979323846;

/* ------------------------------------------------------------------ spectral */
typedef struct { real v[4]; } s4;
static inline s4 S4(real a) { s4 r = {{a, a, a, a}}; return r; }
static inline s4 smul(s4 a, s4 b) { s4 r; for (int k = 0; k < 4; ++k) r.v[k] = a.v[k] * b.v[k]; return r; }
static inline s4 sscale(s4 a, real s) { s4 r; for (int k = 0; k < 4; ++k) r.v[k] = a.v[k] * s; return r; }
static inline real smax(s4 a) { return RFMAX(RFMAX(a.v[0], a.v[1]), RFMAX(a.v[2], a.v[3])); }

typedef struct { uint32_t bin[4]; real weight[4]; } lanes_t;

static lanes_t lanes_of(const vo_render_scene *sc, real u) {
    lanes_t l;
    for (int k = 0; k < 4; ++k) {
        real x = (u + (real)k) * (real)0.25 * (real)1024;
        uint32_t i = (uint32_t)x;
        if (i > 1023u) i = 1023u;
        real low = (real)sc->icdf[i], width = (real)sc->icdf[i + 1] - low;
        l.weight[k] = (real)0.25 * (real)1024 * width;
        real lambda = low + (x - (real)i) * width;
        real b = lambda - (real)360 + (real)0.5;
        if (b < 0) b = 0;
        if (b > 470) b = 470;
        uint32_t bin = (uint32_t)b;
        l.bin[k] = bin > 470u ? 470u : bin;
    }
    return l;
}
static v3 spectrum_rgb(const vo_render_scene *sc, const lanes_t *l, s4 s) {
    v3 rgb = V(0, 0, 0);
    for (int k = 0; k < 4; ++k) {
        const float *c = sc->cmf + 4 * l->bin[k];
        rgb = add(rgb, mulv(ld(c), s.v[k] * l->weight[k]));
    }
    return rgb;
}
static s4 reflectance(const vo_render_scene *sc, const lanes_t *l, const float *L) {
    if (L[3] == 2.0f) return S4((real)L[0]);
    s4 r;
    for (int k = 0; k < 4; ++k) {
        real x = (real)sc->cmf[4 * l->bin[k] + 3];
        real series = (real)L[0] + (real)2 * (real)L[1] * x + (real)2 * (real)L[2] * ((real)2 * x * x - (real)1);
        r.v[k] = RATAN(series) * (real)0.318309886183790671 + (real)0.5;
    }
    return r;
}
static s4 emission_spectrum(const vo_render_scene *sc, const lanes_t *l, v3 e) {
    e = V(RFMAX(e.x, 0), RFMAX(e.y, 0), RFMAX(e.z, 0));
    real hi = RFMAX(e.x, RFMAX(e.y, e.z));
    if (!(hi > 0)) return S4(0);
    real lo = RFMIN(e.x, RFMIN(e.y, e.z)), mid = e.x + e.y + e.z - hi - lo;
    uint32_t pi = e.x == hi ? 0u : (e.y == hi ? 1u : 2u);
    uint32_t si = pi != 2u && e.z == lo ? 2u : (pi != 1u && e.y == lo ? 1u : 0u);
    real ws = RFMAX(mid - lo, 0), wp = RFMAX(hi - mid, 0);
    s4 r;
    for (int k = 0; k < 4; ++k) {
        const float *a = sc->basis + 8 * l->bin[k], *b = a + 4;
        r.v[k] = (lo + ws * (real)b[si] + wp * (real)a[pi]) * (real)a[3];
    }
    return r;
}

/* ------------------------------------------------------------------ tolerances, sampling */
static inline real ray_epsilon(v3 p) { return RFMAX((real)1e-7, (real)9.536743e-7 * maxabs(p)); }
static inline real ray_t_min(v3 o) { return ray_epsilon(o) * (real)0.25; }
static inline v3 ray_origin(v3 p, v3 ng, v3 d, real error) {
    real off = RFMAX(error, ray_epsilon(p));
    return add(p, mulv(ng, dot(d, ng) >= 0 ? off : -off));
}
static inline real endpoint_tolerance(v3 o, v3 e, real d) {
    return RFMAX((real)2 * ray_epsilon(e), (real)1.907349e-6 * (maxabs(o) + maxabs(e) + d));
}
static inline real power_heuristic(real pf, real pg) {
    real s = RFMAX(pf, pg);
    if (s <= 0) return 0;
    real f = pf / s, g = pg / s;
    return f * f / (f * f + g * g);
}
static inline real emission_weight(int prev_delta, int prev_nee, int prev_mis, real bsdf_pdf, real light_pdf) {
    if (prev_delta || !prev_nee || light_pdf <= 0) return 1;
    return prev_mis ? power_heuristic(bsdf_pdf, light_pdf) : 0;
}
static v3 cosine_hemisphere(v3 n, real r0, real r1) {
    real phi = (real)6.28318530717958647692 * r0;
    real ct = RSQRT(r1 + (real)(0.5 / 16777216.0)), st = RSQRT(RFMAX((real)0, (real)1 - ct * ct));
    v3 up = RFABS(n.z) < (real)0.999 ? V(0, 0, 1) : V(1, 0, 0);
    v3 u = norm(cross(up, n)), v = cross(n, u);
    return norm(add(add(mulv(u, RCOS(phi) * st), mulv(v, RSIN(phi) * st)), mulv(n, ct)));
}

/* ------------------------------------------------------------------ scene access */
typedef struct { v3 a, b, c, na, nb, nc; real uv[3][2]; uint32_t tag; } tri_t;
static tri_t tri_of(const vo_render_scene *sc, uint32_t f) {
    const uint32_t *fc = sc->geo->faces + 4 * (size_t)f;
    const float *va = sc->geo->vertices + 8 * (size_t)fc[0], *vb = sc->geo->vertices + 8 * (size_t)fc[1],
                *vc = sc->geo->vertices + 8 * (size_t)fc[2];
    tri_t t;
    t.a = ld(va); t.b = ld(vb); t.c = ld(vc);
    t.na = ld(va + 3); t.nb = ld(vb + 3); t.nc = ld(vc + 3);
    const float *vs[3] = {va, vb, vc};
    for (int k = 0; k < 3; ++k) { t.uv[k][0] = (real)vs[k][6]; t.uv[k][1] = (real)vs[k][7]; }
    t.tag = fc[3];
    return t;
}
typedef struct {
    uint32_t type;
    const float *lagrange;
    v3 emission;
    real weight, roughness, ior, metalness, transmission, coat, anisotropy, fuzz;
} mat_t;
static mat_t mat_of(const vo_render_scene *sc, uint32_t slot) {
    static const float grey[4] = {0.7f, 0, 0, 2.0f};
    mat_t m;
    if (slot >= sc->material_count) {
        m.type = 0; m.lagrange = grey; m.emission = V(0, 0, 0); m.weight = 0;
        m.roughness = (real)0.5; m.ior = (real)1.5; m.metalness = 0; m.transmission = 0; m.coat = 0; m.anisotropy = 0;
        m.fuzz = 0;
        return m;
    }
    const float *r = sc->materials + VO_MATERIAL_FLOATS * (size_t)slot;
    memcpy(&m.type, r, 4);
    m.lagrange = r + 4;
    m.emission = ld(r + 8);
    m.weight = (real)r[11];
    m.roughness = (real)r[12];
    m.ior = (real)r[13];
    m.metalness = (real)r[14];
    m.transmission = (real)r[15];
    m.coat = (real)r[16];
    m.anisotropy = (real)r[17];
    m.fuzz = (real)r[18];
    if (r[8] > 0 || r[9] > 0 || r[10] > 0) m.type = 3;
    return m;
}
/* is_delta (main.swift:3285-3287). */
static int is_delta(const mat_t *m) { return m->type == 2u || (m->type == 1u && m->roughness < (real)0.02); }
static real imported_geometry(v3 p, v3 pos, const tri_t *t) {
    v3 d = sub(pos, p);
    real d2 = dot(d, d);
    v3 n = cross(sub(t->b, t->a), sub(t->c, t->a));
    return d2 > (real)1e-12 && dot(n, n) > (real)1e-20 ? RFMAX((real)0, dot(norm(n), mulv(neg(d), (real)1 / RSQRT(d2)))) / d2 : 0;
}
static float f_of(real x) { return (float)x; }

typedef struct { v3 p, n, ng, tan; real error, t; int front; uint32_t slot; } surf_t;

#include "openpbr_impl.h"

/* Closest hit through the oracle's reference traversal; returns 0 on a miss. */
static int trace(const vo_render_scene *sc, v3 o, v3 d, real tmin, surf_t *s, uint32_t *face_out) {
    float fo[3] = {f_of(o.x), f_of(o.y), f_of(o.z)}, fd[3] = {f_of(d.x), f_of(d.y), f_of(d.z)};
    vo_hit h = sc->geo->nodes ? vo_blas_closest(sc->geo, fo, fd, f_of(tmin), 1e20f, 0, 0)
                              : vo_brute_closest(sc->geo, fo, fd, f_of(tmin), 1e20f);
    if (h.face < 0) return 0;
    tri_t t = tri_of(sc, (uint32_t)h.face);
    real b1 = (real)h.b1, b2 = (real)h.b2, b0 = (real)1 - b1 - b2;
    s->p = add(add(mulv(t.a, b0), mulv(t.b, b1)), mulv(t.c, b2));
    v3 ng = norm(cross(sub(t.b, t.a), sub(t.c, t.a)));
    s->front = dot(ng, d) < 0;
    s->ng = s->front ? ng : neg(ng);
    v3 smooth = add(add(mulv(t.na, b0), mulv(t.nb, b1)), mulv(t.nc, b2));
    s->n = dot(smooth, smooth) > (real)1e-12 ? norm(smooth) : s->ng;
    if (dot(s->n, s->ng) < 0) s->n = neg(s->n);
    /* the tangent dP/du, made perpendicular to the shading normal (mesh_hit_attributes, main.swift:1596-1599) */
    v3 e1 = sub(t.b, t.a), e2 = sub(t.c, t.a);
    real d1x = t.uv[1][0] - t.uv[0][0], d1y = t.uv[1][1] - t.uv[0][1];
    real d2x = t.uv[2][0] - t.uv[0][0], d2y = t.uv[2][1] - t.uv[0][1];
    real det = d1x * d2y - d1y * d2x;
    v3 tg = RFABS(det) > (real)1e-8 ? mulv(sub(mulv(e1, d2y), mulv(e2, d1y)), (real)1 / det) : e1;
    tg = sub(tg, mulv(s->n, dot(tg, s->n)));
    if (dot(tg, tg) < (real)1e-12) tg = cross(s->n, RFABS(s->n.y) < (real)0.9 ? V(0, 1, 0) : V(1, 0, 0));
    s->tan = norm(tg);
    s->error = RFMAX((real)1e-7, (real)1.907349e-6 * RFMAX(maxabs(t.a), RFMAX(maxabs(t.b), maxabs(t.c))));
    s->slot = t.tag;
    s->t = (real)h.t;
    *face_out = (uint32_t)h.face;
    return 1;
}
static int occluded(const vo_render_scene *sc, v3 o, v3 d, real tmin, real tmax) {
    float fo[3] = {f_of(o.x), f_of(o.y), f_of(o.z)}, fd[3] = {f_of(d.x), f_of(d.y), f_of(d.z)};
    return sc->geo->nodes ? vo_blas_any(sc->geo, fo, fd, f_of(tmin), f_of(tmax), 0, 0)
                          : vo_brute_any(sc->geo, fo, fd, f_of(tmin), f_of(tmax));
}

/* sample_direct_light, imported branch (main.swift:2888-2917): draws 4 numbers; pdf is the
 * solid-angle pdf (0: no sample).  light = face + 2 (LightSample.isDirectional). */
/* The pixel's random numbers (Sampler, main.swift:213-293): the PCG stream, or in Z mode the Z++
 * sampler's draws (oracle/zsampler.c) under the pixel's key, each sampling event (a path vertex and a
 * stream) owning 32 dimensions. */
typedef struct { uint32_t pcg, key, dim; int z; uint32_t sub, sub_bits; } rng_t;
void vo_z_draw(uint32_t key, uint32_t *dimension, uint32_t sub, uint32_t sub_bits, int kind, float *out);
uint32_t vo_z_pixel_key(uint32_t x, uint32_t y, uint32_t frame_index, uint32_t sample_index, uint32_t model);
enum { Z_BSDF = 0, Z_NEE = 1, Z_ROULETTE = 2, Z_DI = 4, Z_GI_BSDF = 5, Z_GI_NEE = 6, Z_WAVELENGTH = 12 };
static real rr(rng_t *r) {
    if (!r->z) return (real)vo_rand_f(&r->pcg);
    float f;
    vo_z_draw(r->key, &r->dim, r->sub, r->sub_bits, 1, &f);
    return (real)f;
}
static void rr2(rng_t *r, real *a, real *b) {
    if (!r->z) {
        *a = (real)vo_rand_f(&r->pcg);
        *b = (real)vo_rand_f(&r->pcg);
        return;
    }
    float f[2];
    vo_z_draw(r->key, &r->dim, r->sub, r->sub_bits, 2, f);
    *a = (real)f[0];
    *b = (real)f[1];
}
static void rr3(rng_t *r, real *a, real *b, real *c) {
    if (!r->z) {
        *a = (real)vo_rand_f(&r->pcg);
        *b = (real)vo_rand_f(&r->pcg);
        *c = (real)vo_rand_f(&r->pcg);
        return;
    }
    float f[3];
    vo_z_draw(r->key, &r->dim, r->sub, r->sub_bits, 3, f);
    *a = (real)f[0];
    *b = (real)f[1];
    *c = (real)f[2];
}
/* sampler_event (main.swift:293-295). */
static void rng_event(rng_t *r, uint32_t vertex, uint32_t stream) {
    r->dim = (vertex * 16u + stream) * 32u;
    r->sub = r->sub_bits = 0u;
}
/* sampler_candidate (main.swift:296-299): candidate `index` of `bits` low frame-index bits. */
static void rng_candidate(rng_t *r, uint32_t vertex, uint32_t stream, uint32_t index, uint32_t bits) {
    rng_event(r, vertex, stream);
    r->sub = index;
    r->sub_bits = bits;
}

/* ------------------------------------------------------------------ environment map
 * An image environment (Uniforms.environment.z = 1) without a sun (sunParams.w = 0), in an imported
 * scene (sceneIndex 6): eval_environment :2123-2135 (bilinear, wrapping in longitude), the sampling
 * tables' cell density environment_cell_pdf :2102-2112, environment_image_pdf :2114-2121, the binary
 * search environment_cdf_index :2088-2096, the image branch of sample_direct_light :2934-2949,
 * environment_mixture_pdf :2707-2713 (sun probability 0), eval_environment_pdf :2880-2883 and
 * imported_light_probability :2716-2719. */
enum { ENV_COMPENSATED = 0, ENV_ORIGINAL = 1 };
static const float TWO_PI_F = 6.28318530717958647692f;
static int env_on(const vo_render_scene *sc) { return sc->env != NULL && sc->env_strength > 0.0f; }
/* The probability that a light sample takes the triangle emitters rather than the environment. */
static real imported_probability(const vo_render_scene *sc) {
    if (sc->emitter_count == 0) return 0;
    return env_on(sc) ? (real)0.5 : (real)1;
}
static const float *env_texel(const vo_render_scene *sc, int x, int y) {
    return sc->env + 4 * ((size_t)y * sc->env_w + (size_t)x);
}
/* eval_environment: the map at direction d, bilinear between texel centres (the sampler's linear
 * filter with repeat addressing; latitude clamped half a texel inside), times the strength. */
static v3 eval_environment(const vo_render_scene *sc, v3 d) {
    if (!env_on(sc)) return V(0, 0, 0);
    const real w = (real)sc->env_w, h = (real)sc->env_h;
    real dy = d.y < (real)-1 ? (real)-1 : (d.y > (real)1 ? (real)1 : d.y);
    real u = RATAN2(d.z, d.x) / (real)TWO_PI_F + (real)0.5 + (real)sc->env_rotation / (real)TWO_PI_F;
    u = u - RFLOOR(u);  /* repeat addressing */
    real v = RACOS(dy) / PI_R;
    const real vlo = (real)0.5 / h, vhi = (real)1 - (real)0.5 / h;
    v = v < vlo ? vlo : (v > vhi ? vhi : v);
    real fx = u * w - (real)0.5, fy = v * h - (real)0.5;
    fy = fy < 0 ? 0 : (fy > h - 1 ? h - 1 : fy);  /* inside the map whatever the rounding of 0.5 / h */
    real x0f = RFLOOR(fx), y0f = RFLOOR(fy);
    real ax = fx - x0f, ay = fy - y0f;
    long long x0 = (long long)x0f, y0 = (long long)y0f;
    int xa = (int)(((x0 % (long long)sc->env_w) + sc->env_w) % sc->env_w), xb = (xa + 1) % (int)sc->env_w;
    int ya = (int)(y0 < 0 ? 0 : (y0 >= (long long)sc->env_h ? sc->env_h - 1 : y0));
    int yb = ya + 1 < (int)sc->env_h ? ya + 1 : (int)sc->env_h - 1;
    const float *t00 = env_texel(sc, xa, ya), *t10 = env_texel(sc, xb, ya), *t01 = env_texel(sc, xa, yb),
                *t11 = env_texel(sc, xb, yb);
    real c[3];
    for (int k = 0; k < 3; ++k) {
        real top = (real)t00[k] * ((real)1 - ax) + (real)t10[k] * ax;
        real bot = (real)t01[k] * ((real)1 - ax) + (real)t11[k] * ax;
        c[k] = top * ((real)1 - ay) + bot * ay;
    }
    const real st = (real)sc->env_strength;
    return V(c[0] * st, c[1] * st, c[2] * st);
}
static real env_cell_pdf(const vo_render_scene *sc, uint32_t x, uint32_t y, real sin_theta, uint32_t table) {
    const uint32_t w = sc->env_w, h = sc->env_h, cond = y + table * h;
    real row = (real)sc->env_rows[2 * y + table];
    real prev_row = y == 0 ? 0 : (real)sc->env_rows[2 * (y - 1) + table];
    real col = (real)sc->env_cols[(size_t)cond * w + x];
    real prev_col = x == 0 ? 0 : (real)sc->env_cols[(size_t)cond * w + x - 1];
    real p = RFMAX((real)0, row - prev_row) * RFMAX((real)0, col - prev_col);
    if (vo_get_fault() == 11) return p * (real)((float)(w * h)) / ((real)2 * PI_R * PI_R);  /* planted */
    return p * (real)((float)(w * h)) / ((real)2 * PI_R * PI_R * RFMAX((real)1e-5, sin_theta));
}
static real env_image_pdf(const vo_render_scene *sc, v3 wi, uint32_t table) {
    const uint32_t w = sc->env_w, h = sc->env_h;
    real cy = wi.y < (real)-1 ? (real)-1 : (wi.y > (real)1 ? (real)1 : wi.y);
    real theta = RACOS(cy);
    real u = RATAN2(wi.z, wi.x) / (real)TWO_PI_F + (real)0.5 + (real)sc->env_rotation / (real)TWO_PI_F;
    real v = theta / PI_R;
    u = u - RFLOOR(u);
    if (u > (real)0x1.fffffep-1f) u = (real)0x1.fffffep-1f;
    v = v < 0 ? 0 : (v > (real)1 - (real)1e-7 ? (real)1 - (real)1e-7 : v);
    uint32_t x = (uint32_t)(u * (real)w), y = (uint32_t)(v * (real)h);
    if (x > w - 1) x = w - 1;
    if (y > h - 1) y = h - 1;
    return env_cell_pdf(sc, x, y, RSIN(theta), table);
}
/* eval_environment_pdf: the density of a light sample's environment proposal in `table`. */
static real env_pdf(const vo_render_scene *sc, v3 wi, uint32_t table) {
    if (!env_on(sc)) return 0;
    return ((real)1 - imported_probability(sc)) * env_image_pdf(sc, wi, table);
}
/* environment_cdf_index: the first entry >= value among count entries (at most 16 halvings). */
static uint32_t env_cdf_index(const float *cdf, uint32_t count, size_t base, size_t step, real value) {
    uint32_t low = 0, high = (count > 1 ? count : 1) - 1;
    for (uint32_t it = 0; it < 16 && low < high; ++it) {
        uint32_t mid = (low + high) / 2;
        if ((real)cdf[base + (size_t)mid * step] >= value) high = mid;
        else low = mid + 1;
    }
    return low;
}

typedef struct { v3 position, wi, emission; real pdf; uint32_t light; } light_sample;
/* The environment branch of sample_direct_light: 4 draws (two 2D constituents in Z mode). */
static light_sample sample_environment(const vo_render_scene *sc, v3 p, rng_t *seed, uint32_t table) {
    light_sample ls;
    memset(&ls, 0, sizeof ls);
    const uint32_t w = sc->env_w, h = sc->env_h;
    real cx, cy;
    rr2(seed, &cx, &cy);
    uint32_t y = env_cdf_index(sc->env_rows, h, table, 2, cx);
    uint32_t x = env_cdf_index(sc->env_cols, w, (size_t)(y + table * h) * w, 1, cy);
    real ox, oy;
    rr2(seed, &ox, &oy);
    /* The in-texel offset stays inside the chosen cell (uv below the cell's far edge). */
    real ux = (real)x / (real)w + ox / (real)w, uy = (real)y / (real)h + oy / (real)h;
    real ex = RNEXTAFTER(((real)x + 1) / (real)w, (real)0), ey = RNEXTAFTER(((real)y + 1) / (real)h, (real)0);
    ux = ux < ex ? ux : ex;
    uy = uy < ey ? uy : ey;
    real theta = uy * PI_R, phi = (ux - (real)0.5 - (real)sc->env_rotation / (real)TWO_PI_F) * (real)TWO_PI_F;
    real st = RSIN(theta);
    ls.wi = norm(V(st * RCOS(phi), RCOS(theta), st * RSIN(phi)));
    real image = env_cell_pdf(sc, x, y, st, table);
    ls.position = add(p, mulv(ls.wi, (real)1e6));
    ls.light = 1u;
    ls.pdf = ((real)1 - imported_probability(sc)) * image;
    ls.emission = eval_environment(sc, ls.wi);
    return ls;
}

static light_sample sample_direct_light(const vo_render_scene *sc, v3 p, rng_t *seed, uint32_t table) {
    light_sample ls;
    memset(&ls, 0, sizeof ls);
    const real pimp = imported_probability(sc);
    if (vo_get_fault() == 12) table = ENV_COMPENSATED;  /* planted */
    if (!(pimp > 0 && rr(seed) < pimp)) return env_on(sc) ? sample_environment(sc, p, seed, table) : ls;
    /* Emitter choice and the point on it: one 3D constituent in Z mode (main.swift:2886-2894). */
    real value, ry, rz;
    rr3(seed, &value, &ry, &rz);
    uint32_t face;
    float tree_pmf = 0.0f;
    if (sc->tree) {
        float fp[3] = {f_of(p.x), f_of(p.y), f_of(p.z)};
        uint32_t position = vo_light_tree_sample(sc->tree, fp, f_of(value), &tree_pmf);
        if (position == 0xffffffffu) return ls;
        face = sc->tree->leaves[3 * (size_t)position];
    } else {
        uint32_t low = 0, high = sc->emitter_count - 1;
        while (low < high) {
            uint32_t mid = (low + high) / 2;
            if ((real)sc->emitter_cdf[mid] >= value) high = mid;
            else low = mid + 1;
        }
        face = sc->emitter_face[low];
    }
    tri_t t = tri_of(sc, face);
    real root = RSQRT(ry);
    ls.position = add(add(mulv(t.a, (real)1 - root), mulv(t.b, root * ((real)1 - rz))), mulv(t.c, root * rz));
    v3 delta = sub(ls.position, p);
    real dist = len(delta);
    ls.wi = mulv(delta, (real)1 / RFMAX(dist, (real)1e-8));
    real geometry = imported_geometry(p, ls.position, &t);
    mat_t em = mat_of(sc, t.tag);
    ls.emission = em.emission;
    ls.light = face + 2u;
    real area_pdf;
    if (sc->tree) {
        real area = (real)0.5 * len(cross(sub(t.b, t.a), sub(t.c, t.a)));
        area_pdf = area > 0 ? (real)tree_pmf / area : 0;
    } else {
        area_pdf = sc->emitter_total > 0 ? em.weight / (real)sc->emitter_total : 0;
    }
    ls.pdf = geometry > (real)1e-12 ? pimp * area_pdf / geometry : 0;
    return ls;
}

/* light_visible (main.swift:3254-3264): shadow ray from the offset origin to the point, stopped
 * short by the endpoint tolerance; toward the environment (light 1) along wi for 1e6 less twice the
 * ray epsilon. */
static int light_visible_l(const vo_render_scene *sc, const surf_t *s, v3 wi, v3 pos, uint32_t light) {
    v3 o = ray_origin(s->p, s->ng, wi, s->error);
    if (light == 1u) {
        real tol = (real)2 * ray_epsilon(s->p);
        if ((real)1e6 <= tol) return 0;
        return !occluded(sc, o, norm(wi), ray_t_min(o), (real)1e6 - tol);
    }
    v3 dd = sub(pos, o);
    real d = len(dd), tol = endpoint_tolerance(o, pos, d);
    if (d <= tol) return 0;
    return !occluded(sc, o, norm(dd), ray_t_min(o), d - tol);
}
static int light_visible(const vo_render_scene *sc, const surf_t *s, v3 wi, v3 pos) {
    return light_visible_l(sc, s, wi, pos, 0u);
}

/* eval_bsdf_with_pdf (main.swift:2659-2691): f (not cosine-weighted) and the BSDF pdf of wi. */
static s4 eval_bsdf_pdf(const vo_render_scene *sc, const lanes_t *l, const surf_t *s, const mat_t *m, v3 wo, v3 wi,
                        real *pdf) {
    if (m->type == 0u) {
        *pdf = dot(s->n, wo) <= 0 ? 0 : RFMAX((real)0, dot(s->n, wi)) / PI_R;
        real cosine = RFABS(dot(s->n, wi));
        if (!(dot(wi, s->ng) > 0) || cosine < (real)1e-7 || dot(s->n, wo) <= 0 || !(dot(s->n, wi) > 0)) return S4(0);
        return sscale(reflectance(sc, l, m->lagrange), (real)1 / PI_R);
    }
    *pdf = 0;
    real cosine = RFABS(dot(s->n, wi));
    if (cosine < (real)1e-7 || dot(s->n, wo) <= 0) return S4(0);
    pbr_t p;
    pbr_inputs in = {(real)m->type, m->roughness, m->ior, m->metalness, m->transmission, m->coat, m->anisotropy, s->tan,
                     m->fuzz};
    pbr_prepare(&p, &in, s->n, wo, !s->front, reflectance(sc, l, m->lagrange));
    *pdf = pbr_pdf(&p, wi);
    if (!(m->transmission > 0) && dot(wi, s->ng) <= 0) return S4(0);
    return sscale(pbr_eval(&p, wi), (real)1 / cosine);
}

/* sample_bsdf (main.swift:3304-3368): direction, weight = f |cos| / pdf per lane, pdf (0 for the
 * delta branches).  Returns 0 when the path ends here. */
static int sample_bsdf(const vo_render_scene *sc, const lanes_t *l, const surf_t *s, const mat_t *m, v3 in,
                       rng_t *seed, v3 *dir, s4 *weight, real *pdf) {
    *pdf = 0;
    if (m->type == 2u) {
        real ior = m->ior;
        s4 tint = reflectance(sc, l, m->lagrange);
        real eta = s->front ? (real)1 / ior : ior;
        real r0 = ((real)1 - ior) / ((real)1 + ior);
        for (int attempt = 0; attempt < 2; ++attempt) {
            v3 mm = attempt == 0 ? s->n : s->ng;
            real ci = RFMIN(RFMAX(dot(neg(in), mm), (real)0), (real)1);
            real s2t = eta * eta * ((real)1 - ci * ci);
            real c = eta > 1 && s2t < 1 ? RSQRT((real)1 - s2t) : ci;
            real omc = (real)1 - c;
            real fr = r0 * r0 + ((real)1 - r0 * r0) * (omc * omc * omc * omc * omc);
            int reflected = s2t >= 1 || rr(seed) < fr;
            if (reflected) {
                *dir = norm(sub(in, mulv(mm, (real)2 * dot(mm, in))));
                *weight = tint;
            } else {
                *dir = norm(add(mulv(in, eta), mulv(mm, eta * ci - RSQRT((real)1 - s2t))));
                *weight = sscale(tint, eta * eta);
            }
            if ((dot(*dir, s->ng) > 0) == reflected) return 1;
        }
        return 0;
    }
    if (is_delta(m)) {
        v3 rn = sub(in, mulv(s->n, (real)2 * dot(s->n, in)));
        v3 mm = dot(rn, s->ng) > 0 ? s->n : s->ng;
        *dir = norm(sub(in, mulv(mm, (real)2 * dot(mm, in))));
        real ci = RFMIN(RFMAX(dot(neg(in), mm), (real)0), (real)1);
        s4 f0 = reflectance(sc, l, m->lagrange);
        real omc = (real)1 - ci, p5 = omc * omc * omc * omc * omc;
        for (int k = 0; k < 4; ++k) weight->v[k] = f0.v[k] + ((real)1 - f0.v[k]) * p5;
        return dot(*dir, s->ng) > 0;
    }
    if (m->type == 0u) {
        real r0, r1, r2;  /* the lobe-choice draw r0 is kept and unused (main.swift:3354-3356) */
        rr3(seed, &r0, &r1, &r2);
        *dir = cosine_hemisphere(s->n, r1, r2);
        *pdf = RFMAX((real)0, dot(s->n, *dir)) / PI_R;
        *weight = reflectance(sc, l, m->lagrange);
        return *pdf > 0 && dot(*dir, s->ng) > 0;
    }
    pbr_t p;
    pbr_inputs pin = {(real)m->type, m->roughness, m->ior, m->metalness, m->transmission, m->coat, m->anisotropy,
                      s->tan, m->fuzz};
    pbr_prepare(&p, &pin, s->n, neg(in), !s->front, reflectance(sc, l, m->lagrange));
    real r0, r1, r2;
    rr3(seed, &r0, &r1, &r2);
    if (!pbr_sample(&p, r0, r1, r2, dir, weight, pdf)) { *pdf = 0; return 0; }
    if (!(*pdf > 0)) return 0;
    if (!(m->transmission > 0) && dot(*dir, s->ng) <= 0) return 0;
    for (int k = 0; k < 4; ++k)
        if (!isfinite(weight->v[k]) || weight->v[k] < 0) return 0;
    return 1;
}

/* Next-event estimate at surface s: draws 4 numbers, returns the contribution if visible. */
static v3 nee(const vo_render_scene *sc, const lanes_t *l, const surf_t *s, const mat_t *m, v3 wo, s4 thr,
              int use_mis, uint32_t table, rng_t *seed) {
    v3 zero = V(0, 0, 0);
    light_sample ls = sample_direct_light(sc, s->p, seed, table);
    if (!(ls.pdf > 0)) return zero;
    if (!light_visible_l(sc, s, ls.wi, ls.position, ls.light)) return zero;
    real bpdf;
    s4 f = eval_bsdf_pdf(sc, l, s, m, wo, ls.wi, &bpdf);
    real cosine = RFABS(dot(s->n, ls.wi));
    real w = use_mis ? power_heuristic(ls.pdf, bpdf) : 1;
    s4 le = emission_spectrum(sc, l, ls.emission);
    return spectrum_rgb(sc, l, sscale(smul(smul(thr, f), le), cosine * w / ls.pdf));
}

/* The layered BSDF on its own, for tests: params = type, roughness, ior, metalness, transmission, coat,
 * anisotropy, tangent xyz, fuzz;
 * n is the shading normal on the arriving side, wo the direction toward the viewer, rho the base
 * colour per lane.  eval: f |cos| per lane and the aggregate pdf of wi. */
void SUFFIX(vo_bsdf_eval)(const float *params, const float *n, const float *wo, const float *wi, const float *rho,
                          int inside, float *fcos, float *pdf) {
    pbr_t p;
    pbr_inputs in = {params[0], params[1], params[2], params[3], params[4], params[5], params[6], ld(params + 7),
                     params[10]};
    s4 r = {{rho[0], rho[1], rho[2], rho[3]}};
    pbr_prepare(&p, &in, ld(n), ld(wo), inside, r);
    s4 f = pbr_eval(&p, ld(wi));
    for (int k = 0; k < 4; ++k) fcos[k] = (float)f.v[k];
    *pdf = (float)pbr_pdf(&p, ld(wi));
}
/* sample: returns 1 and (direction, f |cos| / pdf per lane, pdf), or 0. */
int SUFFIX(vo_bsdf_sample)(const float *params, const float *n, const float *wo, const float *rho, int inside,
                           const float *u3, float *dir, float *weight, float *pdf) {
    pbr_t p;
    pbr_inputs in = {params[0], params[1], params[2], params[3], params[4], params[5], params[6], ld(params + 7),
                     params[10]};
    s4 r = {{rho[0], rho[1], rho[2], rho[3]}};
    pbr_prepare(&p, &in, ld(n), ld(wo), inside, r);
    v3 d;
    s4 w;
    real pd;
    if (!pbr_sample(&p, u3[0], u3[1], u3[2], &d, &w, &pd)) return 0;
    dir[0] = (float)d.x; dir[1] = (float)d.y; dir[2] = (float)d.z;
    for (int k = 0; k < 4; ++k) weight[k] = (float)w.v[k];
    *pdf = (float)pd;
    return 1;
}

/* ------------------------------------------------------------------ camera and pixel streams */
typedef struct { v3 eye, fw, rt, up; real aspect, fov_scale; } cam_t;

/* Camera basis (host side of the reference, main.swift:11303-11346, and pass 2's frame). */
static cam_t camera_of(const vo_frame *fr, const float *eye) {
    cam_t c;
    c.eye = V(eye[0], eye[1], eye[2]);
    c.fw = norm(sub(V(fr->target[0], fr->target[1], fr->target[2]), c.eye));
    c.rt = norm(cross(c.fw, V(0, 1, 0)));
    c.up = cross(c.rt, c.fw);
    c.aspect = (real)fr->width / (real)fr->height;
    c.fov_scale = RTAN((real)(fr->fov * 0.5f) * PI_R / (real)180);
    return c;
}
/* pixel_sampler (main.swift:508-515). */
static uint32_t pixel_seed(const vo_frame *fr, uint32_t gx, uint32_t gy, uint32_t sample_index) {
    return (gy * fr->width + gx) ^ (sample_index * 1999999973u);
}
/* wavelength_numbers (main.swift:2218-2230): the pixel's wavelength number u, no draw. */
static real pixel_u(uint32_t seed) {
    uint32_t h = vo_pcg_hash(seed ^ 0x5bd1e995u);
    return (real)(h >> 8u) * (real)(1.0 / 16777216.0);
}
/* Camera ray through the jittered pixel (main.swift:7420-7436; frameJitter :8773-8786). */
static v3 camera_dir(const cam_t *c, const vo_frame *fr, uint32_t gx, uint32_t gy, uint32_t sample_index) {
    uint32_t jn = ((sample_index > 1u ? sample_index : 1u) - 1u) % 1024u + 1u;
    float jx = 0, jy = 0;
    {
        uint32_t m = jn;
        float frac = 1.0f;
        while (m > 0u) { frac /= 2.0f; jx += frac * (float)(m % 2u); m /= 2u; }
        m = jn;
        frac = 1.0f;
        while (m > 0u) { frac /= 3.0f; jy += frac * (float)(m % 3u); m /= 3u; }
        jx -= 0.5f;
        jy -= 0.5f;
    }
    real u = (((real)gx + (real)(jx + 0.5f)) / (real)fr->width) * (real)2 - (real)1;
    real v = (((real)gy + (real)(jy + 0.5f)) / (real)fr->height) * (real)2 - (real)1;
    v = -v;
    u *= c->aspect * c->fov_scale;
    v *= c->fov_scale;
    return norm(add(add(c->fw, mulv(c->rt, u)), mulv(c->up, v)));
}

/* The path after its primary vertex s (main.swift:7712-7828): BSDF continuation, emission with
 * emission_weight, NEE with MIS at every later vertex, roulette after bounce 3.  In ReSTIR mode
 * (samplingMode 0) the departure from a diffuse primary is not MIS-weighted: its DI reservoir
 * estimated the whole direct integral (main.swift:7725-7728). */
/* eval_light_pdf for an imported emitter face reached from prev (main.swift:3070-3075, with
 * imported_emitter_area_pdf :2853-2864): importedProbability x area density / geometry. */
static real hit_light_pdf(const vo_render_scene *sc, v3 prev, const surf_t *r, uint32_t face, const mat_t *rm) {
    tri_t t = tri_of(sc, face);
    real g = imported_geometry(prev, r->p, &t);
    if (sc->tree) {
        uint32_t lo = 0, hi = sc->tree->count;
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            if (sc->tree->ids[mid] < face) lo = mid + 1;
            else hi = mid;
        }
        if (lo < sc->tree->count && sc->tree->ids[lo] == face && g > (real)1e-12) {
            float fp[3] = {f_of(prev.x), f_of(prev.y), f_of(prev.z)};
            real area = (real)0.5 * len(cross(sub(t.b, t.a), sub(t.c, t.a)));
            real pmf = (real)vo_light_tree_pmf(sc->tree, fp, sc->tree->positions[lo]);
            return area > 0 ? imported_probability(sc) * (pmf / area) / g : 0;
        }
        return 0;
    }
    if (sc->emitter_count && sc->emitter_total > 0 && rm->weight > 0 && g > (real)1e-12)
        return imported_probability(sc) * (rm->weight / (real)sc->emitter_total) / g;
    return 0;
}
/* A continuation that leaves the scene (main.swift:7751-7758): the environment with the MIS weight
 * of its light-sample density in the compensated table. */
static v3 env_miss(const vo_render_scene *sc, const lanes_t *l, s4 thr, v3 d, int prev_delta, int prev_nee,
                   int prev_mis, real bpdf) {
    if (!env_on(sc)) return V(0, 0, 0);
    real w = emission_weight(prev_delta, prev_nee, prev_mis, bpdf, env_pdf(sc, d, ENV_COMPENSATED));
    return spectrum_rgb(sc, l, sscale(smul(thr, emission_spectrum(sc, l, eval_environment(sc, d))), w));
}

static v3 continue_path_gi(const vo_render_scene *sc, const vo_frame *fr, const lanes_t *l, surf_t s, mat_t m, v3 dir,
                           rng_t *seed, int gi_secondary);
static v3 continue_path(const vo_render_scene *sc, const vo_frame *fr, const lanes_t *l, surf_t s, mat_t m, v3 dir,
                        rng_t *seed) {
    return continue_path_gi(sc, fr, l, s, m, dir, seed, 0);
}
/* gi_secondary: ReSTIR GI holds this path's x1 -> x2 -> light suffix, so a diffuse secondary vertex
 * takes no NEE (restirGISecondary, main.swift:7801-7803). */
static v3 continue_path_gi(const vo_render_scene *sc, const vo_frame *fr, const lanes_t *l, surf_t s, mat_t m, v3 dir,
                           rng_t *seed, int gi_secondary) {
    const int limit = fr->max_depth > 1 ? fr->max_depth - 1 : 0;
    v3 radiance = V(0, 0, 0);
    s4 thr = S4(1);
    v3 cur_dir = dir;
    int sdepth = 0, emission_only = 0;
    uint32_t face;
    for (int bounce = 1;; ++bounce) {
        /* Departure from s: the scattering budget counts non-delta events only; BSDF-only paths
         * (mode 3) take one final continuation that may only reach emission. */
        const int prev_delta = is_delta(&m);
        if (!prev_delta && sdepth >= limit) {
            if (fr->sampling_mode != 3u || emission_only) break;
            emission_only = 1;
        }
        if (!prev_delta) ++sdepth;
        const int prev_nee = !prev_delta && fr->sampling_mode != 3u;
        int prev_mis = fr->sampling_mode == 1u || (fr->sampling_mode == 0u && (bounce > 1 || m.type != 0u));
        v3 nd;
        s4 bw;
        real bpdf;
        rng_event(seed, (uint32_t)bounce, Z_BSDF);  /* vertex `bounce` scatters (main.swift:7730) */
        if (!sample_bsdf(sc, l, &s, &m, cur_dir, seed, &nd, &bw, &bpdf)) break;
        thr = smul(thr, bw);
        if (!(isfinite(thr.v[0]) && isfinite(thr.v[1]) && isfinite(thr.v[2]) && isfinite(thr.v[3])) || smax(thr) <= 0) break;
        v3 prev = s.p;
        v3 o = ray_origin(s.p, s.ng, nd, s.error);
        cur_dir = nd;
        surf_t r;
        if (!trace(sc, o, nd, ray_t_min(o), &r, &face)) {
            radiance = add(radiance, env_miss(sc, l, thr, nd, prev_delta, prev_nee, prev_mis, bpdf));
            break;
        }
        mat_t rm = mat_of(sc, r.slot);
        if (rm.type == 3) {
            real w = emission_weight(prev_delta, prev_nee, prev_mis, bpdf, hit_light_pdf(sc, prev, &r, face, &rm));
            v3 e = r.front ? rm.emission : V(0, 0, 0);
            radiance = add(radiance, spectrum_rgb(sc, l, sscale(smul(thr, emission_spectrum(sc, l, e)), w)));
            break;
        }
        if (emission_only) break;
        const int gi_skip = gi_secondary && bounce == 1 && rm.type == 0u;
        if (!is_delta(&rm) && fr->sampling_mode != 3u && !gi_skip) {
            rng_event(seed, (uint32_t)bounce + 1u, Z_NEE);
            radiance = add(radiance, nee(sc, l, &r, &rm, neg(cur_dir), thr, fr->sampling_mode <= 1u && sdepth < limit,
                                         sdepth < limit ? ENV_COMPENSATED : ENV_ORIGINAL, seed));
        }
        s = r;
        m = rm;
        if (bounce > 3) {
            real survival = RFMIN(RFMAX(smax(thr), (real)0.05), is_delta(&m) ? (real)0.99 : (real)0.95);
            rng_event(seed, (uint32_t)bounce + 1u, Z_ROULETTE);
            if (rr(seed) >= survival) break;
            thr = sscale(thr, (real)1 / survival);
        }
    }
    return radiance;
}

/* One sample of one pixel: the reference's shading_kernel in MIS mode (samplingMode 1). */
static v3 sample_pixel(const vo_render_scene *sc, const vo_frame *fr, const cam_t *cam, uint32_t gx, uint32_t gy,
                       uint32_t sample_index) {
    const int limit = fr->max_depth > 1 ? fr->max_depth - 1 : 0;
    uint32_t pcg = pixel_seed(fr, gx, gy, sample_index);
    rng_t seed = {vo_pcg_hash(pcg ^ 0x27d4eb2fu), 0u, 0u, fr->sampler == 1u};  /* decorrelate_shading_seed (main.swift:106-108) */
    real u = pixel_u(pcg);
    if (seed.z) {
        /* Z mode: the pixel's key (this sample is frame s + 1 of the accumulation), and the wavelength
         * number from dimension 0 of the vertex-0 Z_WAVELENGTH event (main.swift:2221-2226). */
        seed.key = vo_z_pixel_key(gx, gy, sample_index - fr->sample_index0 + 1u, sample_index, fr->z_temporal);
        rng_event(&seed, 0u, Z_WAVELENGTH);
        u = rr(&seed);
    }
    lanes_t l = lanes_of(sc, u);
    v3 dir = camera_dir(cam, fr, gx, gy, sample_index);

    v3 radiance = V(0, 0, 0);
    surf_t s;
    uint32_t face;
    /* A camera ray that leaves the scene sees the environment's exact colour (environment_rgb). */
    if (!trace(sc, cam->eye, dir, ray_t_min(cam->eye), &s, &face)) return eval_environment(sc, dir);
    mat_t m = mat_of(sc, s.slot);
    if (m.type == 3) return s.front ? m.emission : V(0, 0, 0);
    if (!is_delta(&m) && fr->sampling_mode != 3u) {
        rng_event(&seed, 1u, Z_NEE);
        radiance = nee(sc, &l, &s, &m, neg(dir), S4(1), fr->sampling_mode <= 1u && limit > 0,
                       limit > 0 ? ENV_COMPENSATED : ENV_ORIGINAL, &seed);
    }
    radiance = add(radiance, continue_path(sc, fr, &l, s, m, dir, &seed));
    if (!(isfinite(radiance.x) && isfinite(radiance.y) && isfinite(radiance.z))) radiance = V(0, 0, 0);
    return radiance;
}

/* The environment on its own, for tests: for n directions (3 floats each), the radiance (rgb, strength
 * applied) and the light-sample density of `table` (importedProbability applied). */
void SUFFIX(vo_env_query)(const vo_render_scene *sc, const float *dirs, uint32_t n, uint32_t table, float *rgb,
                          float *pdf) {
    for (uint32_t i = 0; i < n; ++i) {
        v3 d = ld(dirs + 3 * (size_t)i);
        v3 c = eval_environment(sc, d);
        rgb[3 * (size_t)i] = (float)c.x;
        rgb[3 * (size_t)i + 1] = (float)c.y;
        rgb[3 * (size_t)i + 2] = (float)c.z;
        pdf[i] = (float)env_pdf(sc, d, table);
    }
}
/* n light samples at point p from the PCG stream `seed` (sample_direct_light with `table`): direction,
 * solid-angle pdf, emission rgb and light id (0 none, 1 environment, face + 2). */
void SUFFIX(vo_light_samples)(const vo_render_scene *sc, const float *p, uint32_t n, uint32_t seed, uint32_t table,
                              float *wi, float *pdf, float *rgb, uint32_t *light) {
    rng_t r = {seed, 0u, 0u, 0, 0u, 0u};
    const v3 q = ld(p);
    for (uint32_t i = 0; i < n; ++i) {
        light_sample ls = sample_direct_light(sc, q, &r, table);
        wi[3 * (size_t)i] = (float)ls.wi.x;
        wi[3 * (size_t)i + 1] = (float)ls.wi.y;
        wi[3 * (size_t)i + 2] = (float)ls.wi.z;
        pdf[i] = (float)ls.pdf;
        rgb[3 * (size_t)i] = (float)ls.emission.x;
        rgb[3 * (size_t)i + 1] = (float)ls.emission.y;
        rgb[3 * (size_t)i + 2] = (float)ls.emission.z;
        light[i] = ls.light;
    }
}

void SUFFIX(vo_render)(const vo_render_scene *sc, const vo_frame *frame, uint32_t sample_index0, uint32_t samples,
                       const uint32_t *pixels, uint32_t npixels, int threads, double *out_rgb) {
    vo_frame f = *frame;
    f.sample_index0 = sample_index0;  /* the accumulation starts at sample_index0 */
    const vo_frame *fr = &f;
    const cam_t cam = camera_of(fr, fr->eye);
    (void)vo_energy_table(0);  /* bake the BSDF tables outside the parallel region */
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int64_t k = 0; k < (int64_t)npixels; ++k) {
        uint32_t pix = pixels[k], gx = pix % fr->width, gy = pix / fr->width;
        double acc[3] = {0, 0, 0};
        for (uint32_t s = 0; s < samples; ++s) {
            v3 c = sample_pixel(sc, fr, &cam, gx, gy, sample_index0 + s);
            acc[0] += (double)c.x;
            acc[1] += (double)c.y;
            acc[2] += (double)c.z;
        }
        for (int q = 0; q < 3; ++q) out_rgb[3 * k + q] = acc[q] / (double)samples;
    }
}

/* ------------------------------------------------------------------ learned radiance cache
 * The library's cache-terminated estimator (no counterpart in the reference): a path whose pixel and
 * sample hash to a training path (1 in `period`) is the MIS path above; any other path ends at its
 * first vertex reached by a non-delta scatter that is not an emitter, adding throughput x max(y, 0)
 * at its lanes, y the network's answer for the vertex (oracle/nrc.py layout, evaluated with bfloat16
 * weights and layer inputs as the matrix cores consume them) divided by the network's unit: the
 * reciprocal of the mean of the basis table's D65 column over its 471 wavelengths. */
static float bf16r(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}
/* params: the network already rounded to bfloat16. */
static void nrc_forward(const float *params, const float *x, float *y) {
    float e[272], h[64], g[64];
    for (int i = 0; i < 11; ++i)
        for (int b = 0; b < 12; ++b) {
            double a = 3.14159265358979323846 * (double)(1 << b) * (double)x[i];
            e[24 * i + 2 * b] = bf16r((float)sin(a));
            e[24 * i + 2 * b + 1] = bf16r((float)cos(a));
        }
    e[264] = 1.0f;
    for (int k = 265; k < 272; ++k) e[k] = 0.0f;
    const float *w = params;
    for (int o = 0; o < 64; ++o) {
        float acc = 0.0f;
        for (int i = 0; i < 272; ++i) acc += w[o * 272 + i] * e[i];
        h[o] = bf16r(acc > 0.0f ? acc : 0.0f);
    }
    w += 64 * 272;
    for (int l = 0; l < 3; ++l, w += 64 * 64) {
        for (int o = 0; o < 64; ++o) {
            float acc = 0.0f;
            for (int i = 0; i < 64; ++i) acc += w[o * 64 + i] * h[i];
            g[o] = bf16r(acc > 0.0f ? acc : 0.0f);
        }
        memcpy(h, g, sizeof h);
    }
    for (int o = 0; o < 4; ++o) {
        float acc = 0.0f;
        for (int i = 0; i < 64; ++i) acc += w[o * 64 + i] * h[i];
        y[o] = acc;
    }
}
static int nrc_training(uint32_t pix, uint32_t sample_index, uint32_t period) {
    return vo_pcg_hash((pix * 0x9E3779B9u) ^ (sample_index * 0x85EBCA6Bu)) % period == 0u;
}
static v3 sample_pixel_nrc(const vo_render_scene *sc, const vo_frame *fr, const cam_t *cam, uint32_t gx, uint32_t gy,
                           uint32_t sample_index, const float *params, const float *box_lo, const float *box_scale,
                           float inv_unit) {
    const int limit = fr->max_depth > 1 ? fr->max_depth - 1 : 0;
    rng_t seed = {pixel_seed(fr, gx, gy, sample_index), 0u, 0u, fr->sampler == 1u, 0u, 0u};
    real u = pixel_u(seed.pcg);
    if (seed.z) {  /* Z mode: the events of sample_pixel / continue_path */
        seed.key = vo_z_pixel_key(gx, gy, sample_index - fr->sample_index0 + 1u, sample_index, fr->z_temporal);
        rng_event(&seed, 0u, Z_WAVELENGTH);
        u = rr(&seed);
    }
    lanes_t l = lanes_of(sc, u);
    v3 dir = camera_dir(cam, fr, gx, gy, sample_index);
    seed.pcg = vo_pcg_hash(seed.pcg ^ 0x27d4eb2fu);
    v3 radiance = V(0, 0, 0);
    surf_t s;
    uint32_t face;
    if (!trace(sc, cam->eye, dir, ray_t_min(cam->eye), &s, &face)) return eval_environment(sc, dir);
    mat_t m = mat_of(sc, s.slot);
    if (m.type == 3) return s.front ? m.emission : V(0, 0, 0);
    if (!is_delta(&m)) {
        rng_event(&seed, 1u, Z_NEE);
        radiance = nee(sc, &l, &s, &m, neg(dir), S4(1), limit > 0, limit > 0 ? ENV_COMPENSATED : ENV_ORIGINAL, &seed);
    }
    s4 thr = S4(1);
    int sdepth = 0;
    for (int bounce = 1;; ++bounce) {
        const int prev_delta = is_delta(&m);
        if (!prev_delta && sdepth >= limit) break;
        if (!prev_delta) ++sdepth;
        v3 nd;
        s4 bw;
        real bpdf;
        rng_event(&seed, (uint32_t)bounce, Z_BSDF);
        if (!sample_bsdf(sc, &l, &s, &m, dir, &seed, &nd, &bw, &bpdf)) break;
        thr = smul(thr, bw);
        if (!(isfinite(thr.v[0]) && isfinite(thr.v[1]) && isfinite(thr.v[2]) && isfinite(thr.v[3])) || smax(thr) <= 0) break;
        v3 prev = s.p;
        v3 o = ray_origin(s.p, s.ng, nd, s.error);
        dir = nd;
        surf_t r;
        if (!trace(sc, o, nd, ray_t_min(o), &r, &face)) {
            radiance = add(radiance, env_miss(sc, &l, thr, nd, prev_delta, !prev_delta, 1, bpdf));
            break;
        }
        mat_t rm = mat_of(sc, r.slot);
        if (rm.type == 3) {
            real w = emission_weight(prev_delta, !prev_delta, 1, bpdf, hit_light_pdf(sc, prev, &r, face, &rm));
            v3 e = r.front ? rm.emission : V(0, 0, 0);
            radiance = add(radiance, spectrum_rgb(sc, &l, sscale(smul(thr, emission_spectrum(sc, &l, e)), w)));
            break;
        }
        if (!prev_delta) {
            float x[11], y[4];
            x[0] = fminf(fmaxf(((float)r.p.x - box_lo[0]) * box_scale[0], 0.0f), 1.0f);
            x[1] = fminf(fmaxf(((float)r.p.y - box_lo[1]) * box_scale[1], 0.0f), 1.0f);
            x[2] = fminf(fmaxf(((float)r.p.z - box_lo[2]) * box_scale[2], 0.0f), 1.0f);
            x[3] = (1.0f - (float)dir.x) * 0.5f;
            x[4] = (1.0f - (float)dir.y) * 0.5f;
            x[5] = (1.0f - (float)dir.z) * 0.5f;
            x[6] = ((float)r.n.x + 1.0f) * 0.5f;
            x[7] = ((float)r.n.y + 1.0f) * 0.5f;
            x[8] = ((float)r.n.z + 1.0f) * 0.5f;
            x[9] = rm.type == 0u ? 1.0f : fminf(fmaxf((float)rm.roughness, 0.0f), 1.0f);
            x[10] = (float)u;
            nrc_forward(params, x, y);
            s4 v;
            for (int k = 0; k < 4; ++k) v.v[k] = thr.v[k] * (real)(y[k] > 0.0f ? y[k] : 0.0f) * (real)inv_unit;
            radiance = add(radiance, spectrum_rgb(sc, &l, v));
            break;
        }
        if (!is_delta(&rm)) {
            rng_event(&seed, (uint32_t)bounce + 1u, Z_NEE);
            radiance = add(radiance, nee(sc, &l, &r, &rm, neg(dir), thr, sdepth < limit,
                                         sdepth < limit ? ENV_COMPENSATED : ENV_ORIGINAL, &seed));
        }
        s = r;
        m = rm;
        if (bounce > 3) {
            real survival = RFMIN(RFMAX(smax(thr), (real)0.05), is_delta(&m) ? (real)0.99 : (real)0.95);
            rng_event(&seed, (uint32_t)bounce + 1u, Z_ROULETTE);
            if (rr(&seed) >= survival) break;
            thr = sscale(thr, (real)1 / survival);
        }
    }
    if (!(isfinite(radiance.x) && isfinite(radiance.y) && isfinite(radiance.z))) radiance = V(0, 0, 0);
    return radiance;
}

void SUFFIX(vo_render_nrc)(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                           uint32_t period, const float *params, const float *box_lo, const float *box_scale, int threads,
                           double *out_rgb) {
    vo_frame frame = *fr;
    frame.sample_index0 = sample_index0;  /* the accumulation starts at sample_index0 */
    fr = &frame;
    const cam_t cam = camera_of(fr, fr->eye);
    (void)vo_energy_table(0);
    const int64_t npix = (int64_t)fr->width * fr->height;
    float *wb = malloc(30720 * sizeof(float));
    if (!wb) return;
    for (int i = 0; i < 30720; ++i) wb[i] = bf16r(params[i]);
    params = wb;
    double mean = 0.0;
    for (int i = 0; i < 471; ++i) mean += sc->basis[8 * i + 3];
    const float inv_unit = 1.0f / (float)(1.0 / (mean / 471.0));
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int64_t k = 0; k < npix; ++k) {
        uint32_t gx = (uint32_t)(k % fr->width), gy = (uint32_t)(k / fr->width);
        double acc[3] = {0, 0, 0};
        for (uint32_t smp = 0; smp < samples; ++smp) {
            const uint32_t si = sample_index0 + smp;
            v3 c = nrc_training((uint32_t)k, si, period) ? sample_pixel(sc, fr, &cam, gx, gy, si)
                                                          : sample_pixel_nrc(sc, fr, &cam, gx, gy, si, params, box_lo, box_scale, inv_unit);
            acc[0] += (double)c.x;
            acc[1] += (double)c.y;
            acc[2] += (double)c.z;
        }
        for (int q = 0; q < 3; ++q) out_rgb[3 * k + q] = acc[q] / (double)samples;
    }
    free(wb);
}

/* ------------------------------------------------------------------ ReSTIR DI (samplingMode 0)
 * Pass 1 (restir_temporal_kernel, main.swift:4842-4940): G-buffer, RIS over `candidates` light
 * samples (restir_di_initial :4637-4668), temporal merge through backprojection
 * (restir_backproject :4828-4835, restir_same_surface :4816-4823, restir_di_merge :4692-4704),
 * encode (restir_di_encode :4708-4716).  Pass 2 (shading_kernel :7484-7580): four uniform spatial
 * taps in a 16-pixel box (uniform_neighbor :5025-5027, restir2020_compatible :5030-5033), final W,
 * shading with one visibility ray; then the path suffix.  The stored texels hold what the
 * reference's textures hold: rgba32f G-buffer position, rgba16f normal/material type, rgba32f
 * reservoirs. */
typedef struct {
    float gpos[4];      /* position, camera distance (w = -1: no hit) */
    float gnorm[4];     /* normal, material type, each rounded to half precision */
    float pos[3];       /* sample position */
    uint32_t light;     /* face + 2; 0 = no sample */
    float emit_pdf[4];  /* emission rgb, solid-angle pdf */
    float weights[4];   /* weight sum, M, W, wavelength number */
} di_texel;
typedef struct { surf_t s; v3 dir; int hit; } primary_t;
typedef struct { v3 position, emission; uint32_t light; real pdf, u; } di_sample;

/* Round to the nearest half-precision value (ties to even), as an rgba16f store does. */
float vo_half_round(float x);
static float half_round(float x) { return vo_half_round(x); }
#ifdef VO_F32_UNIT
float vo_half_round(float x) {
    if (x == 0.0f || !isfinite(x)) return x;
    float ax = fabsf(x);
    if (ax >= 65520.0f) return copysignf(INFINITY, x);
    int e;
    frexpf(ax, &e);
    int lead = e - 1;
    if (lead < -14) lead = -14;
    float q = ldexpf(1.0f, lead - 10);
    return rintf(x / q) * q;
}
#endif

/* light_geometry for an imported emitter (main.swift:3165-3166). */
/* light_geometry (main.swift:3164-3167): 1 for the environment (a direction), the imported
 * emitter's cosine over distance squared otherwise.  A stored environment sample holds its direction
 * in `position` (restir_di_encode :4714). */
static real light_geometry(const vo_render_scene *sc, v3 p, const di_sample *c) {
    if (c->light == 1u) return 1;
    if (c->light < 2u) return 0;
    tri_t t = tri_of(sc, c->light - 2u);
    return imported_geometry(p, c->position, &t);
}
/* eval_bsdf for DIFFUSE (main.swift:2659-2665). */
static s4 diffuse_f(const vo_render_scene *sc, const lanes_t *l, const mat_t *m, const surf_t *s, v3 wo, v3 wi) {
    if (dot(wi, s->ng) <= 0) return S4(0);
    real c = RFABS(dot(s->n, wi));
    if (c < (real)1e-7 || dot(s->n, wo) <= 0) return S4(0);
    return dot(s->n, wi) > 0 ? sscale(reflectance(sc, l, m->lagrange), (real)1 / PI_R) : S4(0);
}
/* eval_restir_target_pdf (main.swift:3187-3196): |rgb(f L cos)| G at the sample's own wavelengths. */
/* light_spectrum's RGB input (main.swift:2181-2186): an environment sample re-reads the map. */
static v3 di_emission(const vo_render_scene *sc, const di_sample *c) {
    return c->light == 1u ? eval_environment(sc, c->position) : c->emission;
}
static real target_pdf(const vo_render_scene *sc, const surf_t *s, const mat_t *m, v3 raydir, const di_sample *c) {
    if (c->pdf <= 0) return 0;
    v3 dir = c->light == 1u ? c->position : norm(sub(c->position, s->p));
    real ct = RFMAX((real)0, dot(s->n, dir));
    if (ct <= 0) return 0;
    lanes_t l = lanes_of(sc, c->u);
    s4 f = diffuse_f(sc, &l, m, s, neg(raydir), dir);
    v3 rgb = spectrum_rgb(sc, &l, sscale(smul(f, emission_spectrum(sc, &l, di_emission(sc, c))), ct));
    return len(rgb) * light_geometry(sc, s->p, c);
}
static di_sample stored_sample(const di_texel *t) {
    di_sample c;
    c.position = ld(t->pos);
    c.light = t->light;
    c.emission = ld(t->emit_pdf);
    c.pdf = (real)t->emit_pdf[3];
    c.u = (real)t->weights[3];
    return c;
}
/* restir_backproject with the previous sample's camera: the pixel whose unjittered frustum holds p. */
static int backproject(const cam_t *prev, const vo_frame *fr, v3 p, int *px, int *py) {
    v3 d = sub(p, prev->eye);
    real w = dot(d, prev->fw);
    if (!(w > 0)) return 0;
    real x = dot(d, prev->rt) / (w * prev->aspect * prev->fov_scale);
    real y = dot(d, prev->up) / (w * prev->fov_scale);
    real ux = x * (real)0.5 + (real)0.5, uy = y * (real)-0.5 + (real)0.5;
    if (!(ux >= 0 && uy >= 0 && ux < 1 && uy < 1)) return 0;
    *px = (int)(ux * (real)fr->width);
    *py = (int)(uy * (real)fr->height);
    return *px >= 0 && *px < (int)fr->width && *py >= 0 && *py < (int)fr->height;
}

/* ReSTIR GI (IndirectReuse.restirGI, main.swift:3200-3236, 4625, 4718-4812, 7586-7657): the reservoir
 * of a diffuse primary hit x1 holds a secondary point x2 and its one-sample outgoing direct radiance
 * toward x1.  Stored as the reference's textures: position + source pdf and radiance and weights in
 * rgba32f, the normal (+ a valid flag) in rgba16f. */
typedef struct {
    float pos[4];      /* x2, source pdf (area measure) */
    float nrm[4];      /* n2 rounded to half precision, 1 if a sample is held */
    float rad[4];      /* secondary radiance at the sample's four wavelengths */
    float weights[4];  /* weight sum, M, W, wavelength number */
} gi_texel;
typedef struct { v3 position, normal; s4 radiance; real source_pdf, wsum, M, u; } gi_res;

/* gi_geometry (main.swift:3201-3207). */
static real gi_geometry(v3 x1, v3 n1, v3 x2, v3 n2) {
    v3 delta = sub(x2, x1);
    real d2 = dot(delta, delta);
    if (d2 < (real)1e-10) return 0;
    v3 dir = mulv(delta, (real)1 / RSQRT(d2));
    return RFMAX((real)0, dot(n1, dir)) * RFMAX((real)0, dot(n2, neg(dir))) / d2;
}
/* restir_gi_jacobian / restir_gi_accepts_shift (main.swift:3213-3226). */
static int gi_accepts_shift(v3 x1r, v3 x1q, v3 x2, v3 n2) {
    v3 cur = sub(x1r, x2), src = sub(x1q, x2);
    real cd2 = dot(cur, cur), sd2 = dot(src, src);
    if (cd2 < (real)1e-10 || sd2 < (real)1e-10) return 0;
    real cc = RFMAX((real)0, dot(n2, mulv(cur, (real)1 / RSQRT(cd2))));
    real sco = RFMAX((real)0, dot(n2, mulv(src, (real)1 / RSQRT(sd2))));
    real j = cc * sd2 / (sco * cd2);
    if (!isfinite(j)) j = 0;
    return j >= (real)0.1 && j <= (real)10;
}
/* eval_restir_gi_target (main.swift:3230-3236) at a diffuse x1, the sample's wavelengths lanes(u). */
static real gi_target(const vo_render_scene *sc, const surf_t *s, const mat_t *m, v3 raydir, v3 x2, v3 n2, s4 rad, real u) {
    real g = gi_geometry(s->p, s->n, x2, n2);
    if (g <= 0 || !(isfinite(rad.v[0]) && isfinite(rad.v[1]) && isfinite(rad.v[2]) && isfinite(rad.v[3]))) return 0;
    v3 dir = norm(sub(x2, s->p));
    lanes_t l = lanes_of(sc, u);
    s4 f = diffuse_f(sc, &l, m, s, neg(raydir), dir);
    return len(spectrum_rgb(sc, &l, sscale(smul(f, rad), g)));
}
/* gi_connection_visible (main.swift:3238-3251): x1 sees x2. */
static int gi_visible(const vo_render_scene *sc, const surf_t *s, v3 x2) {
    v3 delta = sub(x2, s->p);
    real dist = len(delta);
    if (dist <= (real)2 * ray_epsilon(s->p)) return 0;
    v3 dir = mulv(delta, (real)1 / dist);
    v3 o = ray_origin(s->p, s->ng, dir, s->error);
    v3 ed = sub(x2, o);
    real d = len(ed);
    return !occluded(sc, o, mulv(ed, (real)1 / d), ray_t_min(o), d - endpoint_tolerance(o, x2, d));
}
static s4 gi_rad(const gi_texel *t) { s4 r = {{t->rad[0], t->rad[1], t->rad[2], t->rad[3]}}; return r; }

/* restir_gi_initial (main.swift:4722-4776): x1 -> x2 (BSDF sample, traced) -> a light sample at x2. */
static gi_res gi_initial(const vo_render_scene *sc, const vo_frame *fr, const surf_t *s, const mat_t *m, v3 view,
                         real u, rng_t *seed) {
    const int limit = fr->max_depth > 1 ? fr->max_depth - 1 : 0;
    gi_res r;
    memset(&r, 0, sizeof r);
    r.M = 1;
    r.u = u;
    lanes_t l = lanes_of(sc, u);
    v3 gdir;
    s4 bw;
    real bpdf;
    rng_event(seed, 1u, Z_GI_BSDF);
    if (!sample_bsdf(sc, &l, s, m, view, seed, &gdir, &bw, &bpdf) || !(bpdf > 0)) return r;
    v3 o = ray_origin(s->p, s->ng, gdir, s->error);
    surf_t s2;
    uint32_t face;
    if (!trace(sc, o, gdir, ray_t_min(o), &s2, &face)) return r;
    mat_t m2 = mat_of(sc, s2.slot);
    if (m2.type != 0u) return r;
    s4 rad = S4(0);
    rng_event(seed, 2u, Z_GI_NEE);
    light_sample ls = sample_direct_light(sc, s2.p, seed, limit >= 2 ? ENV_COMPENSATED : ENV_ORIGINAL);
    if (ls.pdf > 0 && light_visible_l(sc, &s2, ls.wi, ls.position, ls.light)) {
        real c2 = RFMAX((real)0, dot(s2.n, ls.wi));
        real pdf2;
        s4 f2 = eval_bsdf_pdf(sc, &l, &s2, &m2, neg(gdir), ls.wi, &pdf2);
        real mis = limit >= 2 ? power_heuristic(ls.pdf, pdf2) : 1;
        rad = sscale(smul(f2, emission_spectrum(sc, &l, ls.emission)), c2 * (mis / ls.pdf));
    }
    v3 gd = sub(s2.p, s->p);
    real d2 = dot(gd, gd);
    real cs = RFMAX((real)0, dot(s2.n, neg(gdir)));
    real src = d2 > (real)1e-10 ? bpdf * cs / d2 : 0;
    if (vo_get_fault() == 10) src = bpdf;  /* planted: the source pdf left in solid-angle measure */
    real ph = gi_target(sc, s, m, view, s2.p, s2.n, rad, u);
    if (src > 0 && ph > 0) {
        r.position = s2.p;
        r.normal = s2.n;
        r.radiance = rad;
        r.source_pdf = src;
        r.wsum = ph / src;
    }
    return r;
}

static void restir_pass1(const vo_render_scene *sc, const vo_frame *fr, const cam_t *cam, const cam_t *prev,
                         int temporal, uint32_t candidates, uint32_t gx, uint32_t gy, uint32_t sample_index,
                         const di_texel *hist, di_texel *cur, primary_t *prim, const gi_texel *ghist, gi_texel *gcur) {
    const uint32_t pix = gy * fr->width + gx;
    rng_t seed = {pixel_seed(fr, gx, gy, sample_index), 0u, 0u, fr->sampler == 1u, 0u, 0u};
    real u = pixel_u(seed.pcg);
    if (seed.z) {
        /* Z mode: the wavelength number as in sample_pixel; the light candidates below are
         * candidates of the vertex-1 Z_DI event, the resampling decisions stay on the PCG stream
         * (restir_di_initial, main.swift:4635-4652). */
        seed.key = vo_z_pixel_key(gx, gy, sample_index - fr->sample_index0 + 1u, sample_index, fr->z_temporal);
        rng_event(&seed, 0u, Z_WAVELENGTH);
        u = rr(&seed);
    }
    uint32_t bits = 0u;
    while ((1u << bits) < candidates) ++bits;
    const v3 dir = camera_dir(cam, fr, gx, gy, sample_index);
    di_texel *o = cur + pix;
    memset(o, 0, sizeof *o);
    gi_texel *g = gcur + pix;
    memset(g, 0, sizeof *g);
    primary_t *pr = prim + pix;
    uint32_t face;
    pr->dir = dir;
    pr->hit = trace(sc, cam->eye, dir, ray_t_min(cam->eye), &pr->s, &face);
    if (!pr->hit) {
        o->gpos[3] = -1.0f;
        o->gnorm[3] = -1.0f;
        return;
    }
    const surf_t *s = &pr->s;
    mat_t m = mat_of(sc, s->slot);
    o->gpos[0] = (float)s->p.x; o->gpos[1] = (float)s->p.y; o->gpos[2] = (float)s->p.z; o->gpos[3] = (float)s->t;
    o->gnorm[0] = half_round((float)s->n.x);
    o->gnorm[1] = half_round((float)s->n.y);
    o->gnorm[2] = half_round((float)s->n.z);
    o->gnorm[3] = half_round((float)m.type);
    if (m.type != 0u) return;
    di_sample sel;
    memset(&sel, 0, sizeof sel);
    real wsum = 0, M = 0;
    for (uint32_t i = 0; i < candidates; ++i) {
        if (seed.z) rng_candidate(&seed, 1u, Z_DI, i, bits);
        light_sample ls = sample_direct_light(sc, s->p, &seed, fr->max_depth > 1 ? ENV_COMPENSATED : ENV_ORIGINAL);
        di_sample c = {ls.light == 1u ? ls.wi : ls.position, ls.emission, ls.light, ls.pdf, u};
        if (vo_get_fault() != 4 || c.pdf > 0) M += 1;
        if (c.pdf > 0) {
            real ph = target_pdf(sc, s, &m, dir, &c);
            real proposal = c.pdf * light_geometry(sc, s->p, &c);
            real w = proposal > 0 ? ph / proposal : 0;
            wsum += w;
            if ((real)vo_rand_f(&seed.pcg) * wsum < w) sel = c;
        }
    }
    int px, py;
    const int back = temporal && backproject(prev, fr, s->p, &px, &py);
    if (back) {
        const di_texel *h = hist + (size_t)py * fr->width + px;
        real hM = (real)h->weights[1], hW = (real)h->weights[2];
        v3 hp = V(h->gpos[0], h->gpos[1], h->gpos[2]);
        v3 hn = V(h->gnorm[0], h->gnorm[1], h->gnorm[2]);
        int same = h->gpos[3] > 0.0f && h->gnorm[3] == (float)m.type && dot(hn, s->n) > (real)0.95 &&
                   len(sub(hp, s->p)) < RFMAX((real)0.01, s->t * (real)0.01);
        if (same && hM > 0 && hW > 0) {
            di_sample hs = stored_sample(h);
            real ph = target_pdf(sc, s, &m, dir, &hs);
            real cm = RFMIN(hM, (real)20);
            real w = ph * hW * cm;
            M += cm;
            wsum += w;
            if ((real)vo_rand_f(&seed.pcg) * wsum < w) sel = hs;
        }
    }
    real ph = target_pdf(sc, s, &m, dir, &sel);
    real W = (M > 0 && ph > 0) ? wsum / (M * ph) : 0;
    o->pos[0] = (float)sel.position.x; o->pos[1] = (float)sel.position.y; o->pos[2] = (float)sel.position.z;
    o->light = sel.light;
    o->emit_pdf[0] = (float)sel.emission.x; o->emit_pdf[1] = (float)sel.emission.y;
    o->emit_pdf[2] = (float)sel.emission.z; o->emit_pdf[3] = (float)sel.pdf;
    o->weights[0] = (float)wsum; o->weights[1] = (float)M; o->weights[2] = (float)W; o->weights[3] = (float)sel.u;
    /* ReSTIR GI (main.swift:4975-5005): depth 1 has no indirect bounce. */
    const int limit = fr->max_depth > 1 ? fr->max_depth - 1 : 0;
    if (fr->indirect_reuse != 0u || limit < 1) return;
    gi_res gr = gi_initial(sc, fr, s, &m, dir, u, &seed);
    if (back) {
        const di_texel *h = hist + (size_t)py * fr->width + px;
        const gi_texel *gh = ghist + (size_t)py * fr->width + px;
        v3 hp = V(h->gpos[0], h->gpos[1], h->gpos[2]);
        v3 hn = V(h->gnorm[0], h->gnorm[1], h->gnorm[2]);
        int same = h->gpos[3] > 0.0f && h->gnorm[3] == (float)m.type && dot(hn, s->n) > (real)0.95 &&
                   len(sub(hp, s->p)) < RFMAX((real)0.01, s->t * (real)0.01);
        if (same && gh->weights[1] > 0.0f && gh->weights[2] > 0.0f && gh->pos[3] > 0.0f && gh->nrm[3] > 0.0f) {
            /* restir_gi_merge (main.swift:4784-4801) */
            v3 op = ld(gh->pos), on = ld(gh->nrm);
            if (gi_accepts_shift(s->p, hp, op, on)) {
                real ph = gi_target(sc, s, &m, dir, op, on, gi_rad(gh), (real)gh->weights[3]);
                real cm = RFMIN((real)gh->weights[1], (real)20);
                real w = ph * (real)gh->weights[2] * cm;
                gr.M += cm;
                gr.wsum += w;
                if ((real)vo_rand_f(&seed.pcg) * gr.wsum < w) {
                    gr.position = op;
                    gr.normal = on;
                    gr.radiance = gi_rad(gh);
                    gr.source_pdf = (real)gh->pos[3];
                    gr.u = (real)gh->weights[3];
                }
            }
        }
    }
    /* restir_gi_encode (main.swift:4803-4811) */
    real gt = gi_target(sc, s, &m, dir, gr.position, gr.normal, gr.radiance, gr.u);
    real gW = gr.M > 0 && gt > 0 ? gr.wsum / (gr.M * gt) : 0;
    g->pos[0] = (float)gr.position.x; g->pos[1] = (float)gr.position.y; g->pos[2] = (float)gr.position.z;
    g->pos[3] = (float)gr.source_pdf;
    g->nrm[0] = half_round((float)gr.normal.x);
    g->nrm[1] = half_round((float)gr.normal.y);
    g->nrm[2] = half_round((float)gr.normal.z);
    g->nrm[3] = gr.source_pdf > 0 ? 1.0f : 0.0f;
    for (int k = 0; k < 4; ++k) g->rad[k] = (float)gr.radiance.v[k];
    g->weights[0] = (float)gr.wsum; g->weights[1] = (float)gr.M; g->weights[2] = (float)gW; g->weights[3] = (float)gr.u;
}

static v3 restir_pass2(const vo_render_scene *sc, const vo_frame *fr, uint32_t gx, uint32_t gy, uint32_t sample_index,
                       const di_texel *cur, const primary_t *prim, const gi_texel *gcur) {
    const int limit = fr->max_depth > 1 ? fr->max_depth - 1 : 0;
    const uint32_t pix = gy * fr->width + gx;
    rng_t seed = {pixel_seed(fr, gx, gy, sample_index), 0u, 0u, fr->sampler == 1u, 0u, 0u};
    real u = pixel_u(seed.pcg);
    if (seed.z) {
        seed.key = vo_z_pixel_key(gx, gy, sample_index - fr->sample_index0 + 1u, sample_index, fr->z_temporal);
        rng_event(&seed, 0u, Z_WAVELENGTH);
        u = rr(&seed);
    }
    lanes_t l = lanes_of(sc, u);
    seed.pcg = vo_pcg_hash(seed.pcg ^ 0x27d4eb2fu);
    const di_texel *c = cur + pix;
    v3 radiance = V(0, 0, 0);
    if (c->gpos[3] <= 0.0f) return eval_environment(sc, prim[pix].dir);
    const surf_t s = prim[pix].s;
    const v3 dir = prim[pix].dir;
    mat_t m = mat_of(sc, s.slot);
    if (m.type == 3) return s.front ? m.emission : V(0, 0, 0);
    if (m.type == 0u) {
        di_sample sel = stored_sample(c);
        real wsum = target_pdf(sc, &s, &m, dir, &sel) * (real)c->weights[2] * (real)c->weights[1];
        real M = (real)c->weights[1];
        for (int i = 0; i < 4; ++i) {
            real rx = (real)vo_rand_f(&seed.pcg), ry = (real)vo_rand_f(&seed.pcg);
            int nx = (int)gx + (int)((rx * (real)2 - (real)1) * (real)16);
            int ny = (int)gy + (int)((ry * (real)2 - (real)1) * (real)16);
            if (nx < 0 || nx >= (int)fr->width || ny < 0 || ny >= (int)fr->height) continue;
            const di_texel *n = cur + (size_t)ny * fr->width + nx;
            v3 nn = V(n->gnorm[0], n->gnorm[1], n->gnorm[2]);
            int compatible = n->gpos[3] > 0.0f && n->gnorm[3] == (float)m.type && dot(s.n, nn) > (real)0.95 &&
                             RFABS((real)c->gpos[3] - (real)n->gpos[3]) < (real)0.05 * (real)c->gpos[3];
            if (!compatible || !(n->weights[1] > 0.0f && n->weights[2] > 0.0f)) continue;
            di_sample ns = stored_sample(n);
            real w = target_pdf(sc, &s, &m, dir, &ns) * (real)n->weights[2] * (real)n->weights[1];
            wsum += w;
            if (vo_get_fault() != 3) M += (real)n->weights[1];
            if ((real)vo_rand_f(&seed.pcg) * wsum < w) sel = ns;
        }
        real fp = target_pdf(sc, &s, &m, dir, &sel);
        real W = (M > 0 && fp > 0) ? wsum / (M * fp) : 0;
        if (W > 0) {
            v3 wi = sel.light == 1u ? sel.position : norm(sub(sel.position, s.p));
            if (light_visible_l(sc, &s, wi, sel.position, sel.light)) {
                real ct = RFMAX((real)0, dot(s.n, wi));
                lanes_t sl = lanes_of(sc, sel.u);
                s4 f = diffuse_f(sc, &sl, &m, &s, neg(dir), wi);
                s4 v = sscale(sscale(smul(sscale(f, ct), emission_spectrum(sc, &sl, di_emission(sc, &sel))), W),
                              light_geometry(sc, s.p, &sel));
                radiance = add(radiance, spectrum_rgb(sc, &sl, v));
            }
        }
        /* Spatial ReSTIR GI reuse, uniform taps (main.swift:7586-7657). */
        if (fr->indirect_reuse == 0u && limit >= 1) {
            const gi_texel *gc = gcur + pix;
            v3 spos = ld(gc->pos), snrm = ld(gc->nrm);
            s4 srad = gi_rad(gc);
            real su = (real)gc->weights[3];
            int sflag = gc->nrm[3] > 0.0f;
            real gt = sflag ? gi_target(sc, &s, &m, dir, spos, snrm, srad, su) : 0;
            real gws = gt * (real)gc->weights[2] * (real)gc->weights[1];
            real gM = (real)gc->weights[1];
            /* Test hook 9 turns the GI taps off (the estimator of initial + temporal reuse alone). */
            const int taps = vo_get_fault() == 9 ? 0 : 4;
            for (int i = 0; i < taps; ++i) {
                real rx = (real)vo_rand_f(&seed.pcg), ry = (real)vo_rand_f(&seed.pcg);
                int nx = (int)gx + (int)((rx * (real)2 - (real)1) * (real)16);
                int ny = (int)gy + (int)((ry * (real)2 - (real)1) * (real)16);
                if (nx < 0 || nx >= (int)fr->width || ny < 0 || ny >= (int)fr->height) continue;
                const di_texel *n = cur + (size_t)ny * fr->width + nx;
                const gi_texel *gn = gcur + (size_t)ny * fr->width + nx;
                v3 nn = V(n->gnorm[0], n->gnorm[1], n->gnorm[2]);
                int compatible = n->gpos[3] > 0.0f && n->gnorm[3] == (float)m.type && dot(s.n, nn) > (real)0.95 &&
                                 RFABS((real)c->gpos[3] - (real)n->gpos[3]) < (real)0.05 * (real)c->gpos[3];
                if (!compatible) continue;
                if (gn->weights[1] <= 0.0f || gn->weights[2] <= 0.0f || gn->pos[3] <= 0.0f || gn->nrm[3] <= 0.0f) continue;
                v3 np = ld(gn->pos), nnrm = ld(gn->nrm);
                if (!gi_accepts_shift(s.p, V(n->gpos[0], n->gpos[1], n->gpos[2]), np, nnrm)) continue;
                real t = gi_target(sc, &s, &m, dir, np, nnrm, gi_rad(gn), (real)gn->weights[3]);
                real w = t * (real)gn->weights[2] * (real)gn->weights[1];
                gws += w;
                gM += (real)gn->weights[1];
                if ((real)vo_rand_f(&seed.pcg) * gws < w) {
                    spos = np;
                    snrm = nnrm;
                    srad = gi_rad(gn);
                    su = (real)gn->weights[3];
                    sflag = 1;
                }
            }
            real ft = sflag ? gi_target(sc, &s, &m, dir, spos, snrm, srad, su) : 0;
            real gW = gM > 0 && ft > 0 ? gws / (gM * ft) : 0;
            if (gW > 0 && gi_visible(sc, &s, spos)) {
                v3 wi = norm(sub(spos, s.p));
                lanes_t sl = lanes_of(sc, su);
                s4 f = diffuse_f(sc, &sl, &m, &s, neg(dir), wi);
                real geo = gi_geometry(s.p, s.n, spos, snrm);
                radiance = add(radiance, spectrum_rgb(sc, &sl, sscale(smul(f, srad), geo * gW)));
            }
        }
    } else if (!is_delta(&m)) {
        rng_event(&seed, 1u, Z_NEE);
        radiance = nee(sc, &l, &s, &m, neg(dir), S4(1), limit > 0, limit > 0 ? ENV_COMPENSATED : ENV_ORIGINAL, &seed);
    }
    radiance = add(radiance, continue_path_gi(sc, fr, &l, s, m, dir, &seed,
                                              m.type == 0u && fr->indirect_reuse == 0u && limit >= 1));
    if (!(isfinite(radiance.x) && isfinite(radiance.y) && isfinite(radiance.z))) radiance = V(0, 0, 0);
    return radiance;
}

void SUFFIX(vo_render_restir)(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                              uint32_t candidates, const float *eyes, int threads, double *out_rgb) {
    const size_t npix = (size_t)fr->width * fr->height;
    di_texel *buf[2] = {calloc(npix, sizeof(di_texel)), calloc(npix, sizeof(di_texel))};
    gi_texel *gbuf[2] = {calloc(npix, sizeof(gi_texel)), calloc(npix, sizeof(gi_texel))};
    primary_t *prim = calloc(npix, sizeof(primary_t));
    double *acc = calloc(3 * npix, sizeof(double));
    if (!buf[0] || !buf[1] || !gbuf[0] || !gbuf[1] || !prim || !acc) {
        free(buf[0]); free(buf[1]); free(gbuf[0]); free(gbuf[1]); free(prim); free(acc);
        return;
    }
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    vo_frame frame = *fr;
    frame.sample_index0 = sample_index0;  /* the accumulation starts at sample_index0 */
    fr = &frame;
    cam_t prev = camera_of(fr, fr->eye);
    (void)vo_energy_table(0);
    for (uint32_t k = 0; k < samples; ++k) {
        const uint32_t si = sample_index0 + k;
        const cam_t cam = camera_of(fr, eyes ? eyes + 3 * (size_t)k : fr->eye);
        di_texel *cur = buf[k & 1u];
        const di_texel *hist = buf[(k + 1u) & 1u];
        gi_texel *gcur = gbuf[k & 1u];
        const gi_texel *ghist = gbuf[(k + 1u) & 1u];
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
        for (int64_t p = 0; p < (int64_t)npix; ++p)
            restir_pass1(sc, fr, &cam, &prev, k > 0, candidates, (uint32_t)(p % fr->width), (uint32_t)(p / fr->width),
                         si, hist, cur, prim, ghist, gcur);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
        for (int64_t p = 0; p < (int64_t)npix; ++p) {
            v3 c = restir_pass2(sc, fr, (uint32_t)(p % fr->width), (uint32_t)(p / fr->width), si, cur, prim, gcur);
            acc[3 * p] += (double)c.x;
            acc[3 * p + 1] += (double)c.y;
            acc[3 * p + 2] += (double)c.z;
        }
        prev = cam;
    }
    for (size_t q = 0; q < 3 * npix; ++q) out_rgb[q] = samples ? acc[q] / (double)samples : 0.0;
    free(buf[0]); free(buf[1]); free(gbuf[0]); free(gbuf[1]); free(prim); free(acc);
}
