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
#include <string.h>

#include "oracle.h"

#ifndef REAL
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
typedef struct { v3 a, b, c, na, nb, nc; uint32_t tag; } tri_t;
static tri_t tri_of(const vo_render_scene *sc, uint32_t f) {
    const uint32_t *fc = sc->geo->faces + 4 * (size_t)f;
    const float *va = sc->geo->vertices + 8 * (size_t)fc[0], *vb = sc->geo->vertices + 8 * (size_t)fc[1],
                *vc = sc->geo->vertices + 8 * (size_t)fc[2];
    tri_t t;
    t.a = ld(va); t.b = ld(vb); t.c = ld(vc);
    t.na = ld(va + 3); t.nb = ld(vb + 3); t.nc = ld(vc + 3);
    t.tag = fc[3];
    return t;
}
typedef struct { uint32_t type; const float *lagrange; v3 emission; real weight; } mat_t;
static mat_t mat_of(const vo_render_scene *sc, uint32_t slot) {
    static const float grey[4] = {0.7f, 0, 0, 2.0f};
    mat_t m;
    if (slot >= sc->material_count) {
        m.type = 0; m.lagrange = grey; m.emission = V(0, 0, 0); m.weight = 0;
        return m;
    }
    const float *r = sc->materials + 16 * (size_t)slot;
    memcpy(&m.type, r, 4);
    m.lagrange = r + 4;
    m.emission = ld(r + 8);
    m.weight = (real)r[11];
    if (r[8] > 0 || r[9] > 0 || r[10] > 0) m.type = 3;
    return m;
}
static real imported_geometry(v3 p, v3 pos, const tri_t *t) {
    v3 d = sub(pos, p);
    real d2 = dot(d, d);
    v3 n = cross(sub(t->b, t->a), sub(t->c, t->a));
    return d2 > (real)1e-12 && dot(n, n) > (real)1e-20 ? RFMAX((real)0, dot(norm(n), mulv(neg(d), (real)1 / RSQRT(d2)))) / d2 : 0;
}
static float f_of(real x) { return (float)x; }

typedef struct { v3 p, n, ng; real error; int front; uint32_t slot; } surf_t;
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
    s->error = RFMAX((real)1e-7, (real)1.907349e-6 * RFMAX(maxabs(t.a), RFMAX(maxabs(t.b), maxabs(t.c))));
    s->slot = t.tag;
    *face_out = (uint32_t)h.face;
    return 1;
}
static int occluded(const vo_render_scene *sc, v3 o, v3 d, real tmin, real tmax) {
    float fo[3] = {f_of(o.x), f_of(o.y), f_of(o.z)}, fd[3] = {f_of(d.x), f_of(d.y), f_of(d.z)};
    return sc->geo->nodes ? vo_blas_any(sc->geo, fo, fd, f_of(tmin), f_of(tmax), 0, 0)
                          : vo_brute_any(sc->geo, fo, fd, f_of(tmin), f_of(tmax));
}

/* Next-event estimate at surface s: draws 4 numbers, returns the contribution if visible. */
static v3 nee(const vo_render_scene *sc, const lanes_t *l, const surf_t *s, const mat_t *m, v3 wo, s4 thr,
              int use_mis, uint32_t *seed) {
    v3 zero = V(0, 0, 0);
    if (sc->emitter_count == 0) return zero;
    if (!(vo_rand_f(seed) < 1.0f)) return zero;  /* importedProbability = 1 */
    real value = (real)vo_rand_f(seed), ry = (real)vo_rand_f(seed), rz = (real)vo_rand_f(seed);
    uint32_t face;
    float tree_pmf = 0.0f;
    if (sc->tree) {
        float fp[3] = {f_of(s->p.x), f_of(s->p.y), f_of(s->p.z)};
        uint32_t position = vo_light_tree_sample(sc->tree, fp, f_of(value), &tree_pmf);
        if (position == 0xffffffffu) return zero;
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
    v3 pos = add(add(mulv(t.a, (real)1 - root), mulv(t.b, root * ((real)1 - rz))), mulv(t.c, root * rz));
    v3 delta = sub(pos, s->p);
    real dist = len(delta);
    v3 wi = mulv(delta, (real)1 / RFMAX(dist, (real)1e-8));
    real geometry = imported_geometry(s->p, pos, &t);
    mat_t em = mat_of(sc, t.tag);
    real area_pdf;
    if (sc->tree) {
        real area = (real)0.5 * len(cross(sub(t.b, t.a), sub(t.c, t.a)));
        area_pdf = area > 0 ? (real)tree_pmf / area : 0;
    } else {
        area_pdf = sc->emitter_total > 0 ? em.weight / (real)sc->emitter_total : 0;
    }
    real pdf = geometry > (real)1e-12 ? area_pdf / geometry : 0;
    if (!(pdf > 0)) return zero;
    /* light_visible */
    v3 o = ray_origin(s->p, s->ng, wi, s->error);
    v3 dd = sub(pos, o);
    real d = len(dd), tol = endpoint_tolerance(o, pos, d);
    if (d <= tol) return zero;
    if (occluded(sc, o, norm(dd), ray_t_min(o), d - tol)) return zero;
    /* eval_bsdf_with_pdf, DIFFUSE */
    real bpdf = dot(s->n, wo) <= 0 ? 0 : RFMAX((real)0, dot(s->n, wi)) / PI_R;
    real cosine = RFABS(dot(s->n, wi));
    if (!(dot(wi, s->ng) > 0) || cosine < (real)1e-7 || dot(s->n, wo) <= 0 || !(dot(s->n, wi) > 0)) return zero;
    s4 f = sscale(reflectance(sc, l, m->lagrange), (real)1 / PI_R);
    real w = use_mis ? power_heuristic(pdf, bpdf) : 1;
    s4 le = emission_spectrum(sc, l, em.emission);
    return spectrum_rgb(sc, l, sscale(smul(smul(thr, f), le), cosine * w / pdf));
}

/* One sample of one pixel: the reference's shading_kernel in MIS mode (samplingMode 1). */
static v3 sample_pixel(const vo_render_scene *sc, const vo_frame *fr, const real *camera, uint32_t gx, uint32_t gy,
                       uint32_t sample_index) {
    const int limit = fr->max_depth > 1 ? fr->max_depth - 1 : 0;
    uint32_t seed = (gy * fr->width + gx) ^ (sample_index * 1999999973u);
    uint32_t h = vo_pcg_hash(seed ^ 0x5bd1e995u);
    real wl = (real)(h >> 8u) * (real)(1.0 / 16777216.0);
    lanes_t l = lanes_of(sc, wl);
    /* frameJitter */
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
    u *= camera[12] * camera[13];
    v *= camera[13];
    v3 eye = V(camera[0], camera[1], camera[2]), fw = V(camera[3], camera[4], camera[5]),
       rt = V(camera[6], camera[7], camera[8]), up = V(camera[9], camera[10], camera[11]);
    v3 dir = norm(add(add(fw, mulv(rt, u)), mulv(up, v)));
    seed = vo_pcg_hash(seed ^ 0x27d4eb2fu);

    v3 radiance = V(0, 0, 0);
    surf_t s;
    uint32_t face;
    if (!trace(sc, eye, dir, ray_t_min(eye), &s, &face)) return radiance;
    mat_t m = mat_of(sc, s.slot);
    if (m.type == 3) return s.front ? m.emission : V(0, 0, 0);
    s4 thr = S4(1);
    radiance = nee(sc, &l, &s, &m, neg(dir), thr, fr->sampling_mode <= 1u && limit > 0, &seed);
    v3 cur_dir = dir;
    int sdepth = 0;
    for (int bounce = 1;; ++bounce) {
        /* Departure from s (Lambert: never delta). */
        if (sdepth >= limit) break;
        ++sdepth;
        vo_rand_f(&seed);
        real r1 = (real)vo_rand_f(&seed), r2 = (real)vo_rand_f(&seed);
        v3 nd = cosine_hemisphere(s.n, r1, r2);
        real bpdf = RFMAX((real)0, dot(s.n, nd)) / PI_R;
        if (!(bpdf > 0 && dot(nd, s.ng) > 0)) break;
        thr = smul(thr, reflectance(sc, &l, m.lagrange));
        if (!(isfinite(thr.v[0]) && isfinite(thr.v[1]) && isfinite(thr.v[2]) && isfinite(thr.v[3])) || smax(thr) <= 0) break;
        v3 prev = s.p;
        v3 o = ray_origin(s.p, s.ng, nd, s.error);
        cur_dir = nd;
        surf_t r;
        if (!trace(sc, o, nd, ray_t_min(o), &r, &face)) break;
        mat_t rm = mat_of(sc, r.slot);
        if (rm.type == 3) {
            tri_t t = tri_of(sc, face);
            real g = imported_geometry(prev, r.p, &t);
            real light_pdf = 0;
            if (sc->tree) {
                /* imported_emitter_area_pdf with a tree (main.swift:2853-2860) */
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
                    light_pdf = area > 0 ? pmf / area / g : 0;
                }
            } else if (sc->emitter_count && sc->emitter_total > 0 && rm.weight > 0 && g > (real)1e-12) {
                light_pdf = (rm.weight / (real)sc->emitter_total) / g;
            }
            real w = emission_weight(0, fr->sampling_mode != 3u, fr->sampling_mode == 1u, bpdf, light_pdf);
            v3 e = r.front ? rm.emission : V(0, 0, 0);
            radiance = add(radiance, spectrum_rgb(sc, &l, sscale(smul(thr, emission_spectrum(sc, &l, e)), w)));
            break;
        }
        radiance = add(radiance, nee(sc, &l, &r, &rm, neg(cur_dir), thr, fr->sampling_mode <= 1u && sdepth < limit, &seed));
        s = r;
        m = rm;
        if (bounce > 3) {
            real survival = RFMIN(RFMAX(smax(thr), (real)0.05), (real)0.95);
            if ((real)vo_rand_f(&seed) >= survival) break;
            thr = sscale(thr, (real)1 / survival);
        }
    }
    if (!(isfinite(radiance.x) && isfinite(radiance.y) && isfinite(radiance.z))) radiance = V(0, 0, 0);
    return radiance;
}

void SUFFIX(vo_render)(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                       const uint32_t *pixels, uint32_t npixels, int threads, double *out_rgb) {
    /* Camera basis (host side of the reference, main.swift:11303-11346, and pass 2's frame). */
    real camera[14];
    v3 eye = V(fr->eye[0], fr->eye[1], fr->eye[2]);
    v3 fw = norm(sub(V(fr->target[0], fr->target[1], fr->target[2]), eye));
    v3 rt = norm(cross(fw, V(0, 1, 0)));
    v3 up = cross(rt, fw);
    camera[0] = eye.x; camera[1] = eye.y; camera[2] = eye.z;
    camera[3] = fw.x; camera[4] = fw.y; camera[5] = fw.z;
    camera[6] = rt.x; camera[7] = rt.y; camera[8] = rt.z;
    camera[9] = up.x; camera[10] = up.y; camera[11] = up.z;
    camera[12] = (real)fr->width / (real)fr->height;
    camera[13] = RTAN((real)(fr->fov * 0.5f) * PI_R / (real)180);
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int64_t k = 0; k < (int64_t)npixels; ++k) {
        uint32_t pix = pixels[k], gx = pix % fr->width, gy = pix / fr->width;
        double acc[3] = {0, 0, 0};
        for (uint32_t s = 0; s < samples; ++s) {
            v3 c = sample_pixel(sc, fr, camera, gx, gy, sample_index0 + s);
            acc[0] += (double)c.x;
            acc[1] += (double)c.y;
            acc[2] += (double)c.z;
        }
        for (int q = 0; q < 3; ++q) out_rgb[3 * k + q] = acc[q] / (double)samples;
    }
}
