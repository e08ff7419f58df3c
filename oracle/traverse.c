/*
 * traverse.c -- oracle (test infrastructure only): the reference's software 4-wide BLAS
 * traversal (main.swift:1760-1838) and a brute-force closest/any hit (tests/Fix_accel.swift:131-141),
 * batched over SoA rays with OpenMP.  Built with -ffp-contract=off.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include "oracle.h"

void vo__wray_make(const float *d, int *k, float *s);
int vo__wray_tri(const int *k, const float *s, const float *o, const float *a, const float *b, const float *c,
                 float tmin, float tmax, float *t, float *b1, float *b2);

static inline float max_abs3(const float *p) { return fmaxf(fabsf(p[0]), fmaxf(fabsf(p[1]), fabsf(p[2]))); }

/* main.swift:1760-1765 wide_ray: origins shifted by +-e, parallel axes |d| < 1e-8. */
typedef struct { float low[3], high[3], inv[3]; int par[3]; } wide_ray_t;
static wide_ray_t wide_ray(const float *o, const float *d, float e) {
    wide_ray_t w;
    for (int a = 0; a < 3; ++a) {
        w.low[a] = o[a] + e;
        w.high[a] = o[a] - e;
        w.par[a] = fabsf(d[a]) < 1e-8f;
        w.inv[a] = 1.0f / (w.par[a] ? 1.0f : d[a]);
    }
    return w;
}

/* main.swift:1766-1779 wide_node_hit */
static int wide_node_hit(const vo_wide_node *n, const wide_ray_t *w, float tmin, float tmax, float *nearT) {
    int mask = 0;
    for (int k = 0; k < 4; ++k) {
        const float lo[3] = {n->lox[k], n->loy[k], n->loz[k]}, hi[3] = {n->hix[k], n->hiy[k], n->hiz[k]};
        float l[3], h[3];
        for (int a = 0; a < 3; ++a) {
            float x = (lo[a] - w->low[a]) * w->inv[a], y = (hi[a] - w->high[a]) * w->inv[a];
            l[a] = w->par[a] ? -1e30f : fminf(x, y);
            h[a] = w->par[a] ? 1e30f : fmaxf(x, y);
        }
        float nt = fmaxf(fmaxf(l[0], l[1]), fmaxf(l[2], tmin));
        int hit = nt <= fminf(fminf(h[0], h[1]), fminf(h[2], tmax)) && n->child[k] != 0;
        for (int a = 0; a < 3; ++a)
            if (w->par[a]) hit = hit && !(w->low[a] < lo[a] || w->high[a] > hi[a]);
        nearT[k] = nt;
        if (hit) mask |= 1 << k;
    }
    return mask;
}

/* main.swift:1781-1785 wide_order: insertion sort near to far (stable). */
static int wide_order(int mask, const float *nearT, int *order, float *key) {
    int m = 0;
    for (int k = 0; k < 4; ++k) {
        if (!(mask >> k & 1)) continue;
        int j = m++;
        while (j > 0 && key[j - 1] > nearT[k]) { key[j] = key[j - 1]; order[j] = order[j - 1]; --j; }
        key[j] = nearT[k];
        order[j] = k;
    }
    return m;
}

static inline const float *vtx(const vo_scene *s, uint32_t i) { return s->vertices + 8 * (size_t)i; }

/* main.swift:1788-1812 blas_closest (no cutouts) */
vo_hit vo_blas_closest(const vo_scene *s, const float *o, const float *d, float tmin, float tmax,
                       uint64_t *nv, uint64_t *nt) {
    vo_hit h = {tmax, 0, 0, -1};
    if (!s->nodes || s->face_count == 0) return h;
    int K[3];
    float S[3];
    vo__wray_make(d, K, S);
    wide_ray_t w = wide_ray(o, d, 1.1920929e-6f * (max_abs3(o) + s->root_max));
    float closest = tmax;
    int stack[64], top = 0;
    uint64_t visits = 0, tests = 0;
    stack[top++] = 0;
    while (top > 0) {
        const vo_wide_node *node = &s->nodes[stack[--top]];
        ++visits;
        float nearT[4], key[4];
        int order[4];
        int m = wide_order(wide_node_hit(node, &w, tmin, closest, nearT), nearT, order, key);
        for (int j = 0; j < m; ++j) {
            int k = order[j];
            if (node->child[k] > 0 || key[j] > closest) continue;
            int first = -node->child[k] - 1;
            for (int q = 0; q < node->count[k]; ++q) {
                uint32_t f = (uint32_t)s->order[first + q];
                const uint32_t *fc = s->faces + 4 * (size_t)f;
                float t, b1, b2;
                ++tests;
                if (!vo__wray_tri(K, S, o, vtx(s, fc[0]), vtx(s, fc[1]), vtx(s, fc[2]), tmin, closest, &t, &b1, &b2))
                    continue;
                closest = t;
                h.t = t; h.b1 = b1; h.b2 = b2; h.face = (int32_t)f;
            }
        }
        for (int j = m - 1; j >= 0; --j) {
            int k = order[j];
            if (node->child[k] > 0 && key[j] <= closest && top < 64) stack[top++] = node->child[k];
        }
    }
    if (nv) *nv += visits;
    if (nt) *nt += tests;
    return h;
}

/* main.swift:1815-1838 blas_any (slack off, no cutouts) */
int vo_blas_any(const vo_scene *s, const float *o, const float *d, float tmin, float tmax, uint64_t *nv,
                uint64_t *nt) {
    if (!s->nodes || s->face_count == 0) return 0;
    int K[3];
    float S[3];
    vo__wray_make(d, K, S);
    wide_ray_t w = wide_ray(o, d, 1.1920929e-6f * (max_abs3(o) + s->root_max));
    int stack[64], top = 0, found = 0;
    uint64_t visits = 0, tests = 0;
    stack[top++] = 0;
    while (top > 0 && !found) {
        const vo_wide_node *node = &s->nodes[stack[--top]];
        ++visits;
        float nearT[4];
        int mask = wide_node_hit(node, &w, tmin, tmax, nearT);
        for (int k = 0; k < 4 && !found; ++k) {
            if (!(mask >> k & 1)) continue;
            if (node->child[k] > 0) { if (top < 64) stack[top++] = node->child[k]; continue; }
            int first = -node->child[k] - 1;
            for (int q = 0; q < node->count[k]; ++q) {
                uint32_t f = (uint32_t)s->order[first + q];
                const uint32_t *fc = s->faces + 4 * (size_t)f;
                float t, b1, b2;
                ++tests;
                if (vo__wray_tri(K, S, o, vtx(s, fc[0]), vtx(s, fc[1]), vtx(s, fc[2]), tmin, tmax, &t, &b1, &b2)) {
                    found = 1;
                    break;
                }
            }
        }
    }
    if (nv) *nv += visits;
    if (nt) *nt += tests;
    return found;
}

/* tests/Fix_accel.swift:131-141: every stored triangle, in document order, shrinking the
 * interval; strict t < best keeps the first (smallest) face index among equal t. */
vo_hit vo_brute_closest(const vo_scene *s, const float *o, const float *d, float tmin, float tmax) {
    vo_hit h = {tmax, 0, 0, -1};
    int K[3];
    float S[3];
    vo__wray_make(d, K, S);
    float best = tmax;
    for (uint32_t f = 0; f < s->face_count; ++f) {
        const uint32_t *fc = s->faces + 4 * (size_t)f;
        float t, b1, b2;
        if (!vo__wray_tri(K, S, o, vtx(s, fc[0]), vtx(s, fc[1]), vtx(s, fc[2]), tmin, best, &t, &b1, &b2)) continue;
        best = t;
        h.t = t; h.b1 = b1; h.b2 = b2; h.face = (int32_t)f;
    }
    return h;
}

int vo_brute_any(const vo_scene *s, const float *o, const float *d, float tmin, float tmax) {
    int K[3];
    float S[3];
    vo__wray_make(d, K, S);
    for (uint32_t f = 0; f < s->face_count; ++f) {
        const uint32_t *fc = s->faces + 4 * (size_t)f;
        float t, b1, b2;
        if (vo__wray_tri(K, S, o, vtx(s, fc[0]), vtx(s, fc[1]), vtx(s, fc[2]), tmin, tmax, &t, &b1, &b2)) return 1;
    }
    return 0;
}

void vo_trace_closest_batch(const vo_scene *s, const float *ox, const float *oy, const float *oz,
                            const float *dx, const float *dy, const float *dz, const float *tmin,
                            const float *tmax, int64_t n, int mode, int threads, float *t, float *b1,
                            float *b2, int32_t *face, uint64_t *counters) {
    uint64_t visits = 0, tests = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : visits, tests) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int64_t i = 0; i < n; ++i) {
        float o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
        vo_hit h = mode == 0 ? vo_blas_closest(s, o, d, tmin[i], tmax[i], &visits, &tests)
                             : vo_brute_closest(s, o, d, tmin[i], tmax[i]);
        t[i] = h.t; b1[i] = h.b1; b2[i] = h.b2; face[i] = h.face;
    }
    if (counters) { counters[0] += visits; counters[1] += tests; }
}

void vo_trace_any_batch(const vo_scene *s, const float *ox, const float *oy, const float *oz,
                        const float *dx, const float *dy, const float *dz, const float *tmin,
                        const float *tmax, int64_t n, int mode, int threads, uint8_t *occluded,
                        uint64_t *counters) {
    uint64_t visits = 0, tests = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : visits, tests) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int64_t i = 0; i < n; ++i) {
        float o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
        occluded[i] = (uint8_t)(mode == 0 ? vo_blas_any(s, o, d, tmin[i], tmax[i], &visits, &tests)
                                          : vo_brute_any(s, o, d, tmin[i], tmax[i]));
    }
    if (counters) { counters[0] += visits; counters[1] += tests; }
}

/* Threads an unbounded batch call uses (OpenMP's default team size). */
int vo_max_threads(void) { return omp_get_max_threads(); }
