/*
 * woop.c -- oracle (test infrastructure only): PCG hash and the watertight ray/triangle test.
 * Compiled with -ffp-contract=off: every product below rounds on its own, as the reference's
 * `#pragma METAL fp contract(off)` requires for the edge functions (main.swift:1558).
 */
#include <math.h>
#include "oracle.h"

static int g_fault = 0;
void vo_set_fault(int fault) { g_fault = fault; }
int vo_get_fault(void) { return g_fault; }

/* main.swift:91-95 pcg_hash: PCG-RXS-M-XS output of one LCG step. */
uint32_t vo_pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

/* main.swift:97-101 rand_f: 24 high bits of the rehashed seed, scaled into [0, 1). */
float vo_rand_f(uint32_t *seed) {
    *seed = vo_pcg_hash(*seed);
    return (float)(*seed >> 8u) * (1.0f / 16777216.0f);
}

/* main.swift:1557-1562 watertight_edge: p.x*q.y - p.y*q.x with each product rounded; when the
 * rounded products tie, the exact rounding errors (FMA residuals) decide the sign. */
float vo_watertight_edge(float px, float py, float qx, float qy) {
    float x = px * qy, y = py * qx, e = x - y;
    if (e == 0.0f && g_fault != 1) e = fmaf(px, qy, -x) - fmaf(py, qx, -y);
    return e;
}

/* main.swift:1537-1550 watertight_ray: kz = axis of largest |d| (ties as the ternary resolves
 * them), kx/ky the next two cyclically, swapped when d[kz] < 0; shear S = d[k]/d[kz], scale
 * 1/d[kz].  The rows x = e_kx - S_x e_kz, y = e_ky - S_y e_kz, z = e_kz / d[kz] have exact
 * zero/unit entries, so dot(row, A) reduces to A[kx] - fl(S_x A[kz]) (no fusion). */
typedef struct { int kx, ky, kz; float sx, sy, sz; } wray;

static wray make_wray(const float *d) {
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    if (d[kz] < 0) { int s = kx; kx = ky; ky = s; }
    wray w;
    w.kx = kx; w.ky = ky; w.kz = kz;
    w.sx = d[kx] / d[kz];
    w.sy = d[ky] / d[kz];
    w.sz = 1.0f / d[kz];
    return w;
}

/* main.swift:1564-1578 intersect_mesh_triangle: two-sided; edge values U (c,b), V (a,c),
 * W (b,a); a miss when signs differ; det = U+V+W, 0 rejects; t = T/det with
 * T = U*Az + V*Bz + W*Cz evaluated left to right; accept t in [tmin, tmax). */
static int wtri(const wray *w, const float *o, const float *pa, const float *pb, const float *pc,
                float tmin, float tmax, float *t, float *b1, float *b2) {
    float A[3] = {pa[0] - o[0], pa[1] - o[1], pa[2] - o[2]};
    float B[3] = {pb[0] - o[0], pb[1] - o[1], pb[2] - o[2]};
    float C[3] = {pc[0] - o[0], pc[1] - o[1], pc[2] - o[2]};
    float ax = A[w->kx] - w->sx * A[w->kz], ay = A[w->ky] - w->sy * A[w->kz];
    float bx = B[w->kx] - w->sx * B[w->kz], by = B[w->ky] - w->sy * B[w->kz];
    float cx = C[w->kx] - w->sx * C[w->kz], cy = C[w->ky] - w->sy * C[w->kz];
    float U = vo_watertight_edge(cx, cy, bx, by);
    float V = vo_watertight_edge(ax, ay, cx, cy);
    if ((U < 0 && V > 0) || (U > 0 && V < 0)) return 0;
    float W = vo_watertight_edge(bx, by, ax, ay);
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return 0;
    float det = U + V + W;
    if (det == 0.0f) return 0;
    float az = w->sz * A[w->kz], bz = w->sz * B[w->kz], cz = w->sz * C[w->kz];
    float T = U * az + V * bz + W * cz;
    float inv = 1.0f / det;
    float tt = T * inv;
    if (!(tt >= tmin && tt < tmax)) return 0;
    *t = tt;
    *b1 = V * inv;
    *b2 = W * inv;
    if (g_fault == 2) { float s = *b1; *b1 = *b2; *b2 = s; }
    return 1;
}

int vo_intersect_triangle(const float *o, const float *d, const float *a, const float *b, const float *c,
                          float tmin, float tmax, float *t, float *b1, float *b2) {
    wray w = make_wray(d);
    return wtri(&w, o, a, b, c, tmin, tmax, t, b1, b2);
}

/* Exposed to traverse.c */
void vo__wray_make(const float *d, int *k, float *s) {
    wray w = make_wray(d);
    k[0] = w.kx; k[1] = w.ky; k[2] = w.kz;
    s[0] = w.sx; s[1] = w.sy; s[2] = w.sz;
}
int vo__wray_tri(const int *k, const float *s, const float *o, const float *a, const float *b, const float *c,
                 float tmin, float tmax, float *t, float *b1, float *b2) {
    wray w;
    w.kx = k[0]; w.ky = k[1]; w.kz = k[2];
    w.sx = s[0]; w.sy = s[1]; w.sz = s[2];
    return wtri(&w, o, a, b, c, tmin, tmax, t, b1, b2);
}
