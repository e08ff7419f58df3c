/*
 * lighttree.c -- oracle (test infrastructure only): the reference's light BVH.
 *
 * Builder: Sources/LightTree.swift:112-218 (LightTree.build / split / bucket / partition / node /
 * octahedral) with Bounds, coneUnion, angleBetween and cost (:19-83), in double precision.
 * Walk: main.swift:2743-2844 (light_tree_axis, light_tree_node, light_tree_importance,
 * light_tree_sample, light_tree_pmf), in float.  Built with -ffp-contract=off.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define BUCKETS 12
#define NO_PARENT 0x7ffffffu
static const double PI_D = 3.14159265358979323846;

typedef struct { double v[3]; } d3;
static d3 mk(double x, double y, double z) { d3 r = {{x, y, z}}; return r; }
static d3 dadd(d3 a, d3 b) { return mk(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
static d3 dsub(d3 a, d3 b) { return mk(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
static d3 dmul(d3 a, double s) { return mk(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
static double ddot(d3 a, d3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
static d3 dcross(d3 a, d3 b) {
    return mk(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
static double dlen(d3 a) { return sqrt(ddot(a, a)); }
static d3 dmin(d3 a, d3 b) { return mk(fmin(a.v[0], b.v[0]), fmin(a.v[1], b.v[1]), fmin(a.v[2], b.v[2])); }
static d3 dmax(d3 a, d3 b) { return mk(fmax(a.v[0], b.v[0]), fmax(a.v[1], b.v[1]), fmax(a.v[2], b.v[2])); }
static double clamp1(double c) { return fmin(1.0, fmax(-1.0, c)); }

typedef struct { d3 lo, hi; double phi; d3 w; double cos_o, cos_e; } lbounds;
static lbounds lb_empty(void) {
    lbounds b;
    b.lo = mk(INFINITY, INFINITY, INFINITY);
    b.hi = mk(-INFINITY, -INFINITY, -INFINITY);
    b.phi = 0;
    b.w = mk(0, 0, 1);
    b.cos_o = 1;
    b.cos_e = 1;
    return b;
}
static d3 centroid(const lbounds *b) { return dmul(dadd(b->lo, b->hi), 0.5); }

/* LightTree.swift:69-72 angleBetween */
static double angle_between(d3 a, d3 b) {
    if (ddot(a, b) < 0) return PI_D - 2 * asin(fmin(1, dlen(dadd(a, b)) / 2));
    return 2 * asin(fmin(1, dlen(dsub(b, a)) / 2));
}
/* LightTree.swift:55-67 coneUnion; the quaternion rotation as Rodrigues' formula. */
static void cone_union(d3 wa, double ca, d3 wb, double cb, d3 *w, double *c) {
    double ta = acos(clamp1(ca)), tb = acos(clamp1(cb)), td = angle_between(wa, wb);
    if (fmin(td + tb, PI_D) <= ta) { *w = wa; *c = ca; return; }
    if (fmin(td + ta, PI_D) <= tb) { *w = wb; *c = cb; return; }
    double to = (ta + td + tb) / 2;
    if (to >= PI_D) { *w = wa; *c = -1; return; }
    d3 axis = dcross(wa, wb);
    if (ddot(axis, axis) == 0) { *w = wa; *c = -1; return; }
    d3 k = dmul(axis, 1.0 / dlen(axis));
    double ang = to - ta, co = cos(ang), si = sin(ang);
    d3 r = dadd(dadd(dmul(wa, co), dmul(dcross(k, wa), si)), dmul(k, ddot(k, wa) * (1 - co)));
    *w = dmul(r, 1.0 / dlen(r));
    *c = cos(to);
}
/* LightTree.swift:41-51 union */
static lbounds lb_union(lbounds a, lbounds b) {
    if (a.phi == 0) return b;
    if (b.phi == 0) return a;
    lbounds r;
    r.lo = dmin(a.lo, b.lo);
    r.hi = dmax(a.hi, b.hi);
    r.phi = a.phi + b.phi;
    cone_union(a.w, a.cos_o, b.w, b.cos_o, &r.w, &r.cos_o);
    r.cos_e = fmin(a.cos_e, b.cos_e);
    return r;
}
/* LightTree.swift:76-88 cost (axis < 0: none) */
static double lb_cost(const lbounds *b, d3 diag, int axis) {
    if (!(b->phi > 0)) return 0;
    double to = acos(clamp1(b->cos_o)), te = acos(clamp1(b->cos_e));
    double tw = fmin(to + te, PI_D);
    double so = sqrt(fmax(0, 1 - b->cos_o * b->cos_o));
    double m = 2 * PI_D * (1 - b->cos_o) + PI_D / 2 * (2 * tw * so - cos(to - 2 * tw) - 2 * to * so + b->cos_o);
    double kr = 1;
    if (axis >= 0) kr = fmax(diag.v[0], fmax(diag.v[1], diag.v[2])) / fmax(diag.v[axis], 1e-300);
    d3 d = dsub(b->hi, b->lo);
    double area = 2 * (d.v[0] * d.v[1] + d.v[0] * d.v[2] + d.v[1] * d.v[2]);
    return b->phi * m * kr * area;
}
/* LightTree.swift:28-38 Bounds(Emitter) */
static lbounds lb_emitter(const float *tri, double power, double cos_emission) {
    d3 a = mk(tri[0], tri[1], tri[2]), b = mk(tri[3], tri[4], tri[5]), c = mk(tri[6], tri[7], tri[8]);
    lbounds r = lb_empty();
    r.lo = dmin(a, dmin(b, c));
    r.hi = dmax(a, dmax(b, c));
    d3 n = dcross(dsub(b, a), dsub(c, a));
    double len = dlen(n);
    r.phi = len > 0 && isfinite(power) ? fmax(0, power) : 0;
    r.w = len > 0 ? dmul(n, 1.0 / len) : mk(0, 0, 1);
    r.cos_o = 1;
    r.cos_e = fmin(1, fmax(0, cos_emission));
    return r;
}

typedef struct { d3 lo, hi; } cents;
/* LightTree.swift:175-178 bucket */
static int bucket(const lbounds *b, int axis, const cents *c) {
    double lo = c->lo.v[axis], extent = c->hi.v[axis] - lo;
    int k = (int)((double)BUCKETS * (centroid(b).v[axis] - lo) / extent);
    return k < BUCKETS - 1 ? k : BUCKETS - 1;
}

static uint32_t quant(double v, double lo, double extent, int up) {
    if (!(extent > 0) || !isfinite(v)) return up ? 65535u : 0u;
    double x = (v - lo) / extent * 65535;
    return (uint32_t)fmin(65535.0, fmax(0.0, up ? ceil(x) : floor(x)));
}
static uint32_t quant_cos(double c) { return (uint32_t)fmin(65535.0, fmax(0.0, floor((c + 1) / 2 * 65535))); }
/* LightTree.swift:208-215 octahedral */
static uint32_t octahedral(d3 v) {
    double s = fabs(v.v[0]) + fabs(v.v[1]) + fabs(v.v[2]);
    double px = v.v[0] / s, py = v.v[1] / s;
    if (v.v[2] < 0) {
        double nx = (1 - fabs(py)) * (px >= 0 ? 1 : -1), ny = (1 - fabs(px)) * (py >= 0 ? 1 : -1);
        px = nx;
        py = ny;
    }
    uint32_t ex = (uint32_t)fmin(65535.0, fmax(0.0, round((px * 0.5 + 0.5) * 65535)));
    uint32_t ey = (uint32_t)fmin(65535.0, fmax(0.0, round((py * 0.5 + 0.5) * 65535)));
    return ex | ey << 16;
}

typedef struct { uint32_t lo, hi, parent, second, level; } lwork;

int32_t vo_light_tree_build(const float *tris, const double *power, uint32_t count, int max_leaf, double leaf_ratio,
                            vo_light_tree *out) {
    memset(out, 0, sizeof *out);
    lbounds *bounds = (lbounds *)malloc((count ? count : 1) * sizeof(lbounds));
    d3 glo = mk(INFINITY, INFINITY, INFINITY), ghi = mk(-INFINITY, -INFINITY, -INFINITY);
    for (uint32_t i = 0; i < count; ++i) {
        bounds[i] = lb_emitter(tris + 9 * (size_t)i, power[i], 0.0);
        glo = dmin(glo, bounds[i].lo);
        ghi = dmax(ghi, bounds[i].hi);
    }
    for (int k = 0; k < 3 && count; ++k) {
        out->grid_lo[k] = nextafterf((float)glo.v[k], -INFINITY);
        out->grid_hi[k] = nextafterf((float)ghi.v[k], INFINITY);
    }
    d3 g_lo = mk(out->grid_lo[0], out->grid_lo[1], out->grid_lo[2]);
    d3 g_ext = dsub(mk(out->grid_hi[0], out->grid_hi[1], out->grid_hi[2]), g_lo);
    uint32_t cap_nodes = count ? 2 * count : 1;
    out->words = (uint32_t *)calloc((size_t)8 * cap_nodes, 4);
    out->order = (uint32_t *)malloc((count ? count : 1) * 4);
    out->leaf_node = (uint32_t *)calloc(count ? count : 1, 4);
    for (uint32_t i = 0; i < count; ++i) out->order[i] = i;
    lwork *work = (lwork *)malloc(sizeof(lwork) * (2 * (size_t)count + 2));
    int top = 0;
    if (count) { lwork w0 = {0, count, NO_PARENT, 0, 0}; work[top++] = w0; }
    while (top > 0) {
        lwork it = work[--top];
        uint32_t index = out->node_count++;
        if (it.level > out->depth) out->depth = it.level;
        if (it.second) out->words[8 * (size_t)it.parent + 6] = index;
        lbounds total = lb_empty();
        cents cen = {mk(INFINITY, INFINITY, INFINITY), mk(-INFINITY, -INFINITY, -INFINITY)};
        for (uint32_t i = it.lo; i < it.hi; ++i) {
            const lbounds *b = &bounds[out->order[i]];
            total = lb_union(total, *b);
            cen.lo = dmin(cen.lo, centroid(b));
            cen.hi = dmax(cen.hi, centroid(b));
        }
        if (total.phi == 0)
            for (uint32_t i = it.lo; i < it.hi; ++i) {
                total.lo = dmin(total.lo, bounds[out->order[i]].lo);
                total.hi = dmax(total.hi, bounds[out->order[i]].hi);
            }
        uint32_t n = it.hi - it.lo;
        /* LightTree.swift:153-171 split */
        int have = 0, best_axis = 0, best_bucket = 0;
        double best_cost = INFINITY;
        if (n > 1) {
            d3 diag = dsub(total.hi, total.lo);
            for (int axis = 0; axis < 3; ++axis) {
                if (!(cen.hi.v[axis] > cen.lo.v[axis])) continue;
                lbounds bb[BUCKETS], below[BUCKETS], above[BUCKETS];
                for (int k = 0; k < BUCKETS; ++k) bb[k] = lb_empty();
                for (uint32_t i = it.lo; i < it.hi; ++i) {
                    int k = bucket(&bounds[out->order[i]], axis, &cen);
                    bb[k] = lb_union(bb[k], bounds[out->order[i]]);
                }
                for (int k = 0; k < BUCKETS; ++k) below[k] = k == 0 ? bb[0] : lb_union(below[k - 1], bb[k]);
                for (int k = BUCKETS - 1; k >= 0; --k) above[k] = k == BUCKETS - 1 ? bb[k] : lb_union(above[k + 1], bb[k]);
                for (int k = 0; k < BUCKETS - 1; ++k) {
                    if (!(below[k].phi > 0 || above[k + 1].phi > 0)) continue;
                    double c = lb_cost(&below[k], diag, axis) + lb_cost(&above[k + 1], diag, axis);
                    if (c < best_cost) { have = 1; best_cost = c; best_axis = axis; best_bucket = k; }
                }
            }
        }
        double unsplit = lb_cost(&total, dsub(total.hi, total.lo), -1);
        int leaf = n == 1 || ((int)n <= max_leaf && (total.phi == 0 || (have ? best_cost : INFINITY) >= leaf_ratio * unsplit));
        uint32_t child = 0, leaf_count = 0;
        uint32_t mid = it.lo + n / 2;
        if (leaf) {
            for (uint32_t i = it.lo; i < it.hi; ++i) out->leaf_node[i] = index;
            child = it.lo;
            leaf_count = n;
        } else if (have) {
            /* LightTree.swift:180-187 partition: Swift's partition(by:) two-index exchange. */
            uint32_t lo = it.lo, hi = it.hi, *v = out->order;
            for (;;) {
                int done = 0;
                for (;;) {
                    if (lo == hi) { done = 1; break; }
                    if (bucket(&bounds[v[lo]], best_axis, &cen) > best_bucket) break;
                    ++lo;
                }
                if (done) break;
                for (;;) {
                    --hi;
                    if (lo == hi) { done = 1; break; }
                    if (!(bucket(&bounds[v[hi]], best_axis, &cen) > best_bucket)) break;
                }
                if (done) break;
                uint32_t t = v[lo]; v[lo] = v[hi]; v[hi] = t;
                ++lo;
            }
            if (lo != it.lo && lo != it.hi) mid = lo;
        }
        /* LightTree.swift:192-205 node */
        uint32_t lq[3], hq[3];
        for (int k = 0; k < 3; ++k) {
            lq[k] = quant(total.lo.v[k], g_lo.v[k], g_ext.v[k], 0);
            hq[k] = quant(total.hi.v[k], g_lo.v[k], g_ext.v[k], 1);
        }
        float phi = (float)total.phi;
        uint32_t phib;
        memcpy(&phib, &phi, 4);
        uint32_t *w = out->words + 8 * (size_t)index;
        w[0] = octahedral(total.w);
        w[1] = phib;
        w[2] = quant_cos(total.cos_o) | quant_cos(total.cos_e) << 16;
        w[3] = lq[0] | lq[1] << 16;
        w[4] = lq[2] | hq[0] << 16;
        w[5] = hq[1] | hq[2] << 16;
        w[6] = child;
        w[7] = it.parent | leaf_count << 27;
        if (leaf) continue;
        lwork right = {mid, it.hi, index, 1, it.level + 1}, left = {it.lo, mid, index, 0, it.level + 1};
        work[top++] = right;
        work[top++] = left;
    }
    free(work);
    free(bounds);
    return (int32_t)out->node_count;
}

void vo_light_tree_free(vo_light_tree *t) {
    free(t->words);
    free(t->order);
    free(t->leaf_node);
    memset(t, 0, sizeof *t);
}

/* ------------------------------------------------------------------ walk (float) */
typedef struct { float lo[3], hi[3], w[3], phi, cos_o, cos_e; uint32_t child, parent, count; } ltnode;

static void decode(const vo_light_view *t, uint32_t i, ltnode *n) {
    const uint32_t *a = t->nodes + 8 * (size_t)i, *b = a + 4;
    /* light_tree_axis main.swift:2744-2749 */
    float fx = (float)(a[0] & 0xffffu) / 65535.0f * 2.0f - 1.0f, fy = (float)(a[0] >> 16) / 65535.0f * 2.0f - 1.0f;
    float vx = fx, vy = fy, vz = 1.0f - fabsf(fx) - fabsf(fy);
    if (vz < 0.0f) {
        float nx = (1.0f - fabsf(fy)) * (fx >= 0.0f ? 1.0f : -1.0f), ny = (1.0f - fabsf(fx)) * (fy >= 0.0f ? 1.0f : -1.0f);
        vx = nx;
        vy = ny;
    }
    float l = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
    n->w[0] = vx * l; n->w[1] = vy * l; n->w[2] = vz * l;
    memcpy(&n->phi, &a[1], 4);
    n->cos_o = (float)(a[2] & 0xffffu) / 65535.0f * 2.0f - 1.0f;
    n->cos_e = (float)(a[2] >> 16) / 65535.0f * 2.0f - 1.0f;
    float ql[3] = {(float)(a[3] & 0xffffu), (float)(a[3] >> 16), (float)(b[0] & 0xffffu)};
    float qh[3] = {(float)(b[0] >> 16), (float)(b[1] & 0xffffu), (float)(b[1] >> 16)};
    for (int k = 0; k < 3; ++k) {
        n->lo[k] = t->lo[k] + ql[k] * t->scale[k];
        n->hi[k] = t->lo[k] + qh[k] * t->scale[k];
    }
    n->child = b[2];
    n->parent = b[3] & NO_PARENT;
    n->count = b[3] >> 27;
}
static float leaf_power(const vo_light_view *t, uint32_t position) {
    float p;
    memcpy(&p, &t->leaves[3 * (size_t)position + 1], 4);
    return p;
}
static float leaf_sum(const vo_light_view *t, const ltnode *n) {
    float s = 0.0f;
    for (uint32_t i = 0; i < n->count; ++i) s += leaf_power(t, n->child + i);
    return s;
}
static float cos_sub(float sa, float ca, float sb, float cb) { return ca > cb ? 1.0f : ca * cb + sa * sb; }
static float sin_sub(float sa, float ca, float sb, float cb) { return ca > cb ? 0.0f : sa * cb - ca * sb; }
static float dist2(const float *a, const float *b) {
    float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}
/* light_tree_importance main.swift:2785-2798 */
float vo_light_tree_importance(const vo_light_view *t, uint32_t i, const float *p) {
    ltnode n;
    decode(t, i, &n);
    if (!(n.phi > 0.0f)) return 0.0f;
    float pc[3] = {0.5f * (n.lo[0] + n.hi[0]), 0.5f * (n.lo[1] + n.hi[1]), 0.5f * (n.lo[2] + n.hi[2])};
    float diag[3] = {n.hi[0] - n.lo[0], n.hi[1] - n.lo[1], n.hi[2] - n.lo[2]};
    float dl = sqrtf(diag[0] * diag[0] + diag[1] * diag[1] + diag[2] * diag[2]);
    float d2 = fmaxf(dist2(p, pc), fmaxf(0.5f * dl, 1e-20f));
    if (p[0] >= n.lo[0] && p[1] >= n.lo[1] && p[2] >= n.lo[2] && p[0] <= n.hi[0] && p[1] <= n.hi[1] && p[2] <= n.hi[2])
        return n.phi / d2;
    float inv = 1.0f / sqrtf(dist2(p, pc));
    float wi[3] = {(p[0] - pc[0]) * inv, (p[1] - pc[1]) * inv, (p[2] - pc[2]) * inv};
    float cw = n.w[0] * wi[0] + n.w[1] * wi[1] + n.w[2] * wi[2], sw = sqrtf(fmaxf(0.0f, 1.0f - cw * cw));
    float s2 = dist2(pc, n.hi) / dist2(pc, p);
    float cb = s2 >= 1.0f ? -1.0f : sqrtf(fmaxf(0.0f, 1.0f - s2)), sb = sqrtf(fmaxf(0.0f, 1.0f - cb * cb));
    float so = sqrtf(fmaxf(0.0f, 1.0f - n.cos_o * n.cos_o));
    float cx = cos_sub(sw, cw, so, n.cos_o), sx = sin_sub(sw, cw, so, n.cos_o);
    float cp = cos_sub(sx, cx, sb, cb);
    return cp <= n.cos_e ? 0.0f : n.phi * cp / d2;
}
/* light_tree_sample main.swift:2802-2826 */
uint32_t vo_light_tree_sample(const vo_light_view *t, const float *p, float u, float *pmf) {
    *pmf = 1.0f;
    uint32_t index = 0;
    for (uint32_t depth = 0; depth <= 64; ++depth) {
        ltnode n;
        decode(t, index, &n);
        if (n.count != 0u) {
            float sum = leaf_sum(t, &n), target = u * sum, cum = 0.0f, chosen = 0.0f;
            uint32_t position = 0xffffffffu;
            for (uint32_t i = 0; i < n.count && sum > 0.0f; ++i) {
                float power = leaf_power(t, n.child + i);
                if (!(power > 0.0f)) continue;
                position = n.child + i;
                chosen = power;
                cum += power;
                if (target < cum) break;
            }
            if (position == 0xffffffffu) break;
            *pmf *= chosen / sum;
            return position;
        }
        float l = vo_light_tree_importance(t, index + 1u, p), r = vo_light_tree_importance(t, n.child, p);
        if (!(l + r > 0.0f)) break;
        float pl = l / (l + r);
        if (u < pl) {
            u = fminf(u / pl, 0x1.fffffep-1f);
            index = index + 1u;
            *pmf *= pl;
        } else {
            u = fminf((u - pl) / (1.0f - pl), 0x1.fffffep-1f);
            index = n.child;
            *pmf *= r / (l + r);
        }
    }
    *pmf = 0.0f;
    return 0xffffffffu;
}
/* light_tree_pmf main.swift:2828-2844 */
float vo_light_tree_pmf(const vo_light_view *t, const float *p, uint32_t position) {
    uint32_t index = t->leaves[3 * (size_t)position + 2];
    ltnode node;
    decode(t, index, &node);
    float sum = leaf_sum(t, &node);
    float pmf = sum > 0.0f ? leaf_power(t, position) / sum : 0.0f;
    for (uint32_t depth = 0; depth <= 64 && pmf > 0.0f; ++depth) {
        uint32_t parent = node.parent;
        if (parent == NO_PARENT) return pmf;
        ltnode up;
        decode(t, parent, &up);
        float l = vo_light_tree_importance(t, parent + 1u, p), r = vo_light_tree_importance(t, up.child, p);
        float mine = index == parent + 1u ? l : r;
        if (!(l + r > 0.0f) || !(mine > 0.0f)) return 0.0f;
        pmf *= mine / (l + r);
        index = parent;
        node = up;
    }
    return 0.0f;
}
