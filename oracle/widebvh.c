/*
 * widebvh.c -- oracle (test infrastructure only): restatement of the reference's 4-wide
 * binned-SAH BVH builder, Sources/Acceleration.swift:106-351 (WideBVH), and of the per-face
 * bounds, Sources/Acceleration.swift:445-458.  Float arithmetic follows the Swift source
 * operation by operation (no contraction: built with -ffp-contract=off).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define BINS 16 /* Acceleration.swift:109 */

void vo_free(void *p) { free(p); }

/* Acceleration.swift:445-458 */
void vo_mesh_bounds(const float *v, const uint32_t *f, uint32_t n, float *b) {
    for (uint32_t i = 0; i < n; ++i) {
        const float *pa = v + 8 * (size_t)f[4 * i], *pb = v + 8 * (size_t)f[4 * i + 1], *pc = v + 8 * (size_t)f[4 * i + 2];
        for (int a = 0; a < 3; ++a) {
            float l = fminf(fminf(pa[a], pb[a]), pc[a]), h = fmaxf(fmaxf(pa[a], pb[a]), pc[a]);
            b[6 * (size_t)i + a] = l;
            b[6 * (size_t)i + 3 + a] = h;
        }
    }
}

typedef struct {
    int32_t start, end;
    float lo[3], hi[3], clo[3], chi[3];
} range_t;

/* Acceleration.swift:130-133 WideBVH.area: half the surface area of max(h - l, 0). */
static float area3(const float *l, const float *h) {
    float ex = fmaxf(h[0] - l[0], 0.0f), ey = fmaxf(h[1] - l[1], 0.0f), ez = fmaxf(h[2] - l[2], 0.0f);
    return ex * ey + ey * ez + ez * ex;
}
static float range_area(const range_t *r) { return area3(r->lo, r->hi); }

static void range_empty(range_t *r, int32_t s, int32_t e) {
    r->start = s; r->end = e;
    for (int a = 0; a < 3; ++a) {
        r->lo[a] = FLT_MAX; r->hi[a] = -FLT_MAX; r->clo[a] = FLT_MAX; r->chi[a] = -FLT_MAX;
    }
}

typedef struct {
    const float *bounds;
    int32_t *order;
    int32_t leaf, median;
    float binLo[3 * BINS][3], binHi[3 * BINS][3], binCLo[3 * BINS][3], binCHi[3 * BINS][3];
    int32_t binCount[3 * BINS];
    float suffix[BINS];
} builder_t;

/* Acceleration.swift:154-163 measure */
static range_t measure(const builder_t *B, int32_t s, int32_t e) {
    range_t r;
    range_empty(&r, s, e);
    for (int32_t k = s; k < e; ++k) {
        const float *b = B->bounds + 6 * (size_t)B->order[k];
        for (int a = 0; a < 3; ++a) {
            float l = b[a], h = b[3 + a], c = (l + h) * 0.5f;
            r.lo[a] = fminf(r.lo[a], l); r.hi[a] = fmaxf(r.hi[a], h);
            r.clo[a] = fminf(r.clo[a], c); r.chi[a] = fmaxf(r.chi[a], c);
        }
    }
    return r;
}

static __thread const float *t_sort_bounds;
static __thread int t_sort_axis;
static int median_cmp(const void *x, const void *y) {
    int32_t i = *(const int32_t *)x, j = *(const int32_t *)y;
    const float *b = t_sort_bounds;
    float cx = b[6 * (size_t)i + t_sort_axis] + b[6 * (size_t)i + 3 + t_sort_axis];
    float cy = b[6 * (size_t)j + t_sort_axis] + b[6 * (size_t)j + 3 + t_sort_axis];
    if (cx == cy) return i < j ? -1 : (i > j ? 1 : 0);
    return cx < cy ? -1 : 1;
}

static int32_t bin_of(float c, float origin, float s) {
    /* Swift Int(Float) truncates toward zero; then clamped to [0, bins-1]. */
    float f = (c - origin) * s;
    long long v = (long long)f;
    if (v < 0) v = 0;
    if (v > BINS - 1) v = BINS - 1;
    return (int32_t)v;
}

/* Acceleration.swift:165-244 split */
static void split(builder_t *B, const range_t *r, range_t *left, range_t *right) {
    float extent[3] = {r->chi[0] - r->clo[0], r->chi[1] - r->clo[1], r->chi[2] - r->clo[2]};
    float emax = fmaxf(fmaxf(extent[0], extent[1]), extent[2]);
    int32_t count = r->end - r->start;
    if (!B->median && emax > 0) {
        float scale[3];
        for (int a = 0; a < 3; ++a) {
            float s = (float)BINS / extent[a];
            scale[a] = (extent[a] > 0 && isfinite(s)) ? s : 0.0f;
        }
        for (int k = 0; k < 3 * BINS; ++k) {
            for (int a = 0; a < 3; ++a) {
                B->binLo[k][a] = FLT_MAX; B->binHi[k][a] = -FLT_MAX;
                B->binCLo[k][a] = FLT_MAX; B->binCHi[k][a] = -FLT_MAX;
            }
            B->binCount[k] = 0;
        }
        for (int32_t k = r->start; k < r->end; ++k) {
            const float *b = B->bounds + 6 * (size_t)B->order[k];
            float l[3] = {b[0], b[1], b[2]}, h[3] = {b[3], b[4], b[5]}, c[3];
            int slot[3];
            for (int a = 0; a < 3; ++a) c[a] = (l[a] + h[a]) * 0.5f;
            for (int a = 0; a < 3; ++a) slot[a] = a * BINS + bin_of(c[a], r->clo[a], scale[a]);
            for (int a = 0; a < 3; ++a) {
                int s = slot[a];
                B->binCount[s] += 1;
                for (int q = 0; q < 3; ++q) {
                    B->binLo[s][q] = fminf(B->binLo[s][q], l[q]); B->binHi[s][q] = fmaxf(B->binHi[s][q], h[q]);
                    B->binCLo[s][q] = fminf(B->binCLo[s][q], c[q]); B->binCHi[s][q] = fmaxf(B->binCHi[s][q], c[q]);
                }
            }
        }
        float bestCost = FLT_MAX;
        int best = -1, bin = 0;
        for (int a = 0; a < 3; ++a) {
            if (!(scale[a] > 0)) continue;
            float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            int32_t cnt = 0;
            for (int b = BINS - 1; b > 0; --b) {
                int k = a * BINS + b;
                if (B->binCount[k] > 0)
                    for (int q = 0; q < 3; ++q) { l[q] = fminf(l[q], B->binLo[k][q]); h[q] = fmaxf(h[q], B->binHi[k][q]); }
                cnt += B->binCount[k];
                B->suffix[b] = cnt > 0 ? area3(l, h) * (float)cnt : 0.0f;
            }
            for (int q = 0; q < 3; ++q) { l[q] = FLT_MAX; h[q] = -FLT_MAX; }
            cnt = 0;
            for (int b = 0; b < BINS - 1; ++b) {
                int k = a * BINS + b;
                if (B->binCount[k] > 0)
                    for (int q = 0; q < 3; ++q) { l[q] = fminf(l[q], B->binLo[k][q]); h[q] = fmaxf(h[q], B->binHi[k][q]); }
                cnt += B->binCount[k];
                if (cnt == 0 || cnt == count) continue;
                float cost = area3(l, h) * (float)cnt + B->suffix[b + 1];
                if (cost < bestCost) { bestCost = cost; best = a; bin = b; }
            }
        }
        if (best >= 0) {
            float s = scale[best], origin = r->clo[best];
            int32_t i = r->start, j = r->end - 1;
            while (i <= j) {
                int32_t t = B->order[i];
                float c = (B->bounds[6 * (size_t)t + best] + B->bounds[6 * (size_t)t + 3 + best]) * 0.5f;
                if (bin_of(c, origin, s) <= bin) i += 1;
                else { int32_t x = B->order[i]; B->order[i] = B->order[j]; B->order[j] = x; j -= 1; }
            }
            if (i > r->start && i < r->end) {
                range_empty(left, r->start, i);
                range_empty(right, i, r->end);
                for (int b = 0; b < BINS; ++b) {
                    int k = best * BINS + b;
                    if (B->binCount[k] <= 0) continue;
                    range_t *d = b <= bin ? left : right;
                    for (int q = 0; q < 3; ++q) {
                        d->lo[q] = fminf(d->lo[q], B->binLo[k][q]); d->hi[q] = fmaxf(d->hi[q], B->binHi[k][q]);
                        d->clo[q] = fminf(d->clo[q], B->binCLo[k][q]); d->chi[q] = fmaxf(d->chi[q], B->binCHi[k][q]);
                    }
                }
                return;
            }
        }
    }
    /* Acceleration.swift:232-243: object median on the widest centroid axis. */
    int32_t middle = (r->start + r->end) / 2;
    if (emax > 0) {
        int a = extent[0] > extent[1] ? (extent[0] > extent[2] ? 0 : 2) : (extent[1] > extent[2] ? 1 : 2);
        t_sort_bounds = B->bounds;
        t_sort_axis = a;
        qsort(B->order + r->start, (size_t)count, sizeof(int32_t), median_cmp);
    }
    *left = measure(B, r->start, middle);
    *right = measure(B, middle, r->end);
}

typedef struct { vo_wide_node *v; int32_t n, cap; } nodevec;
static void nv_push(nodevec *a, const vo_wide_node *x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = (vo_wide_node *)realloc(a->v, (size_t)a->cap * sizeof(vo_wide_node));
    }
    a->v[a->n++] = *x;
}

typedef struct { range_t r; int32_t parent, slot, level; } work_t;
typedef struct { work_t *v; int32_t n, cap; } workvec;
static void wv_push(workvec *a, const work_t *x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = (work_t *)realloc(a->v, (size_t)a->cap * sizeof(work_t));
    }
    a->v[a->n++] = *x;
}

/* Acceleration.swift:247-284 run */
static int32_t run(builder_t *B, const range_t *root, int32_t depth, nodevec *nodes, int32_t deferThreshold,
                   workvec *deferred) {
    int32_t depthReached = depth;
    workvec work = {0};
    work_t w0 = {*root, -1, 0, depth};
    wv_push(&work, &w0);
    while (work.n > 0) {
        work_t it = work.v[--work.n];
        if (it.level > depthReached) depthReached = it.level;
        range_t children[4];
        int nc = 1;
        children[0] = it.r;
        while (nc < 4) {
            int best = -1;
            for (int k = 0; k < nc; ++k) {
                int32_t ck = children[k].end - children[k].start;
                if (ck <= B->leaf) continue;
                if (best < 0) { best = k; continue; }
                if (B->median ? ck > children[best].end - children[best].start
                              : range_area(&children[k]) > range_area(&children[best]))
                    best = k;
            }
            if (best < 0) break;
            range_t a, b;
            split(B, &children[best], &a, &b);
            for (int k = nc; k > best + 1; --k) children[k] = children[k - 1];
            children[best] = a;
            children[best + 1] = b;
            nc += 1;
        }
        vo_wide_node node;
        memset(&node, 0, sizeof node);
        int32_t index = nodes->n;
        if (it.parent >= 0) nodes->v[it.parent].child[it.slot] = index;
        for (int k = 0; k < nc; ++k) {
            const range_t *c = &children[k];
            node.lox[k] = c->lo[0]; node.loy[k] = c->lo[1]; node.loz[k] = c->lo[2];
            node.hix[k] = c->hi[0]; node.hiy[k] = c->hi[1]; node.hiz[k] = c->hi[2];
            if (c->end - c->start <= B->leaf) {
                node.child[k] = -(c->start + 1);
                node.count[k] = c->end - c->start;
            }
        }
        nv_push(nodes, &node);
        for (int k = nc - 1; k >= 0; --k) {
            int32_t ck = children[k].end - children[k].start;
            if (ck <= B->leaf) continue;
            work_t w = {children[k], index, k, it.level + 1};
            if (ck < deferThreshold) wv_push(deferred, &w);
            else wv_push(&work, &w);
        }
    }
    free(work.v);
    return depthReached;
}

static int32_t build_once(const float *bounds, int32_t n, int32_t leaf, int median, vo_wide_node **out,
                          int32_t *order, int32_t *depth_out) {
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    builder_t *B = (builder_t *)calloc(1, sizeof(builder_t));
    B->bounds = bounds; B->order = order; B->leaf = leaf; B->median = median;
    nodevec nodes = {0};
    workvec deferred = {0};
    int32_t threshold = n >= 65536 ? n / 16 : 0;
    range_t root = measure(B, 0, n);
    int32_t depth = run(B, &root, 1, &nodes, threshold, &deferred);
    if (deferred.n > 0) {
        nodevec *parts = (nodevec *)calloc((size_t)deferred.n, sizeof(nodevec));
        int32_t *depths = (int32_t *)calloc((size_t)deferred.n, sizeof(int32_t));
        /* Each task partitions only its own item range (Acceleration.swift:322-333). */
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t t = 0; t < deferred.n; ++t) {
            builder_t *L = (builder_t *)calloc(1, sizeof(builder_t));
            L->bounds = bounds; L->order = order; L->leaf = leaf; L->median = median;
            workvec none = {0};
            depths[t] = run(L, &deferred.v[t].r, deferred.v[t].level, &parts[t], 0, &none);
            free(none.v);
            free(L);
        }
        /* Acceleration.swift:336-347: append each subtree, rebasing inner links. */
        for (int32_t t = 0; t < deferred.n; ++t) {
            int32_t offset = nodes.n;
            nodes.v[deferred.v[t].parent].child[deferred.v[t].slot] = offset;
            for (int32_t q = 0; q < parts[t].n; ++q) {
                vo_wide_node nd = parts[t].v[q];
                for (int k = 0; k < 4; ++k) if (nd.child[k] > 0) nd.child[k] += offset;
                nv_push(&nodes, &nd);
            }
            free(parts[t].v);
            if (depths[t] > depth) depth = depths[t];
        }
        free(parts);
        free(depths);
    }
    free(deferred.v);
    free(B);
    *out = nodes.v;
    *depth_out = depth;
    return nodes.n;
}

/* Acceleration.swift:296-300 build(bounds:count:leafSize:depth:) */
int32_t vo_widebvh_build(const float *bounds, int32_t n, int32_t leaf_size, int32_t max_depth,
                         vo_wide_node **nodes_out, int32_t *order_out, int32_t *depth) {
    *nodes_out = NULL;
    *depth = 0;
    if (n <= 0) return 0;
    int32_t count = build_once(bounds, n, leaf_size, 0, nodes_out, order_out, depth);
    if (*depth > max_depth) {
        free(*nodes_out);
        count = build_once(bounds, n, leaf_size, 1, nodes_out, order_out, depth);
    }
    return count;
}
