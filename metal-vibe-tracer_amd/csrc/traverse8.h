// traverse8.h -- per-ray traversal of the 8-wide compressed BVH (layout in bvh8.h), shared by
// every kernel that traces (closest-hit, any-hit, and the integrator's extend/shadow stages).
//
// The node-group stack is abstracted (Stack::push/pop/size) so a kernel keeps it in LDS.
// Arithmetic follows the reference exactly where results depend on it:
//   * boxes: conservative slab test of wide_node_hit (main.swift:1766-1779) on the exactly
//     dequantised child boxes, grown by e = 1.1920929e-6 (|o|max + root max) (main.swift:1792);
//   * triangles: watertight Woop test (main.swift:1537-1582), products rounded separately,
//     correctly rounded divisions, FMA tie-break of equal rounded edge products.
// Among hits at exactly equal t the smaller document face index wins.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VT_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#include <cstring>
#define VT_HD inline
#endif

namespace vt {

struct U4 {
    uint32_t x, y, z, w;
};
struct F4 {
    float x, y, z, w;
};

VT_HD float bits_to_float(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
#endif
}
VT_HD uint32_t float_to_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
#endif
}

struct WoopRay {
    int kx, ky, kz;
    float sx, sy, sz;
};

VT_HD float sel3(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

// main.swift:1537-1550 watertight_ray.
VT_HD WoopRay woop_ray(float dx, float dy, float dz) {
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    float dkz = sel3(dx, dy, dz, kz);
    if (dkz < 0.0f) {
        int s = kx;
        kx = ky;
        ky = s;
    }
    WoopRay w;
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    w.sx = sel3(dx, dy, dz, kx) / dkz;
    w.sy = sel3(dx, dy, dz, ky) / dkz;
    w.sz = 1.0f / dkz;
    return w;
}

// main.swift:1557-1562 watertight_edge.
VT_HD float watertight_edge(float px, float py, float qx, float qy) {
#pragma clang fp contract(off)
    float x = px * qy, y = py * qx, e = x - y;
    if (e == 0.0f) e = __builtin_fmaf(px, qy, -x) - __builtin_fmaf(py, qx, -y);
    return e;
}

// main.swift:1564-1578 without the [tmin, tmax) check (the caller applies it).
VT_HD bool woop_triangle(const WoopRay &w, float ox, float oy, float oz, F4 p0, F4 p1, F4 p2, float &t, float &b1,
                         float &b2) {
#pragma clang fp contract(off)
    float Ax = p0.x - ox, Ay = p0.y - oy, Az = p0.z - oz;
    float Bx = p1.x - ox, By = p1.y - oy, Bz = p1.z - oz;
    float Cx = p2.x - ox, Cy = p2.y - oy, Cz = p2.z - oz;
    float akz = sel3(Ax, Ay, Az, w.kz), bkz = sel3(Bx, By, Bz, w.kz), ckz = sel3(Cx, Cy, Cz, w.kz);
    float ax = sel3(Ax, Ay, Az, w.kx) - w.sx * akz, ay = sel3(Ax, Ay, Az, w.ky) - w.sy * akz;
    float bx = sel3(Bx, By, Bz, w.kx) - w.sx * bkz, by = sel3(Bx, By, Bz, w.ky) - w.sy * bkz;
    float cx = sel3(Cx, Cy, Cz, w.kx) - w.sx * ckz, cy = sel3(Cx, Cy, Cz, w.ky) - w.sy * ckz;
    float U = watertight_edge(cx, cy, bx, by);
    float V = watertight_edge(ax, ay, cx, cy);
    float W = watertight_edge(bx, by, ax, ay);
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (w.sz * akz) + V * (w.sz * bkz) + W * (w.sz * ckz);
    float inv = 1.0f / det;
    t = T * inv;
    b1 = V * inv;
    b2 = W * inv;
    return true;
}

VT_HD uint32_t byte_of(uint32_t lo4, uint32_t hi4, int s) { return ((s < 4 ? lo4 : hi4) >> (8 * (s & 3))) & 0xffu; }

struct TraceRay {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
};

struct TraceResult {
    float t, b1, b2;
    int32_t face;    // -1: miss
    bool occluded;   // any-hit result
    bool overflow;   // the node-group stack filled: the ray ended before its traversal did
    uint32_t nodes;  // node records fetched
    uint32_t tris;   // triangle records tested
};

// Per-ray traversal state, so a ray can be advanced one node visit at a time (the persistent
// kernels refill a lane with a new ray as soon as its ray finishes).
struct TravState {
    WoopRay w;
    float ox, oy, oz, tmin, tmax;
    float e, invx, invy, invz;  // e: the reference's box enlargement for this ray
    uint32_t oct;  // direction sign bits (x 1, y 2, z 4); the closest hit so far is res.t
    // Current node group: first child index and (imask << 8 | visit-order bits still to visit).
    uint32_t cur_base, cur_bits, node;
    TraceResult res;
};

// Starts a ray with box enlargement e (trav_init: the reference's e for a one-level hierarchy).
VT_HD void trav_init_e(TravState &st, const TraceRay &r, float e) {
    st.res.t = r.tmax;
    st.res.b1 = st.res.b2 = 0.0f;
    st.res.face = -1;
    st.res.occluded = false;
    st.res.overflow = false;
    st.res.nodes = st.res.tris = 0;
    st.w = woop_ray(r.dx, r.dy, r.dz);
    st.ox = r.ox;
    st.oy = r.oy;
    st.oz = r.oz;
    st.tmin = r.tmin;
    st.tmax = r.tmax;
    st.e = e;
    // A component below 1e-8 in magnitude (the reference's parallel threshold) is taken as +-1e-8 with
    // its sign: the slab of a parallel axis then spans about -+1e8 x the origin's distance to its
    // planes -- the whole ray when the origin lies inside the (e-enlarged) slab, nothing reachable
    // when it lies at least e outside -- so no separate parallel test is needed.
    const uint32_t sx = float_to_bits(r.dx) >> 31, sy = float_to_bits(r.dy) >> 31, sz = float_to_bits(r.dz) >> 31;
    st.invx = 1.0f / (fabsf(r.dx) < 1e-8f ? (sx ? -1e-8f : 1e-8f) : r.dx);
    st.invy = 1.0f / (fabsf(r.dy) < 1e-8f ? (sy ? -1e-8f : 1e-8f) : r.dy);
    st.invz = 1.0f / (fabsf(r.dz) < 1e-8f ? (sz ? -1e-8f : 1e-8f) : r.dz);
    st.oct = sx | (sy << 1) | (sz << 2);
    st.cur_base = 0;
    st.cur_bits = 0;
    st.node = 0;
}

VT_HD float max_abs3(float x, float y, float z) { return fmaxf(fabsf(x), fmaxf(fabsf(y), fabsf(z))); }

// main.swift:1790-1792: e = 1.1920929e-6 (|origin|max + root max).
VT_HD void trav_init(TravState &st, const TraceRay &r, float root_max) {
    trav_init_e(st, r, 1.1920929e-6f * (max_abs3(r.ox, r.oy, r.oz) + root_max));
}

// The object-space ray of an instance (main.swift:1634-1637 instance_ray): origin and direction by
// the rows of the world-to-object matrix, the direction not renormalised, so t keeps its meaning.
// Each dot product is evaluated as ((r0 x + r1 y) + r2 z) (+ r3), every operation rounded.
VT_HD TraceRay instance_ray(const float (&local)[3][4], const TraceRay &r, float tmax) {
#pragma clang fp contract(off)
    TraceRay l;
    l.ox = ((local[0][0] * r.ox + local[0][1] * r.oy) + local[0][2] * r.oz) + local[0][3];
    l.oy = ((local[1][0] * r.ox + local[1][1] * r.oy) + local[1][2] * r.oz) + local[1][3];
    l.oz = ((local[2][0] * r.ox + local[2][1] * r.oy) + local[2][2] * r.oz) + local[2][3];
    l.dx = (local[0][0] * r.dx + local[0][1] * r.dy) + local[0][2] * r.dz;
    l.dy = (local[1][0] * r.dx + local[1][1] * r.dy) + local[1][2] * r.dz;
    l.dz = (local[2][0] * r.dx + local[2][1] * r.dy) + local[2][2] * r.dz;
    l.tmin = r.tmin;
    l.tmax = tmax;
    return l;
}

// Node half of a step: tests st.node's 8 child boxes, pushes the current group if inner children were
// hit, and moves st.node to the next node near to far.  Returns the triangles of the leaf children hit
// (trimask: bits relative to tri_base, at most 24) and true when no node is left to visit (the
// triangles returned must still be tested).  The next node does not depend on those triangles, so
// testing them after the move gives the same result as testing them first.
template <bool STATS, class Stack>
VT_HD bool trav_nodes(TravState &st, const U4 *__restrict__ nodes, Stack &stack, uint32_t &trimask,
                      uint32_t &tri_base) {
    const U4 *np = nodes + 5 * (size_t)st.node;
    const U4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3], w4 = np[4];
    if (STATS) ++st.res.nodes;
    const float px = bits_to_float(w0.x), py = bits_to_float(w0.y), pz = bits_to_float(w0.z);
    const float sx = bits_to_float((w0.w & 0xffu) << 23);
    const float sy = bits_to_float(((w0.w >> 8) & 0xffu) << 23);
    const float sz = bits_to_float(((w0.w >> 16) & 0xffu) << 23);
    const uint32_t imask = w0.w >> 24;
    // Slab distances of plane q of an axis: ((q s + p) - o -+ e) / d = q (s / d) + (p - o -+ e) / d, one
    // fma per plane.  The near plane of an axis is its low plane for a positive direction, its high
    // plane otherwise, so the byte words are chosen once per node and no per-child min / max is needed.
    // The boxes are the reference's e-enlarged ones; the rounding differs from the reference's
    // (plane - o -+ e) / d by a few ulps of terms no larger than the scene, far inside the enlargement,
    // so no box the reference enters is culled.
    const bool nx = (st.oct & 1u) != 0u, ny = (st.oct & 2u) != 0u, nz = (st.oct & 4u) != 0u;
    const float bx = sx * st.invx, by = sy * st.invy, bz = sz * st.invz;
    const float dx = px - st.ox, dy = py - st.oy, dz = pz - st.oz;
    const float alx = (dx - st.e) * st.invx, ahx = (dx + st.e) * st.invx;
    const float aly = (dy - st.e) * st.invy, ahy = (dy + st.e) * st.invy;
    const float alz = (dz - st.e) * st.invz, ahz = (dz + st.e) * st.invz;
    const float anx = nx ? ahx : alx, afx = nx ? alx : ahx;
    const float any_ = ny ? ahy : aly, afy = ny ? aly : ahy;
    const float anz = nz ? ahz : alz, afz = nz ? alz : ahz;
    const uint32_t nxa = nx ? w3.z : w2.x, nxb = nx ? w3.w : w2.y, fxa = nx ? w2.x : w3.z, fxb = nx ? w2.y : w3.w;
    const uint32_t nya = ny ? w4.x : w2.z, nyb = ny ? w4.y : w2.w, fya = ny ? w2.z : w4.x, fyb = ny ? w2.w : w4.y;
    const uint32_t nza = nz ? w4.z : w3.x, nzb = nz ? w4.w : w3.y, fza = nz ? w3.x : w4.z, fzb = nz ? w3.y : w4.w;
    uint32_t hitbits = 0;
    trimask = 0;
    tri_base = w1.y;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const uint32_t meta = byte_of(w1.z, w1.w, s);
        const float tnx = __builtin_fmaf((float)byte_of(nxa, nxb, s), bx, anx);
        const float tny = __builtin_fmaf((float)byte_of(nya, nyb, s), by, any_);
        const float tnz = __builtin_fmaf((float)byte_of(nza, nzb, s), bz, anz);
        const float tfx = __builtin_fmaf((float)byte_of(fxa, fxb, s), bx, afx);
        const float tfy = __builtin_fmaf((float)byte_of(fya, fyb, s), by, afy);
        const float tfz = __builtin_fmaf((float)byte_of(fza, fzb, s), bz, afz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, st.tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, st.res.t));
        if (tn <= tf && meta != 0u) {
            if (meta & 0x80u) hitbits |= 1u << (s ^ st.oct);
            else trimask |= ((1u << (meta & 3u)) - 1u) << (meta >> 2);
        }
    }
    if (hitbits) {
        if (st.cur_bits & 0xffu) {
            // The launcher sizes the stack to the tree's depth, so a full stack never occurs; if it did,
            // the ray stops (never an out-of-bounds access) and the kernel reports it.
            if (stack.full()) {
                st.res.overflow = true;
                trimask = 0;
                return true;
            }
            stack.push(st.cur_base, st.cur_bits);
        }
        st.cur_base = w1.x;
        st.cur_bits = (imask << 8) | hitbits;
    }
    if ((st.cur_bits & 0xffu) == 0u) {
        if (stack.empty()) return true;
        stack.pop(st.cur_base, st.cur_bits);
    }
    const uint32_t v = __builtin_ctz(st.cur_bits & 0xffu);
    st.cur_bits &= ~(1u << v);
    const uint32_t slot = v ^ st.oct;
    st.node = st.cur_base + __builtin_popcount((st.cur_bits >> 8) & ((1u << slot) - 1u));
    return false;
}

// A hit of the ray at t on face: the closest-hit rule (nearest t in [tmin, current t), equal t going
// to the smaller face index).  For the any-hit query any t in [tmin, tmax) ends the ray.
template <bool ANY>
VT_HD bool trav_accept(const TravState &st, float t, int32_t face) {
    if (ANY) return t >= st.tmin && t < st.tmax;
    return t >= st.tmin && (t < st.res.t || (t == st.res.t && st.res.face >= 0 && face < st.res.face));
}

// Triangle half of a step, one lane alone: tests the triangles trav_nodes returned.  Returns true
// when an any-hit query found a hit.
template <bool ANY, bool STATS>
VT_HD bool trav_tris(TravState &st, const F4 *__restrict__ tris, uint32_t trimask, uint32_t tri_base) {
    while (trimask) {
        const uint32_t b = __builtin_ctz(trimask);
        trimask &= trimask - 1u;
        const F4 *tp = tris + 3 * (size_t)(tri_base + b);
        const F4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
        if (STATS) ++st.res.tris;
        float t, b1, b2;
        if (!woop_triangle(st.w, st.ox, st.oy, st.oz, p0, p1, p2, t, b1, b2)) continue;
        const int32_t face = (int32_t)float_to_bits(p0.w);
        if (!trav_accept<ANY>(st, t, face)) continue;
        if (ANY) {
            st.res.occluded = true;
            return true;
        }
        st.res.t = t;
        st.res.b1 = b1;
        st.res.b2 = b2;
        st.res.face = face;
    }
    return false;
}

// One step of one ray: a node visit and the triangles of the leaves it hits.  Returns true when the
// ray is finished (result in st.res).
template <bool ANY, bool STATS, class Stack>
VT_HD bool trav_step(TravState &st, const U4 *__restrict__ nodes, const F4 *__restrict__ tris, Stack &stack) {
    uint32_t trimask, tri_base;
    const bool done = trav_nodes<STATS>(st, nodes, stack, trimask, tri_base);
    if (trav_tris<ANY, STATS>(st, tris, trimask, tri_base)) return true;
    return done;
}

// Traverses one ray.  nodes: Node8 records as 5 U4 each; tris: TriRecord as 3 F4 each.
template <bool ANY, bool STATS, class Stack>
VT_HD TraceResult traverse(const U4 *__restrict__ nodes, const F4 *__restrict__ tris, float root_max,
                           const TraceRay &r, Stack &stack) {
    TravState st;
    trav_init(st, r, root_max);
    while (!trav_step<ANY, STATS>(st, nodes, tris, stack)) {
    }
    return st.res;
}

}  // namespace vt
