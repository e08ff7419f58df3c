// lighttree.cpp -- host builder of the light BVH (lighttree.h).
#include "lighttree.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace vt {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kBuckets = 12;

struct D3 {
    double x, y, z;
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
D3 d3(double x, double y, double z) { return D3{x, y, z}; }
D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
double length(D3 a) { return std::sqrt(dot(a, a)); }
D3 vmin(D3 a, D3 b) { return d3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
D3 vmax(D3 a, D3 b) { return d3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }
double clamp1(double c) { return std::min(1.0, std::max(-1.0, c)); }

const double kInf = std::numeric_limits<double>::infinity();

struct Bounds {
    D3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
    double phi = 0;
    D3 w{0, 0, 1};
    double cos_o = 1, cos_e = 1;
    D3 centroid() const { return (lo + hi) * 0.5; }
};

// pbrt-v4 AngleBetween.
double angle_between(D3 a, D3 b) {
    if (dot(a, b) < 0) return kPi - 2 * std::asin(std::min(1.0, length(a + b) / 2));
    return 2 * std::asin(std::min(1.0, length(b - a) / 2));
}

// Rotation of v by `angle` about the unit axis k (Rodrigues).
D3 rotate(D3 v, D3 k, double angle) {
    double c = std::cos(angle), s = std::sin(angle);
    return v * c + cross(k, v) * s + k * (dot(k, v) * (1 - c));
}

// pbrt-v4 Union(DirectionCone, DirectionCone).
void cone_union(D3 wa, double ca, D3 wb, double cb, D3 &w, double &c) {
    double ta = std::acos(clamp1(ca)), tb = std::acos(clamp1(cb));
    double td = angle_between(wa, wb);
    if (std::min(td + tb, kPi) <= ta) { w = wa; c = ca; return; }
    if (std::min(td + ta, kPi) <= tb) { w = wb; c = cb; return; }
    double to = (ta + td + tb) / 2;
    if (to >= kPi) { w = wa; c = -1; return; }
    D3 axis = cross(wa, wb);
    if (dot(axis, axis) == 0) { w = wa; c = -1; return; }
    D3 r = rotate(wa, axis * (1.0 / length(axis)), to - ta);
    w = r * (1.0 / length(r));
    c = std::cos(to);
}

// pbrt-v4 Union(LightBounds, LightBounds): a powerless side contributes nothing.
Bounds unite(const Bounds &a, const Bounds &b) {
    if (a.phi == 0) return b;
    if (b.phi == 0) return a;
    Bounds r;
    r.lo = vmin(a.lo, b.lo);
    r.hi = vmax(a.hi, b.hi);
    r.phi = a.phi + b.phi;
    cone_union(a.w, a.cos_o, b.w, b.cos_o, r.w, r.cos_o);
    r.cos_e = std::min(a.cos_e, b.cos_e);
    return r;
}

// pbrt-v4 BVHLightSampler::EvaluateCost (SAOH), axis < 0: no Kr penalty.
double cost(const Bounds &b, D3 diag, int axis) {
    if (!(b.phi > 0)) return 0;
    double to = std::acos(clamp1(b.cos_o)), te = std::acos(clamp1(b.cos_e));
    double tw = std::min(to + te, kPi);
    double so = std::sqrt(std::max(0.0, 1 - b.cos_o * b.cos_o));
    double m_omega = 2 * kPi * (1 - b.cos_o) + kPi / 2 * (2 * tw * so - std::cos(to - 2 * tw) - 2 * to * so + b.cos_o);
    double kr = 1;
    if (axis >= 0) kr = std::max(diag.x, std::max(diag.y, diag.z)) / std::max(diag[axis], 1e-300);
    D3 d = b.hi - b.lo;
    double area = 2 * (d.x * d.y + d.x * d.z + d.y * d.z);
    return b.phi * m_omega * kr * area;
}

Bounds emitter_bounds(const LightEmitter &e) {
    D3 a = d3(e.a[0], e.a[1], e.a[2]), b = d3(e.b[0], e.b[1], e.b[2]), c = d3(e.c[0], e.c[1], e.c[2]);
    Bounds r;
    r.lo = vmin(a, vmin(b, c));
    r.hi = vmax(a, vmax(b, c));
    D3 n = cross(b - a, c - a);
    double len = length(n);
    r.phi = len > 0 && std::isfinite(e.power) ? std::max(0.0, e.power) : 0;
    r.w = len > 0 ? n * (1.0 / len) : d3(0, 0, 1);
    r.cos_o = 1;
    r.cos_e = std::min(1.0, std::max(0.0, e.cos_emission));
    return r;
}

struct Centroids {
    D3 lo, hi;
};

int bucket_of(const Bounds &b, int axis, const Centroids &c) {
    double lo = c.lo[axis], extent = c.hi[axis] - lo;
    return std::min(kBuckets - 1, (int)(kBuckets * (b.centroid()[axis] - lo) / extent));
}

struct Split {
    bool ok = false;
    double cost = kInf;
    int axis = 0, bucket = 0;
};

Split best_split(uint32_t lo_i, uint32_t hi_i, const std::vector<Bounds> &bounds, const std::vector<uint32_t> &order,
                 const Bounds &node, const Centroids &cen) {
    D3 diag = node.hi - node.lo;
    Split best;
    for (int axis = 0; axis < 3; ++axis) {
        if (!(cen.hi[axis] > cen.lo[axis])) continue;
        Bounds bb[kBuckets];
        for (uint32_t i = lo_i; i < hi_i; ++i) {
            int k = bucket_of(bounds[order[i]], axis, cen);
            bb[k] = unite(bb[k], bounds[order[i]]);
        }
        Bounds below[kBuckets], above[kBuckets];
        for (int k = 0; k < kBuckets; ++k) below[k] = k == 0 ? bb[0] : unite(below[k - 1], bb[k]);
        for (int k = kBuckets - 1; k >= 0; --k) above[k] = k == kBuckets - 1 ? bb[k] : unite(above[k + 1], bb[k]);
        for (int k = 0; k < kBuckets - 1; ++k) {
            if (!(below[k].phi > 0 || above[k + 1].phi > 0)) continue;
            double c = cost(below[k], diag, axis) + cost(above[k + 1], diag, axis);
            if (c < best.cost) {
                best.ok = true;
                best.cost = c;
                best.axis = axis;
                best.bucket = k;
            }
        }
    }
    return best;
}

// Swift's MutableCollection.partition(by:): elements for which `second` holds move to the end,
// by the two-index exchange the standard library uses; returns the first index of the second part.
template <class F>
uint32_t swift_partition(std::vector<uint32_t> &v, uint32_t lo, uint32_t hi, F second) {
    for (;;) {
        for (;;) {
            if (lo == hi) return lo;
            if (second(v[lo])) break;
            ++lo;
        }
        for (;;) {
            --hi;
            if (lo == hi) return lo;
            if (!second(v[hi])) break;
        }
        std::swap(v[lo], v[hi]);
        ++lo;
    }
}

uint32_t quant(double v, double lo, double extent, bool up) {
    if (!(extent > 0) || !std::isfinite(v)) return up ? 65535u : 0u;
    double x = (v - lo) / extent * 65535;
    return (uint32_t)std::min(65535.0, std::max(0.0, up ? std::ceil(x) : std::floor(x)));
}
uint32_t quant_cos(double c) { return (uint32_t)std::min(65535.0, std::max(0.0, std::floor((c + 1) / 2 * 65535))); }

uint32_t octahedral(D3 v) {
    double s = std::fabs(v.x) + std::fabs(v.y) + std::fabs(v.z);
    double px = v.x / s, py = v.y / s;
    if (v.z < 0) {
        double nx = (1 - std::fabs(py)) * (px >= 0 ? 1 : -1), ny = (1 - std::fabs(px)) * (py >= 0 ? 1 : -1);
        px = nx;
        py = ny;
    }
    auto e = [](double x) { return (uint32_t)std::min(65535.0, std::max(0.0, std::round((x * 0.5 + 0.5) * 65535))); };
    return e(px) | e(py) << 16;
}

uint32_t float_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

}  // namespace

LightTreeBuilt build_light_tree(const std::vector<LightEmitter> &emitters, int max_leaf, double leaf_ratio) {
    LightTreeBuilt out;
    const uint32_t count = (uint32_t)emitters.size();
    std::vector<Bounds> bounds(count);
    D3 glo{kInf, kInf, kInf}, ghi{-kInf, -kInf, -kInf};
    for (uint32_t i = 0; i < count; ++i) {
        bounds[i] = emitter_bounds(emitters[i]);
        glo = vmin(glo, bounds[i].lo);
        ghi = vmax(ghi, bounds[i].hi);
    }
    if (count > 0) {
        for (int k = 0; k < 3; ++k) {
            out.grid_lo[k] = std::nextafter((float)glo[k], -std::numeric_limits<float>::infinity());
            out.grid_hi[k] = std::nextafter((float)ghi[k], std::numeric_limits<float>::infinity());
        }
    }
    const D3 g_lo = d3(out.grid_lo[0], out.grid_lo[1], out.grid_lo[2]);
    const D3 g_ext = d3(out.grid_hi[0], out.grid_hi[1], out.grid_hi[2]) - g_lo;
    out.order.resize(count);
    for (uint32_t i = 0; i < count; ++i) out.order[i] = i;
    out.leaf_node.assign(count, 0);
    struct Work {
        uint32_t lo, hi, parent;
        bool second;
        uint32_t level;
    };
    std::vector<Work> work;
    if (count > 0) work.push_back({0, count, kLightTreeNoParent, false, 0});
    auto emit_node = [&](const Bounds &b, uint32_t child, uint32_t parent, uint32_t leaf_count) {
        uint32_t lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = quant(b.lo[k], g_lo[k], g_ext[k], false);
            hi[k] = quant(b.hi[k], g_lo[k], g_ext[k], true);
        }
        uint32_t w[8] = {octahedral(b.w), float_bits((float)b.phi), quant_cos(b.cos_o) | quant_cos(b.cos_e) << 16,
                         lo[0] | lo[1] << 16, lo[2] | hi[0] << 16, hi[1] | hi[2] << 16, child,
                         parent | leaf_count << 27};
        out.words.insert(out.words.end(), w, w + 8);
    };
    while (!work.empty()) {
        Work it = work.back();
        work.pop_back();
        const uint32_t index = out.node_count++;
        out.depth = std::max(out.depth, it.level);
        if (it.second) out.words[8 * (size_t)it.parent + 6] = index;
        Bounds total;
        Centroids cen{{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}};
        for (uint32_t i = it.lo; i < it.hi; ++i) {
            const Bounds &b = bounds[out.order[i]];
            total = unite(total, b);
            cen.lo = vmin(cen.lo, b.centroid());
            cen.hi = vmax(cen.hi, b.centroid());
        }
        if (total.phi == 0) {
            for (uint32_t i = it.lo; i < it.hi; ++i) {
                total.lo = vmin(total.lo, bounds[out.order[i]].lo);
                total.hi = vmax(total.hi, bounds[out.order[i]].hi);
            }
        }
        const uint32_t n = it.hi - it.lo;
        Split cand;
        if (n > 1) cand = best_split(it.lo, it.hi, bounds, out.order, total, cen);
        const double unsplit = cost(total, total.hi - total.lo, -1);
        const bool leaf = n == 1 || ((int)n <= max_leaf && (total.phi == 0 || (cand.ok ? cand.cost : kInf) >= leaf_ratio * unsplit));
        if (leaf) {
            for (uint32_t i = it.lo; i < it.hi; ++i) out.leaf_node[i] = index;
            emit_node(total, it.lo, it.parent, n);
            continue;
        }
        uint32_t mid = it.lo + n / 2;
        if (cand.ok) {
            const int axis = cand.axis, bk = cand.bucket;
            uint32_t part = swift_partition(out.order, it.lo, it.hi,
                                            [&](uint32_t e) { return bucket_of(bounds[e], axis, cen) > bk; });
            if (part != it.lo && part != it.hi) mid = part;
        }
        emit_node(total, 0, it.parent, 0);
        work.push_back({mid, it.hi, index, true, it.level + 1});
        work.push_back({it.lo, mid, index, false, it.level + 1});
    }
    return out;
}

}  // namespace vt
