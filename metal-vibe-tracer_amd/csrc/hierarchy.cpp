// The reference's flat binary hierarchy (OBJMesh.hierarchy, Sources/StudioModel.swift:701-760), built
// on the host as the reference builds it, so the node records and the triangle order are the same:
// triangle centres (a + b + c) / 3 in float; a node's bounds are the min / max of its triangles'
// corners; a node of more than four triangles splits at the middle position on the axis of its largest
// extent, by a quickselect of triangle references (Hoare partition around the middle element's centre,
// ties broken by triangle index, at most 64 partition rounds before the remaining range is sorted);
// children are created left then right, so nodes are in pre-order.
#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "bvh8.h"

namespace vt {
namespace {

struct Builder {
    const float *vertices;
    const uint32_t *faces;
    std::vector<float> centres;  // 3 per triangle
    std::vector<uint32_t> order;
    std::vector<MeshNode> nodes;
    int depth = 0;

    const float *corner(uint32_t tri, int k) const { return vertices + 8 * (size_t)faces[4 * (size_t)tri + k]; }

    bool less(uint32_t a, uint32_t b, int axis) const {
        const float ca = centres[3 * (size_t)a + axis], cb = centres[3 * (size_t)b + axis];
        return ca == cb ? a < b : ca < cb;
    }

    // Moves the references of [start, end) so that position `middle` holds the element a sort along
    // `axis` would put there, smaller ones before it and larger ones after.
    void partition(int64_t start, int64_t end, int64_t middle, int axis) {
        int64_t low = start, high = end - 1;
        int rounds = 0;
        auto lt = [&](uint32_t a, uint32_t b) { return less(a, b, axis); };
        while (low < high) {
            if (rounds >= 64) {  // bounded selection time on adversarial geometry
                std::sort(order.begin() + low, order.begin() + high + 1, lt);
                return;
            }
            ++rounds;
            const uint32_t pivot = order[(size_t)((low + high) / 2)];
            int64_t i = low, j = high;
            while (i <= j) {
                while (i <= high && lt(order[(size_t)i], pivot)) ++i;
                while (j >= low && lt(pivot, order[(size_t)j])) --j;
                if (i <= j) {
                    std::swap(order[(size_t)i], order[(size_t)j]);
                    ++i;
                    --j;
                }
            }
            if (middle <= j) high = j;
            else if (middle >= i) low = i;
            else return;
        }
    }

    int32_t node(int64_t start, int64_t end, int level) {
        depth = std::max(depth, level);
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int64_t i = start; i < end; ++i)
            for (int k = 0; k < 3; ++k) {
                const float *p = corner(order[(size_t)i], k);
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::fmin(lo[a], p[a]);
                    hi[a] = std::fmax(hi[a], p[a]);
                }
            }
        const int32_t index = (int32_t)nodes.size();
        MeshNode n{};
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = lo[a];
            n.hi[a] = hi[a];
        }
        nodes.push_back(n);
        if (end - start <= 4) {
            nodes[(size_t)index].links[2] = (int32_t)start;
            nodes[(size_t)index].links[3] = (int32_t)(end - start);
            return index;
        }
        const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        const int axis = ex > ey ? (ex > ez ? 0 : 2) : (ey > ez ? 1 : 2);
        const int64_t middle = (start + end) / 2;
        partition(start, end, middle, axis);
        const int32_t left = node(start, middle, level + 1), right = node(middle, end, level + 1);
        nodes[(size_t)index].links[0] = left;
        nodes[(size_t)index].links[1] = right;
        nodes[(size_t)index].links[2] = axis;
        return index;
    }
};

}  // namespace

MedianBvh build_median_bvh(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count) {
    const auto t0 = std::chrono::steady_clock::now();
    MedianBvh out;
    if (face_count == 0) return out;
    for (uint32_t f = 0; f < face_count; ++f)
        for (int k = 0; k < 3; ++k)
            if (faces[4 * (size_t)f + k] >= vertex_count) throw std::invalid_argument("face references a missing vertex");
    Builder b{vertices, faces, {}, {}, {}, 0};
    b.centres.resize(3 * (size_t)face_count);
    for (uint32_t f = 0; f < face_count; ++f) {
        const float *a = b.corner(f, 0), *p = b.corner(f, 1), *c = b.corner(f, 2);
        for (int k = 0; k < 3; ++k) b.centres[3 * (size_t)f + k] = ((a[k] + p[k]) + c[k]) / 3.0f;
    }
    b.order.resize(face_count);
    for (uint32_t f = 0; f < face_count; ++f) b.order[f] = f;
    b.nodes.reserve(face_count);
    b.node(0, face_count, 1);
    out.nodes = std::move(b.nodes);
    out.order = std::move(b.order);
    out.depth = b.depth;
    const MeshNode &root = out.nodes[0];
    for (int a = 0; a < 3; ++a) out.root_max = std::max(out.root_max, std::max(std::fabs(root.lo[a]), std::fabs(root.hi[a])));
    out.tris.resize(face_count);
    for (uint32_t i = 0; i < face_count; ++i) {
        TriRecord &t = out.tris[i];
        const uint32_t f = out.order[i];
        const float *a = vertices + 8 * (size_t)faces[4 * (size_t)f], *p = vertices + 8 * (size_t)faces[4 * (size_t)f + 1],
                    *c = vertices + 8 * (size_t)faces[4 * (size_t)f + 2];
        for (int k = 0; k < 3; ++k) {
            t.v0[k] = a[k];
            t.v1[k] = p[k];
            t.v2[k] = c[k];
        }
        t.face = f;
        t.pad1 = t.pad2 = 0;
    }
    out.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace vt
