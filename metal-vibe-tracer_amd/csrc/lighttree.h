// lighttree.h -- host builder of the light BVH over emitter triangles.
//
// Method: Conty and Kulla 2018 as pbrt-v4's BVHLightSampler implements it (LightBounds union with
// normal cones, SAOH split cost with 12 buckets per axis), with the reference's leaf rule and compact
// 32-byte node (Sources/LightTree.swift:1-224): a range of at most 8 emitters stays one leaf when its
// best split costs at least 0.75 of the unsplit SAOH; nodes are depth-first (a node's first child
// follows it); bounds are quantised outward to 16 bits in a grid box, cone cosines rounded down.
#pragma once
#include <cstdint>
#include <vector>

namespace vt {

struct LightEmitter {
    float a[3], b[3], c[3];
    double power;        // area x selection weight
    double cos_emission; // 0 for a one-sided Lambertian emitter
};

struct LightTreeBuilt {
    std::vector<uint32_t> words;     // 8 per node
    uint32_t node_count = 0;
    std::vector<uint32_t> order;     // leaf position -> input emitter
    std::vector<uint32_t> leaf_node; // leaf position -> node
    uint32_t depth = 0;
    float grid_lo[3] = {0, 0, 0}, grid_hi[3] = {0, 0, 0};
};

constexpr uint32_t kLightTreeMaxDepth = 64;
constexpr uint32_t kLightTreeMaxNodes = (1u << 27) - 1;
constexpr uint32_t kLightTreeNoParent = kLightTreeMaxNodes;

LightTreeBuilt build_light_tree(const std::vector<LightEmitter> &emitters, int max_leaf = 8, double leaf_ratio = 0.75);

}  // namespace vt
