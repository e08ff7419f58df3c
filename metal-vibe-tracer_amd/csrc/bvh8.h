// bvh8.h -- the 8-wide BVH the gfx950 traversal kernels read, and its host builder.
//
// Layout in HBM (one scene):
//   nodes:  Node8[node_count], 80 bytes each (five 16-byte words, one dwordx4 load each).  A
//           node stores its eight child boxes quantised to 8 bits per plane on a per-node grid
//           p + q * 2^e (power-of-two scale, so dequantisation fma(q, s, p) rounds once and the
//           host verifies every dequantised box encloses its child exactly).  Internal children
//           of a node are contiguous from child_base, in slot order; the leaves' triangles are
//           contiguous from tri_base.
//   tris:   TriRecord[face_count], 48 bytes each: the three corner positions (exact copies of the
//           uploaded vertex positions) and the document face index, in leaf order.
// Children are assigned to slots per ray octant (slot s holds the child nearest for rays whose
// direction signs are s), so visiting hit slots in the order (i ^ octant), i = 0..7, goes
// roughly near to far without sorting.
#pragma once
#include <cstdint>
#include <vector>

namespace vt {

struct alignas(16) Node8 {
    float px, py, pz;       // grid origin
    uint8_t ex, ey, ez;     // biased float exponents of the per-axis scale 2^(e - 127)
    uint8_t imask;          // bit s: slot s is an internal child
    uint32_t child_base;    // index of the first internal child node
    uint32_t tri_base;      // index of the first triangle record of this node's leaves
    uint8_t meta[8];        // 0 empty; 0x80 internal; else (triangle offset << 2) | count (1..3)
    uint8_t qlo[3][8];      // per axis, per slot: low plane
    uint8_t qhi[3][8];      // per axis, per slot: high plane
};
static_assert(sizeof(Node8) == 80, "Node8 is five 16-byte words");

struct alignas(16) TriRecord {
    float v0[3];
    uint32_t face;          // document face index
    float v1[3];
    uint32_t pad1;
    float v2[3];
    uint32_t pad2;
};
static_assert(sizeof(TriRecord) == 48, "TriRecord is three 16-byte words");

struct BuildOptions {
    int leaf_size = 3;
    int bins = 16;
    int max_depth = 24;
    int collapse = 0;  // 0: SAH-optimal collapse of a binary tree; 1: greedy widening; 2: the flat
                       // binary median hierarchy (MedianBvh) instead of an 8-wide tree
};

struct Bvh8 {
    std::vector<Node8> nodes;
    std::vector<TriRecord> tris;
    int depth = 0;
    float root_max = 0;     // max |coordinate| of the scene bounds
    double build_seconds = 0;
};

// Builds the BVH over `face_count` faces (4 u32 each: a, b, c, tag) of `vertices` (8 floats each).
// Deterministic: the result does not depend on the number of threads.
Bvh8 build_bvh8(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count,
                const BuildOptions &opt);

// The reference's flat binary hierarchy (OBJMesh.hierarchy, Sources/StudioModel.swift:701-760), which
// its software trace_scene walks for a published single-level mesh (main.swift:2056-2080): median
// splits on the longest axis of a node's triangle bounds, at most four triangles per leaf, nodes in
// creation (pre-)order.  MeshNode is the reference's 48-byte record: lo, hi (w unused), links = (left,
// right, split axis, 0) for an interior node, (0, 0, first triangle, count) for a leaf.
struct alignas(16) MeshNode {
    float lo[4], hi[4];
    int32_t links[4];
};
static_assert(sizeof(MeshNode) == 48, "MeshNode is three 16-byte words");

struct MedianBvh {
    std::vector<MeshNode> nodes;
    std::vector<TriRecord> tris;   // in hierarchy order (tris[i] is face order[i])
    std::vector<uint32_t> order;
    int depth = 0;                 // levels (root = 1)
    float root_max = 0;            // max |coordinate| of the root bounds
    double build_seconds = 0;
};

MedianBvh build_median_bvh(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count);

// The same hierarchy over n boxes (6 floats each: lo xyz, hi xyz), e.g. instance bounds.  Leaf slots
// name items instead of holding triangle records: tri_base + offset indexes `items`, which receives
// the item numbers in leaf order (Bvh8::tris stays empty).
Bvh8 build_bvh8_boxes(const float *boxes, uint32_t n, const BuildOptions &opt, std::vector<uint32_t> &items);

}  // namespace vt
