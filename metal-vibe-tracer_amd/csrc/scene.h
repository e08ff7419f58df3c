// scene.h -- the published scene (vt_scene) and the context, shared by capi.cpp and renderer.cpp.
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "render_kernels.h"
#include "spectral.h"
#include "trace_kernels.h"
#include "vt.h"

namespace vt {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct BudgetError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Error plumbing of the C ABI (capi.cpp): set the thread's message and return the status; run a
// body, mapping exceptions to statuses.
vt_status fail(vt_status s, const std::string &msg);
vt_status guarded_call(const std::function<vt_status()> &f);

inline void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { reset(); }
    void reset() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        reset();
        if (n == 0) return;
        check(hipMalloc(&ptr, n), "hipMalloc");
        bytes = n;
    }
    void upload(const void *src, size_t n) {
        alloc(n);
        if (n) check(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace vt

struct vt_context {
    int device = 0;
    uint64_t budget = 0;
    uint64_t resident = 0;  // bytes of live scenes
    // Energy-compensation tables of the layered BSDF, baked on the device on first use.
    vt::DeviceBuffer energy_metal, energy_luts;
    bool energy_ready = false;
    // Traversal fault word (TraceScene::fault) of every scene of the context, read by vt_context_check.
    vt::DeviceBuffer fault;
};

namespace vt {
// Bakes ctx->energy_metal / energy_luts once (GPU bakes, host averages and unorm16 quantisation).
void ensure_energy_tables(vt_context *ctx);
// The pinhole camera of the reference's orbit view (eye, target, +y up, vertical fov) and the image
// size into f (renderer.cpp).
void camera_frame(const float eye[3], const float target[3], float fov_degrees, uint32_t width, uint32_t height,
                  FrameArgs &f);
}  // namespace vt

// One published mesh: the reference's MeshStreams + BLAS nodes (Sources/Acceleration.swift:59-64),
// the document-order geometry the shading stage reads, and the material / emitter tables.
struct vt_scene {
    vt_context *ctx = nullptr;
    vt::DeviceBuffer nodes, tris, counters;
    vt::DeviceBuffer vertices, faces;             // document order
    vt::DeviceBuffer materials, emitter_face, emitter_cdf, icdf, cmf, basis;
    vt::DeviceBuffer tree_nodes, tree_leaves, tree_ids, tree_pos, tree_dec;
    // Environment map (rgba32f, lat-long) and its sampling tables (vt_scene_set_environment).
    vt::DeviceBuffer env_image, env_rows, env_cols;
    uint32_t env_w = 0, env_h = 0;
    uint64_t env_accounted = 0;
    uint32_t tree_node_count = 0;
    float tree_lo[3] = {0, 0, 0}, tree_scale[3] = {0, 0, 0};
    bool light_tree = true;
    vt::TraceScene view{};
    vt_scene_info info{};
    uint32_t material_count = 0, emitter_count = 0;
    float emitter_total = 0.0f;
    bool shaded = false;  // materials set
    bool has_delta = false;  // some material is a delta (mirror or smooth glass)
    bool has_layered = false;  // some material is GLOSSY or OPENPBR
    bool mixed = false;        // materials of both shading classes (layered, and Lambert / delta)
    bool box_valid = false;    // box_lo / box_hi computed (the radiance cache's position frame)
    float box_lo[3] = {}, box_hi[3] = {};
    // Identity of the scene's current contents: a process-unique id, bumped by every material or
    // light-tree change (renderers drop ReSTIR history rendered with other contents).
    uint64_t generation = next_generation();
    static uint64_t next_generation() {
        static std::atomic<uint64_t> counter{1};
        return counter.fetch_add(1);
    }
    std::vector<float> host_vertices;
    std::vector<uint32_t> host_faces;
    std::unique_ptr<vt::SpectralTables> spectral;
    uint64_t accounted = 0;
    ~vt_scene() {
        if (ctx) ctx->resident -= accounted + env_accounted;
    }
    vt::RenderScene render_view() const {
        vt::RenderScene r{};
        r.trace = view;
        r.vertices = vertices.as<const float>();
        r.faces = faces.as<const uint32_t>();
        r.materials = materials.as<const float>();
        r.material_count = material_count;
        r.face_count = (uint32_t)(host_faces.size() / 4);
        r.emitter_face = emitter_face.as<const uint32_t>();
        r.emitter_cdf = emitter_cdf.as<const float>();
        r.emitter_count = emitter_count;
        r.emitter_total = emitter_total;
        r.icdf = icdf.as<const float>();
        r.cmf = cmf.as<const float>();
        r.basis = basis.as<const float>();
        r.tree_nodes = tree_nodes.as<const uint32_t>();
        r.tree_dec = tree_dec.as<const float4>();
        r.tree_leaves = tree_leaves.as<const uint32_t>();
        r.tree_ids = tree_ids.as<const uint32_t>();
        r.tree_pos = tree_pos.as<const uint32_t>();
        r.tree_node_count = tree_node_count;
        r.energy_metal = ctx->energy_metal.as<const float>();
        r.energy_luts = ctx->energy_luts.as<const float>();
        r.layered = has_layered ? 1u : 0u;
        r.env = env_image.as<const float4>();
        r.env_rows = env_rows.as<const float>();
        r.env_cols = env_cols.as<const float>();
        r.env_w = env_w;
        r.env_h = env_h;
        for (int k = 0; k < 3; ++k) {
            r.tree_lo[k] = tree_lo[k];
            r.tree_scale[k] = tree_scale[k];
        }
        return r;
    }
};
