// capi.cpp -- the C ABI (include/vt.h) over the host runtime: device context, scene publish
// (MaterialLibrary::setMesh) and the trace entry points.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include <functional>

#include "bvh8.h"
#include "scene.h"
#include "trace_kernels.h"
#include "vt.h"
#include "vt_mesh.h"

namespace {

thread_local std::string g_error;

using vt::BudgetError;
using vt::check;
using vt::DeviceBuffer;
using vt::DeviceError;

}  // namespace

namespace vt {

vt_status fail(vt_status s, const std::string &msg) {
    g_error = msg;
    return s;
}

vt_status guarded_call(const std::function<vt_status()> &f) {
    try {
        g_error.clear();
        return f();
    } catch (const BudgetError &e) {
        return fail(VT_ERROR_OUT_OF_MEMORY, e.what());
    } catch (const DeviceError &e) {
        return fail(VT_ERROR_DEVICE, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(VT_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::bad_alloc &) {
        return fail(VT_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(VT_ERROR_DEVICE, e.what());
    }
}

}  // namespace vt

using vt::fail;
template <class F>
static vt_status guarded(F &&f) {
    return vt::guarded_call(std::function<vt_status()>(std::forward<F>(f)));
}

namespace vt {

// Mirrors MaterialLibrary.setMesh for an indexed mesh (Sources/StudioModel.swift:1480-1504 with
// Sources/Acceleration.swift:652-782): validate, build the hierarchy on the host, check the GPU
// budget with the reference's message, then copy the streams to device memory.
// The scene's hierarchy as the traversal reads it: the 8-wide tree, or with opt.collapse == 2 the
// reference's flat binary median hierarchy.  Built on the host; `tree` receives the node and triangle
// records, the view and info fields the chosen kind.
struct HostTree {
    std::vector<uint8_t> nodes, tris;
    uint32_t node_count = 0, depth = 0, kind = 0, width = 8, node_bytes = 0;
    float root_max = 0;
    double seconds = 0;
};
static HostTree build_tree(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count,
                           const BuildOptions &opt) {
    HostTree t;
    auto copy = [](std::vector<uint8_t> &dst, const void *src, size_t bytes) {
        dst.resize(bytes);
        if (bytes) std::memcpy(dst.data(), src, bytes);
    };
    if (opt.collapse == 2) {
        MedianBvh m = build_median_bvh(vertices, vertex_count, faces, face_count);
        if (m.depth > (int)kMedianStack - 1) throw std::invalid_argument("mesh hierarchy deeper than the traversal stack");
        copy(t.nodes, m.nodes.data(), m.nodes.size() * sizeof(MeshNode));
        copy(t.tris, m.tris.data(), m.tris.size() * sizeof(TriRecord));
        t.node_count = (uint32_t)m.nodes.size();
        t.depth = (uint32_t)m.depth;
        t.kind = 1;
        t.width = 2;
        t.node_bytes = sizeof(MeshNode);
        t.root_max = m.root_max;
        t.seconds = m.build_seconds;
        return t;
    }
    Bvh8 bvh = build_bvh8(vertices, vertex_count, faces, face_count, opt);
    if (bvh.nodes.size() >= (1u << 26)) throw std::invalid_argument("mesh exceeds the node index range");
    copy(t.nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(Node8));
    copy(t.tris, bvh.tris.data(), bvh.tris.size() * sizeof(TriRecord));
    t.node_count = (uint32_t)bvh.nodes.size();
    t.depth = (uint32_t)bvh.depth;
    t.node_bytes = sizeof(Node8);
    t.root_max = bvh.root_max;
    t.seconds = bvh.build_seconds;
    return t;
}
static void publish_tree(vt_scene *s, const HostTree &t, const BuildOptions &opt) {
    s->nodes.upload(t.nodes.data(), t.nodes.size());
    s->tris.upload(t.tris.data(), t.tris.size());
    s->view.nodes = s->nodes.ptr;
    s->view.tris = s->tris.ptr;
    s->view.node_count = t.node_count;
    s->view.depth = t.depth;
    s->view.root_max = t.root_max;
    s->view.kind = t.kind;
    vt_scene_info &i = s->info;
    i.node_count = t.node_count;
    i.depth = t.depth;
    i.leaf_size = t.kind == 1 ? 4u : (uint32_t)opt.leaf_size;
    i.width = t.width;
    i.node_bytes = t.node_bytes;
    i.tri_bytes = sizeof(TriRecord);
    i.root_max = t.root_max;
    i.build_seconds = (float)t.seconds;
}

// Mirrors MaterialLibrary.setMesh for an indexed mesh (Sources/StudioModel.swift:1480-1504 with
// Sources/Acceleration.swift:652-782): validate, build the hierarchy on the host, check the GPU
// budget with the reference's message, then copy the streams to device memory.
std::unique_ptr<vt_scene> set_mesh(vt_context *ctx, const float *vertices, uint32_t vertex_count,
                                   const uint32_t *faces, uint32_t face_count, const BuildOptions &opt) {
    if (face_count > 0 && (!vertices || !faces)) throw std::invalid_argument("vertices and faces must not be NULL");
    if (face_count >= (1u << 30)) throw std::invalid_argument("Stored mesh assets exceed the 1,073,741,824 triangle limit.");
    if (opt.collapse < 0 || opt.collapse > 2) throw std::invalid_argument("collapse must be 0, 1 or 2");
    HostTree tree = build_tree(vertices, vertex_count, faces, face_count, opt);
    const uint64_t bytes = tree.nodes.size() + tree.tris.size() + (uint64_t)vertex_count * 32 + (uint64_t)face_count * 16;
    if (ctx->resident + bytes > ctx->budget)
        throw BudgetError("The imported meshes exceed the safe GPU memory budget.");
    check(hipSetDevice(ctx->device), "hipSetDevice");
    auto s = std::make_unique<vt_scene>();
    publish_tree(s.get(), tree, opt);
    // Document-order geometry for hit attributes and emitter sampling (the reference's MeshStreams).
    s->vertices.upload(vertices, (size_t)vertex_count * 8 * sizeof(float));
    s->faces.upload(faces, (size_t)face_count * 4 * sizeof(uint32_t));
    s->host_vertices.assign(vertices, vertices + (size_t)vertex_count * 8);
    s->host_faces.assign(faces, faces + (size_t)face_count * 4);
    s->ctx = ctx;
    s->accounted = s->nodes.bytes + s->tris.bytes + s->vertices.bytes + s->faces.bytes;
    ctx->resident += s->accounted;
    s->view.fault = ctx->fault.as<uint32_t>();
    s->info.face_count = face_count;
    s->info.vertex_count = vertex_count;
    s->info.device_bytes = bytes;
    return s;
}

// Rebuilds a published scene's hierarchy from its document-order streams with other options (the
// autotuner's leaf-size sweep); materials and the light tree are kept.
void rebuild_bvh(vt_scene *s, const BuildOptions &opt) {
    if (opt.collapse < 0 || opt.collapse > 2) throw std::invalid_argument("collapse must be 0, 1 or 2");
    const uint32_t vertex_count = (uint32_t)(s->host_vertices.size() / 8), face_count = (uint32_t)(s->host_faces.size() / 4);
    HostTree tree = build_tree(s->host_vertices.data(), vertex_count, s->host_faces.data(), face_count, opt);
    check(hipSetDevice(s->ctx->device), "hipSetDevice");
    s->ctx->resident -= s->nodes.bytes + s->tris.bytes;
    s->accounted -= s->nodes.bytes + s->tris.bytes;
    publish_tree(s, tree, opt);
    s->ctx->resident += s->nodes.bytes + s->tris.bytes;
    s->accounted += s->nodes.bytes + s->tris.bytes;
    s->info.device_bytes = tree.nodes.size() + tree.tris.size() + (uint64_t)vertex_count * 32 + (uint64_t)face_count * 16;
    s->generation = vt_scene::next_generation();
}

}  // namespace vt

extern "C" {

const char *vt_last_error(void) { return g_error.c_str(); }


vt_status vt_context_create(const vt_context_desc *desc, vt_context **out) {
    return guarded([&]() -> vt_status {
        if (!out) return fail(VT_ERROR_INVALID_ARGUMENT, "out must not be NULL");
        *out = nullptr;
        int device = desc ? desc->device : 0;
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) return fail(VT_ERROR_INVALID_ARGUMENT, "no such HIP device");
        check(hipSetDevice(device), "hipSetDevice");
        size_t free_b = 0, total = 0;
        check(hipMemGetInfo(&free_b, &total), "hipMemGetInfo");
        auto c = std::make_unique<vt_context>();
        c->device = device;
        c->budget = desc && desc->memory_budget ? desc->memory_budget
                                                : std::max<uint64_t>(256ull << 20, (uint64_t)total / 4);
        const uint32_t zero = 0;
        c->fault.upload(&zero, 4);
        check(vt::scratch_init(device), "hipGetSymbolAddress");
        *out = c.release();
        return VT_OK;
    });
}

void vt_context_destroy(vt_context *ctx) { delete ctx; }

vt_status vt_context_check(vt_context *ctx) {
    return guarded([&]() -> vt_status {
        if (!ctx) return fail(VT_ERROR_INVALID_ARGUMENT, "context must not be NULL");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        uint32_t word = 0;
        check(hipMemcpy(&word, ctx->fault.ptr, 4, hipMemcpyDeviceToHost), "hipMemcpy");
        if (word) {
            const uint32_t zero = 0;
            check(hipMemcpy(ctx->fault.ptr, &zero, 4, hipMemcpyHostToDevice), "hipMemcpy");
            return fail(VT_ERROR_DEVICE, "BVH traversal stack overflow: some rays ended before their traversal did");
        }
        return VT_OK;
    });
}

vt_status vt_scene_set_stack_cap(vt_scene *scene, uint32_t groups) {
    return guarded([&]() -> vt_status {
        if (!scene) return fail(VT_ERROR_INVALID_ARGUMENT, "scene must not be NULL");
        if (groups > 24) return fail(VT_ERROR_INVALID_ARGUMENT, "groups must be 0..24");
        scene->view.stack_cap = groups;
        scene->generation = vt_scene::next_generation();
        return VT_OK;
    });
}

static vt::BuildOptions build_options(const vt_build_options *opts) {
    vt::BuildOptions o;
    if (opts) {
        if (opts->leaf_size) o.leaf_size = (int)opts->leaf_size;
        if (opts->bins) o.bins = (int)opts->bins;
        if (opts->max_depth) o.max_depth = (int)opts->max_depth;
        o.collapse = (int)opts->collapse;
    }
    return o;
}

vt_status vt_scene_rebuild(vt_scene *scene, const vt_build_options *opts) {
    return guarded([&]() -> vt_status {
        if (!scene) return fail(VT_ERROR_INVALID_ARGUMENT, "scene must not be NULL");
        vt::rebuild_bvh(scene, build_options(opts));
        return VT_OK;
    });
}

vt_status vt_hierarchy_build_host(const float *vertices, uint32_t vertex_count, const uint32_t *faces,
                                  uint32_t face_count, uint32_t **order, void **nodes, uint32_t *node_count,
                                  uint32_t *depth) {
    return guarded([&]() -> vt_status {
        if (!order || !nodes || !node_count || !depth) return fail(VT_ERROR_INVALID_ARGUMENT, "outputs must not be NULL");
        if (face_count > 0 && (!vertices || !faces)) return fail(VT_ERROR_INVALID_ARGUMENT, "vertices and faces must not be NULL");
        *order = nullptr;
        *nodes = nullptr;
        vt::MedianBvh m = vt::build_median_bvh(vertices, vertex_count, faces, face_count);
        void *o = std::malloc(std::max<size_t>(1, m.order.size() * sizeof(uint32_t)));
        void *n = std::malloc(std::max<size_t>(1, m.nodes.size() * sizeof(vt::MeshNode)));
        if (!o || !n) { std::free(o); std::free(n); throw std::bad_alloc(); }
        if (!m.order.empty()) std::memcpy(o, m.order.data(), m.order.size() * sizeof(uint32_t));
        if (!m.nodes.empty()) std::memcpy(n, m.nodes.data(), m.nodes.size() * sizeof(vt::MeshNode));
        *order = static_cast<uint32_t *>(o);
        *nodes = n;
        *node_count = (uint32_t)m.nodes.size();
        *depth = (uint32_t)m.depth;
        return VT_OK;
    });
}

vt_status vt_bvh8_build_host(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count,
                             const vt_build_options *opts, void **nodes, uint32_t *node_count, void **tris,
                             uint32_t *depth, float *root_max) {
    return guarded([&]() -> vt_status {
        if (!nodes || !node_count || !tris || !depth || !root_max)
            return fail(VT_ERROR_INVALID_ARGUMENT, "outputs must not be NULL");
        if (face_count > 0 && (!vertices || !faces)) return fail(VT_ERROR_INVALID_ARGUMENT, "vertices and faces must not be NULL");
        *nodes = nullptr;
        *tris = nullptr;
        vt::Bvh8 b = vt::build_bvh8(vertices, vertex_count, faces, face_count, build_options(opts));
        void *n = std::malloc(std::max<size_t>(1, b.nodes.size() * sizeof(vt::Node8)));
        void *t = std::malloc(std::max<size_t>(1, b.tris.size() * sizeof(vt::TriRecord)));
        if (!n || !t) { std::free(n); std::free(t); throw std::bad_alloc(); }
        if (!b.nodes.empty()) std::memcpy(n, b.nodes.data(), b.nodes.size() * sizeof(vt::Node8));
        if (!b.tris.empty()) std::memcpy(t, b.tris.data(), b.tris.size() * sizeof(vt::TriRecord));
        *nodes = n;
        *tris = t;
        *node_count = (uint32_t)b.nodes.size();
        *depth = (uint32_t)b.depth;
        *root_max = b.root_max;
        return VT_OK;
    });
}

void vt_host_free(void *p) { std::free(p); }

vt_status vt_scene_upload_indexed(vt_context *ctx, const float *vertices, uint32_t vertex_count, const uint32_t *faces,
                                  uint32_t face_count, const vt_build_options *opts, vt_scene **out) {
    return guarded([&]() -> vt_status {
        if (!ctx || !out) return fail(VT_ERROR_INVALID_ARGUMENT, "context and out must not be NULL");
        *out = nullptr;
        *out = vt::set_mesh(ctx, vertices, vertex_count, faces, face_count, build_options(opts)).release();
        return VT_OK;
    });
}

vt_status vt_scene_upload_triangles(vt_context *ctx, const float *triangles, uint32_t count,
                                    const vt_build_options *opts, vt_scene **out) {
    return guarded([&]() -> vt_status {
        if (!ctx || !out) return fail(VT_ERROR_INVALID_ARGUMENT, "context and out must not be NULL");
        *out = nullptr;
        float *v = nullptr;
        uint32_t *f = nullptr, nv = 0, nf = 0;
        const vt_status st = vt_mesh_from_triangles(triangles, count, &v, &nv, &f, &nf);
        if (st != VT_OK) return st;
        std::unique_ptr<float, void (*)(void *)> vg(v, std::free);
        std::unique_ptr<uint32_t, void (*)(void *)> fg(f, std::free);
        *out = vt::set_mesh(ctx, v, nv, f, nf, build_options(opts)).release();
        return VT_OK;
    });
}

void vt_scene_destroy(vt_scene *scene) { delete scene; }

vt_status vt_scene_get_info(const vt_scene *scene, vt_scene_info *info) {
    if (!scene || !info) return fail(VT_ERROR_INVALID_ARGUMENT, "scene and info must not be NULL");
    *info = scene->info;
    return VT_OK;
}

static bool rays_ok(const vt_ray_soa *r) {
    return r && r->ox && r->oy && r->oz && r->dx && r->dy && r->dz && r->tmin && r->tmax;
}

static vt::RayArgs ray_args(const vt_ray_soa *r) {
    return vt::RayArgs{r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax};
}

vt_status vt_trace_closest(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, const vt_hit_soa *hits, void *stream) {
    return guarded([&]() -> vt_status {
        if (!scene) return fail(VT_ERROR_INVALID_ARGUMENT, "scene must not be NULL");
        if (n == 0) return VT_OK;
        if (!rays_ok(rays)) return fail(VT_ERROR_INVALID_ARGUMENT, "ray arrays must not be NULL");
        if (!hits || !hits->t || !hits->b1 || !hits->b2 || !hits->face)
            return fail(VT_ERROR_INVALID_ARGUMENT, "hit arrays must not be NULL");
        check(hipSetDevice(scene->ctx->device), "hipSetDevice");
        vt::HitArgs h{hits->t, hits->b1, hits->b2, hits->face, hits->instance};
        check(vt::launch_trace(scene->view, ray_args(rays), h, nullptr, n, false, nullptr, (hipStream_t)stream),
              "trace_closest launch");
        return VT_OK;
    });
}

vt_status vt_trace_any(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, uint8_t *occluded, void *stream) {
    return guarded([&]() -> vt_status {
        if (!scene) return fail(VT_ERROR_INVALID_ARGUMENT, "scene must not be NULL");
        if (n == 0) return VT_OK;
        if (!rays_ok(rays) || !occluded) return fail(VT_ERROR_INVALID_ARGUMENT, "ray and result arrays must not be NULL");
        check(hipSetDevice(scene->ctx->device), "hipSetDevice");
        vt::HitArgs h{};
        check(vt::launch_trace(scene->view, ray_args(rays), h, occluded, n, true, nullptr, (hipStream_t)stream),
              "trace_any launch");
        return VT_OK;
    });
}

vt_status vt_trace_stats_noalloc(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, int any,
                                 const vt_hit_soa *hits, uint8_t *occluded, uint64_t *counters, void *stream) {
    return guarded([&]() -> vt_status {
        if (!scene || !counters) return fail(VT_ERROR_INVALID_ARGUMENT, "scene and counters must not be NULL");
        if (scene->view.kind != 0u)
            return fail(VT_ERROR_UNSUPPORTED, "traversal counters are kept for the 8-wide hierarchy only");
        if (n && !rays_ok(rays)) return fail(VT_ERROR_INVALID_ARGUMENT, "ray arrays must not be NULL");
        vt::HitArgs h{};
        if (n && any && !occluded) return fail(VT_ERROR_INVALID_ARGUMENT, "occluded must not be NULL");
        if (n && !any) {
            if (!hits || !hits->t || !hits->b1 || !hits->b2 || !hits->face)
                return fail(VT_ERROR_INVALID_ARGUMENT, "hit arrays must not be NULL");
            h = vt::HitArgs{hits->t, hits->b1, hits->b2, hits->face, hits->instance};
        }
        check(hipSetDevice(scene->ctx->device), "hipSetDevice");
        const vt::RayArgs r = n ? ray_args(rays) : vt::RayArgs{};
        check(vt::launch_trace_stats_scratch(scene->ctx->device, scene->view, r, h, any ? occluded : nullptr, n, any != 0,
                                             (unsigned long long *)counters, (hipStream_t)stream),
              "trace_stats_noalloc launch");
        return VT_OK;
    });
}

vt_status vt_trace_stats(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, int any, uint64_t counters[2],
                         void *stream) {
    return guarded([&]() -> vt_status {
        if (!scene || !counters) return fail(VT_ERROR_INVALID_ARGUMENT, "scene and counters must not be NULL");
        counters[0] = counters[1] = 0;
        if (scene->view.kind != 0u)
            return fail(VT_ERROR_UNSUPPORTED, "traversal counters are kept for the 8-wide hierarchy only");
        if (n == 0) return VT_OK;
        if (!rays_ok(rays)) return fail(VT_ERROR_INVALID_ARGUMENT, "ray arrays must not be NULL");
        check(hipSetDevice(scene->ctx->device), "hipSetDevice");
        const size_t cbytes = 128 * sizeof(unsigned long long);
        if (!scene->counters.ptr) {
            check(hipMalloc(&scene->counters.ptr, cbytes), "hipMalloc");
            scene->counters.bytes = cbytes;
        }
        hipStream_t s = (hipStream_t)stream;
        check(hipMemsetAsync(scene->counters.ptr, 0, cbytes, s), "hipMemsetAsync");
        // Scratch outputs: the counting variant writes results like the timed kernels do.
        DeviceBuffer out;
        const size_t per = any ? 1 : (4 * sizeof(float) + sizeof(int32_t));
        check(hipMalloc(&out.ptr, per * (size_t)n + 64), "hipMalloc");
        out.bytes = per * (size_t)n + 64;
        char *base = (char *)out.ptr;
        vt::HitArgs h{};
        uint8_t *occ = nullptr;
        if (any) {
            occ = (uint8_t *)base;
        } else {
            h.t = (float *)base;
            h.b1 = h.t + n;
            h.b2 = h.b1 + n;
            h.face = (int32_t *)(h.b2 + n);
            h.instance = nullptr;
        }
        check(vt::launch_trace(scene->view, ray_args(rays), h, occ, n, any != 0,
                               (unsigned long long *)scene->counters.ptr, s),
              "trace_stats launch");
        unsigned long long host[128];
        check(hipMemcpyAsync(host, scene->counters.ptr, cbytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
        for (int k = 0; k < 64; ++k) {
            counters[0] += host[2 * k];
            counters[1] += host[2 * k + 1];
        }
        return VT_OK;
    });
}

}  // extern "C"
