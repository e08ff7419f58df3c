// tlas.cpp -- two-level scenes (vt_tlas): instances of published scenes under object-to-world
// transforms, an 8-wide hierarchy over their world bounds, and the trace entry points over them.
//
// Follows MeshSceneLayout.addInstance / sceneBlob (Sources/Acceleration.swift:470-573): the same
// transform checks and messages, world bounds from the transformed corners of the asset's bounds
// grown by 2^-20 of the instance's reach, the per-ray growth of the instance boxes
// 2^-18 (1 + max condition number) (|origin|max + max reach), and the instance hierarchy built with
// one instance per leaf.  The matrix work is done in double and rounded once to float.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh8.h"
#include "scene.h"
#include "trace_kernels.h"
#include "vt.h"

struct vt_tlas {
    vt_context *ctx = nullptr;
    std::vector<vt_instance> list;
    std::vector<uint32_t> leaf_items;               // instance index of each leaf position
    std::vector<std::array<float, 12>> local_rows;  // world-to-object rows per instance
    std::vector<const vt_scene *> assets;           // distinct assets
    std::vector<uint64_t> asset_generation;         // their generation when the records were written
    vt::DeviceBuffer nodes, records;
    vt::TlasScene view{};
    vt_tlas_info info{};
    uint64_t accounted = 0;
    ~vt_tlas() {
        if (ctx) ctx->resident -= accounted;
    }
};

namespace vt {
namespace {

struct AssetBounds {
    double lo[3], hi[3];
    double max_abs;
    bool empty;
};

// Bounds of the vertices the asset's faces use (its BLAS root box).
AssetBounds asset_bounds(const vt_scene *s) {
    AssetBounds b{};
    const size_t nf = s->host_faces.size() / 4;
    b.empty = nf == 0;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = HUGE_VAL;
        b.hi[a] = -HUGE_VAL;
    }
    for (size_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            const float *v = s->host_vertices.data() + 8 * (size_t)s->host_faces[4 * f + c];
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = std::min(b.lo[a], (double)v[a]);
                b.hi[a] = std::max(b.hi[a], (double)v[a]);
            }
        }
    b.max_abs = 0;
    if (!b.empty)
        for (int a = 0; a < 3; ++a) b.max_abs = std::max(b.max_abs, std::max(std::fabs(b.lo[a]), std::fabs(b.hi[a])));
    return b;
}

// Infinity norm (max row sum of |entries|) of a 3x3 matrix.
double norm_inf(const double m[3][3]) {
    double n = 0;
    for (int r = 0; r < 3; ++r) n = std::max(n, std::fabs(m[r][0]) + std::fabs(m[r][1]) + std::fabs(m[r][2]));
    return n;
}

// Inverse of the affine map x -> A x + t by cofactors: rows of [A^-1 | -A^-1 t], each entry rounded
// once to float.  det receives det(A).
void affine_inverse(const float w[12], float out[12], double &det, double inv[3][3]) {
    const double a00 = w[0], a01 = w[1], a02 = w[2], a10 = w[4], a11 = w[5], a12 = w[6];
    const double a20 = w[8], a21 = w[9], a22 = w[10];
    const double c[3][3] = {{a11 * a22 - a12 * a21, a02 * a21 - a01 * a22, a01 * a12 - a02 * a11},
                            {a12 * a20 - a10 * a22, a00 * a22 - a02 * a20, a02 * a10 - a00 * a12},
                            {a10 * a21 - a11 * a20, a01 * a20 - a00 * a21, a00 * a11 - a01 * a10}};
    det = a00 * c[0][0] + a01 * c[1][0] + a02 * c[2][0];
    const double t[3] = {w[3], w[7], w[11]};
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) inv[r][k] = c[r][k] / det;
        const double tr = -(inv[r][0] * t[0] + inv[r][1] * t[1] + inv[r][2] * t[2]);
        for (int k = 0; k < 3; ++k) out[4 * r + k] = (float)inv[r][k];
        out[4 * r + 3] = (float)tr;
    }
}

float round_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -HUGE_VALF);
    return f;
}
float round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, HUGE_VALF);
    return f;
}

// Instance records in leaf order from the assets' current hierarchies.
void write_records(vt_tlas *t) {
    std::vector<InstanceRecord> rec(t->leaf_items.size());
    uint32_t blas_depth = 1;
    for (size_t p = 0; p < rec.size(); ++p) {
        const uint32_t k = t->leaf_items[p];
        const vt_scene *s = t->list[k].mesh;
        InstanceRecord &r = rec[p];
        std::memset(&r, 0, sizeof r);
        for (int row = 0; row < 3; ++row)
            for (int c = 0; c < 4; ++c) r.local[row][c] = t->local_rows[k][4 * row + c];
        r.nodes = s->view.nodes;
        r.tris = s->view.tris;
        r.root_max = s->view.root_max;
        r.index = k;
        r.faces = s->view.node_count ? s->info.face_count : 0u;
        blas_depth = std::max(blas_depth, s->view.depth);
    }
    check(hipSetDevice(t->ctx->device), "hipSetDevice");
    t->ctx->resident -= t->records.bytes;
    t->accounted -= t->records.bytes;
    t->records.upload(rec.data(), rec.size() * sizeof(InstanceRecord));
    t->ctx->resident += t->records.bytes;
    t->accounted += t->records.bytes;
    t->view.instances = t->records.as<const InstanceRecord>();
    t->view.blas_depth = blas_depth;
    t->info.blas_depth = blas_depth;
    for (size_t a = 0; a < t->assets.size(); ++a) t->asset_generation[a] = t->assets[a]->generation;
}

std::unique_ptr<vt_tlas> build_tlas(vt_context *ctx, const vt_instance *list, uint32_t count) {
    auto t0 = std::chrono::steady_clock::now();
    auto t = std::make_unique<vt_tlas>();
    t->ctx = ctx;
    t->list.assign(list, list + count);
    t->local_rows.resize(count);
    std::vector<float> boxes(6 * (size_t)count);
    double condition = 0, reach_max = 0;
    uint64_t rendered = 0;
    std::vector<std::pair<const vt_scene *, AssetBounds>> seen;
    for (uint32_t k = 0; k < count; ++k) {
        const vt_instance &in = list[k];
        if (!in.mesh) throw std::invalid_argument("instance " + std::to_string(k) + ": mesh must not be NULL");
        if (in.mesh->view.kind != 0u)
            throw std::invalid_argument("instance " + std::to_string(k) + ": instanced meshes need an 8-wide hierarchy "
                                        "(build option collapse 0 or 1)");
        if (in.mesh->ctx != ctx) throw std::invalid_argument("instance " + std::to_string(k) + ": mesh of another context");
        for (int q = 0; q < 12; ++q)
            if (!std::isfinite(in.world[q])) throw std::invalid_argument("Singular USD transform.");
        const AssetBounds *ab = nullptr;
        for (auto &e : seen)
            if (e.first == in.mesh) ab = &e.second;
        if (!ab) {
            seen.push_back({in.mesh, asset_bounds(in.mesh)});
            ab = &seen.back().second;
            t->assets.push_back(in.mesh);
        }
        // Sources/Acceleration.swift:477-499: scale range, singular transforms, scene extent.
        const double scale = std::sqrt((double)in.world[0] * in.world[0] + (double)in.world[4] * in.world[4] +
                                       (double)in.world[8] * in.world[8]);
        if (!std::isfinite(scale) || scale < 0.000001 || scale > 1000000.0)
            throw std::invalid_argument("Combined hierarchy scale exceeds the supported range.");
        double det = 0, inv[3][3];
        float local[12];
        affine_inverse(in.world, local, det, inv);
        if (!std::isfinite(det) || std::fabs(det) <= 1e-18) throw std::invalid_argument("Singular USD transform.");
        for (int q = 0; q < 12; ++q)
            if (!std::isfinite(local[q])) throw std::invalid_argument("Singular USD transform.");
        std::memcpy(t->local_rows[k].data(), local, sizeof local);
        const double lin[3][3] = {{in.world[0], in.world[1], in.world[2]},
                                  {in.world[4], in.world[5], in.world[6]},
                                  {in.world[8], in.world[9], in.world[10]}};
        const double n = norm_inf(lin);
        const double tmax = std::max(std::fabs((double)in.world[3]), std::max(std::fabs((double)in.world[7]), std::fabs((double)in.world[11])));
        const double reach = tmax + n * ab->max_abs;
        if (!(reach < 1e8)) {
            const vt_scene *s = in.mesh;
            for (size_t v = 0; v < s->host_vertices.size() / 8; ++v) {
                const float *p = s->host_vertices.data() + 8 * v;
                for (int r = 0; r < 3; ++r) {
                    const double x = lin[r][0] * p[0] + lin[r][1] * p[1] + lin[r][2] * p[2] + in.world[4 * r + 3];
                    if (!std::isfinite(x) || !(std::fabs(x) < 1e8))
                        throw std::invalid_argument("Hierarchy transform exceeds the supported scene extent.");
                }
            }
        }
        condition = std::max(condition, n * norm_inf(inv));
        reach_max = std::max(reach_max, reach);
        rendered += in.mesh->info.face_count;
        // World bounds: the transformed corners of the asset's bounds, rounded outward, grown by
        // 2^-20 of the reach (an empty asset gets a point box at its translation).
        double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        for (int c = 0; c < 8; ++c) {
            const double p[3] = {ab->empty ? 0.0 : ((c & 1) ? ab->hi[0] : ab->lo[0]),
                                 ab->empty ? 0.0 : ((c & 2) ? ab->hi[1] : ab->lo[1]),
                                 ab->empty ? 0.0 : ((c & 4) ? ab->hi[2] : ab->lo[2])};
            for (int r = 0; r < 3; ++r) {
                const double x = lin[r][0] * p[0] + lin[r][1] * p[1] + lin[r][2] * p[2] + in.world[4 * r + 3];
                lo[r] = std::min(lo[r], x);
                hi[r] = std::max(hi[r], x);
            }
        }
        const double margin = std::ldexp(reach, -20);
        for (int r = 0; r < 3; ++r) {
            boxes[6 * (size_t)k + r] = round_down(lo[r] - margin);
            boxes[6 * (size_t)k + 3 + r] = round_up(hi[r] + margin);
        }
    }
    BuildOptions opt;
    opt.leaf_size = 1;
    opt.max_depth = 17;  // 16 node groups at most pending on the instance level
    Bvh8 top = build_bvh8_boxes(boxes.data(), count, opt, t->leaf_items);
    t->asset_generation.resize(t->assets.size());
    check(hipSetDevice(ctx->device), "hipSetDevice");
    const uint64_t bytes = top.nodes.size() * sizeof(Node8) + (uint64_t)count * sizeof(InstanceRecord);
    if (ctx->resident + bytes > ctx->budget) throw BudgetError("The imported meshes exceed the safe GPU memory budget.");
    t->nodes.upload(top.nodes.data(), top.nodes.size() * sizeof(Node8));
    t->accounted = t->nodes.bytes;
    ctx->resident += t->nodes.bytes;
    t->view.nodes = t->nodes.ptr;
    t->view.node_count = (uint32_t)top.nodes.size();
    t->view.depth = (uint32_t)std::max(top.depth, 1);
    t->view.enlarge = (float)std::ldexp(1.0 + condition, -18);
    t->view.reach = (float)reach_max;
    t->view.fault = ctx->fault.as<uint32_t>();
    write_records(t.get());
    vt_tlas_info &i = t->info;
    i.instance_count = count;
    i.node_count = t->view.node_count;
    i.depth = t->view.depth;
    i.rendered_triangles = rendered;
    i.enlarge = t->view.enlarge;
    i.reach = t->view.reach;
    i.device_bytes = bytes;
    i.build_seconds = (float)std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return t;
}

// Rewrites the records when an asset's hierarchy was rebuilt since they were written.
void refresh(vt_tlas *t) {
    for (size_t a = 0; a < t->assets.size(); ++a)
        if (t->assets[a]->generation != t->asset_generation[a]) {
            if (t->assets[a]->view.kind != 0u)
                throw std::invalid_argument("an instanced mesh was rebuilt as the flat binary hierarchy (collapse 2); "
                                            "instances need an 8-wide hierarchy");
            check(hipSetDevice(t->ctx->device), "hipSetDevice");
            check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            write_records(t);
            return;
        }
}

}  // namespace
}  // namespace vt

extern "C" {

vt_status vt_tlas_create(vt_context *ctx, const vt_instance *instances, uint32_t count, vt_tlas **out) {
    return vt::guarded_call([&]() -> vt_status {
        if (!ctx || !out) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "context and out must not be NULL");
        *out = nullptr;
        if (count > 0 && !instances) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "instances must not be NULL");
        if (count >= (1u << 24)) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "too many instances");
        *out = vt::build_tlas(ctx, instances, count).release();
        return VT_OK;
    });
}

void vt_tlas_destroy(vt_tlas *tlas) { delete tlas; }

vt_status vt_tlas_get_info(const vt_tlas *tlas, vt_tlas_info *info) {
    if (!tlas || !info) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "tlas and info must not be NULL");
    *info = tlas->info;
    return VT_OK;
}

static vt::RayArgs ray_args(const vt_ray_soa *r) {
    return vt::RayArgs{r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax};
}

vt_status vt_tlas_trace_closest(vt_tlas *tlas, const vt_ray_soa *rays, uint32_t n, const vt_hit_soa *hits,
                                void *stream) {
    return vt::guarded_call([&]() -> vt_status {
        if (!tlas || !rays || !hits) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "tlas, rays and hits must not be NULL");
        if (!hits->t || !hits->b1 || !hits->b2 || !hits->face)
            return vt::fail(VT_ERROR_INVALID_ARGUMENT, "hit arrays must not be NULL");
        vt::refresh(tlas);
        vt::check(hipSetDevice(tlas->ctx->device), "hipSetDevice");
        vt::HitArgs h{hits->t, hits->b1, hits->b2, hits->face, hits->instance};
        vt::check(vt::launch_trace_tlas(tlas->view, ray_args(rays), h, nullptr, n, false, (hipStream_t)stream),
                  "two-level closest-hit launch");
        return VT_OK;
    });
}

vt_status vt_tlas_trace_any(vt_tlas *tlas, const vt_ray_soa *rays, uint32_t n, uint8_t *occluded, void *stream) {
    return vt::guarded_call([&]() -> vt_status {
        if (!tlas || !rays || !occluded) return vt::fail(VT_ERROR_INVALID_ARGUMENT, "tlas, rays and occluded must not be NULL");
        vt::refresh(tlas);
        vt::check(hipSetDevice(tlas->ctx->device), "hipSetDevice");
        vt::HitArgs h{nullptr, nullptr, nullptr, nullptr, nullptr};
        vt::check(vt::launch_trace_tlas(tlas->view, ray_args(rays), h, occluded, n, true, (hipStream_t)stream),
                  "two-level any-hit launch");
        return VT_OK;
    });
}

}  // extern "C"
