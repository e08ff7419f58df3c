// renderer.cpp -- materials, emitter list and the wavefront frame driver behind vt_render_frame.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "lighttree.h"
#include "nrc.h"
#include "render_kernels.h"
#include "scene.h"
#include "spectral.h"
#include "uplift.h"
#include "vt.h"

using vt::check;
using vt::DeviceBuffer;
using vt::fail;

namespace {

// Path-state arrays of one queue: 8 ray arrays, then id, seed, wl, t0-t3, state, pdf, px, py, pz.
constexpr int kQueueArrays = 8 + 12;
constexpr int kShadowArrays = 8 + 4 + 4;  // rays, id, rgb, lane contributions (float4)

struct QueueStorage {
    DeviceBuffer buf;
    vt::PathQueue view(uint32_t cap) const {
        char *b = buf.as<char>();
        auto f = [&](int k) { return reinterpret_cast<float *>(b + (size_t)k * cap * 4); };
        auto u = [&](int k) { return reinterpret_cast<uint32_t *>(b + (size_t)k * cap * 4); };
        vt::PathQueue q;
        q.ox = f(0); q.oy = f(1); q.oz = f(2); q.dx = f(3); q.dy = f(4); q.dz = f(5); q.tmin = f(6); q.tmax = f(7);
        q.id = u(8); q.seed = u(9); q.wl = f(10);
        q.t0 = f(11); q.t1 = f(12); q.t2 = f(13); q.t3 = f(14);
        q.state = u(15); q.pdf = f(16); q.px = f(17); q.py = f(18); q.pz = f(19);
        return q;
    }
};

}  // namespace

struct vt_renderer {
    vt_context *ctx = nullptr;
    uint32_t width = 0, height = 0, frame_index = 0;
    uint32_t capacity = 0;  // paths per wavefront
    DeviceBuffer accum, radiance, hits, occluded, counters, shadow, perm;
    QueueStorage q[2];
    // ReSTIR DI: two sets of G-buffer + reservoir arrays (72 B per pixel each), ping-ponged per
    // sample; restir_history counts consecutive mode-0 samples of this image (reservoirHistory).
    DeviceBuffer restir[2];
    int restir_cur = 0;
    uint32_t restir_history = 0;
    // ReSTIR GI: reservoirs (56 B per pixel each) ping-ponged with the DI ones, and the per-pixel
    // work arrays (141 B); gi_history: the history reservoirs were written by a GI sample.
    DeviceBuffer gi_res[2], gi_work;
    bool gi_history = false;
    uint64_t restir_generation = 0;  // vt_scene::generation the history was rendered with
    uint32_t *host_count = nullptr;  // pinned: live paths, read between vertices of delta scenes
    // Learned radiance cache: the attached network and its queues (sized with the path capacity).
    vt_nrc *cache = nullptr;
    DeviceBuffer nrc_queries, nrc_records, nrc_sacc;
    uint32_t nrc_capacity = 0, nrc_rcap = 0;
    vt::NrcDevice cache_device() const { return vt::nrc_device(cache); }
    uint32_t cache_next_step() { return vt::nrc_next_step(cache); }
    ~vt_renderer() {
        if (host_count) (void)hipHostFree(host_count);
    }
    float prev_camera[14] = {};  // eye, forward, right, up, fov_scale, aspect of the last sample
    vt_frame_fn on_frame = nullptr;  // completion callbacks (vt_renderer_set_callbacks)
    vt_error_fn on_error = nullptr;
    void *callback_user = nullptr;
};

namespace {

void ensure_capacity(vt_renderer *r, uint32_t cap) {
    if (r->capacity >= cap) return;
    r->q[0].buf.alloc((size_t)kQueueArrays * cap * 4);
    r->q[1].buf.alloc((size_t)kQueueArrays * cap * 4);
    r->shadow.alloc((size_t)kShadowArrays * cap * 4);
    r->hits.alloc((size_t)4 * cap * 4);
    r->occluded.alloc(cap);
    r->radiance.alloc((size_t)3 * cap * 4);
    r->perm.alloc((size_t)cap * 4);
    r->capacity = cap;
}

vt::RestirBuffers restir_view(const DeviceBuffer &b, uint32_t npix) {
    char *p = b.as<char>();
    vt::RestirBuffers v;
    v.gpos = reinterpret_cast<float4 *>(p);
    v.gnorm = reinterpret_cast<uint2 *>(p + (size_t)16 * npix);
    v.posdir = reinterpret_cast<float4 *>(p + (size_t)24 * npix);
    v.emitpdf = reinterpret_cast<float4 *>(p + (size_t)40 * npix);
    v.weights = reinterpret_cast<float4 *>(p + (size_t)56 * npix);
    return v;
}

vt::GiBuffers gi_view(const DeviceBuffer &b, uint32_t npix) {
    char *p = b.as<char>();
    vt::GiBuffers v;
    v.pos = reinterpret_cast<float4 *>(p);
    v.rad = reinterpret_cast<float4 *>(p + (size_t)16 * npix);
    v.weights = reinterpret_cast<float4 *>(p + (size_t)32 * npix);
    v.nrm = reinterpret_cast<uint2 *>(p + (size_t)48 * npix);
    return v;
}

// Work arrays: three float4 arrays, 23 four-byte arrays, the visibility bytes.
constexpr size_t kGiWorkBytes = 3 * 16 + 23 * 4 + 1;

vt::GiWork gi_work_view(const DeviceBuffer &b, uint32_t npix) {
    char *p = b.as<char>();
    vt::GiWork w;
    w.light_wi = reinterpret_cast<float4 *>(p);
    w.light_le = reinterpret_cast<float4 *>(p + (size_t)16 * npix);
    w.rgb = reinterpret_cast<float4 *>(p + (size_t)32 * npix);
    char *a = p + (size_t)48 * npix;
    auto f = [&](int k) { return reinterpret_cast<float *>(a + (size_t)k * npix * 4); };
    w.ox = f(0); w.oy = f(1); w.oz = f(2); w.dx = f(3); w.dy = f(4); w.dz = f(5); w.tmin = f(6); w.tmax = f(7);
    w.t = f(8); w.b1 = f(9); w.b2 = f(10);
    w.face = reinterpret_cast<int32_t *>(f(11));
    w.sox = f(12); w.soy = f(13); w.soz = f(14); w.sdx = f(15); w.sdy = f(16); w.sdz = f(17);
    w.stmin = f(18); w.stmax = f(19);
    w.stage = reinterpret_cast<uint32_t *>(f(20));
    w.seed = reinterpret_cast<uint32_t *>(f(21));
    w.bpdf = f(22);
    w.occluded = reinterpret_cast<uint8_t *>(a + (size_t)23 * npix * 4);
    return w;
}

vt::ShadowQueue shadow_view(const vt_renderer *r) {
    char *b = r->shadow.as<char>();
    const size_t cap = r->capacity;
    auto f = [&](int k) { return reinterpret_cast<float *>(b + (size_t)k * cap * 4); };
    vt::ShadowQueue s;
    s.ox = f(0); s.oy = f(1); s.oz = f(2); s.dx = f(3); s.dy = f(4); s.dz = f(5); s.tmin = f(6); s.tmax = f(7);
    s.id = reinterpret_cast<uint32_t *>(f(8));
    s.r = f(9); s.g = f(10); s.b = f(11);
    s.s = reinterpret_cast<float4 *>(f(12));
    return s;
}

// Cache queues: queries (11 inputs, throughput, wavelength, id, 4 outputs) per path slot; records
// (11 inputs, throughput, snapshot, id, 4 targets) for the one cache vertex of each training path,
// sized at twice the expected number of training paths (1 in `period`, chosen by a hash).
void ensure_nrc_buffers(vt_renderer *r, uint32_t period) {
    // Rounded to 64 entries so every float4 array of the layout stays 16-byte aligned.
    const uint32_t cap = (r->capacity + 63u) & ~63u;
    const uint32_t rcap = ((r->capacity / period * 2 + 1024) + 63u) & ~63u;
    if (r->nrc_capacity >= cap && r->nrc_rcap >= rcap) return;
    r->nrc_queries.alloc((size_t)cap * (11 + 4 + 1 + 1 + 4) * 4);
    r->nrc_records.alloc((size_t)rcap * (11 + 4 + 4 + 1 + 4) * 4);
    r->nrc_sacc.alloc((size_t)cap * 16);
    r->nrc_capacity = cap;
    r->nrc_rcap = rcap;
}

vt::NrcArgs nrc_view(vt_renderer *r, uint32_t *counters) {
    vt::NrcArgs a{};
    const size_t cap = r->nrc_capacity, rcap = r->nrc_rcap;
    char *q = r->nrc_queries.as<char>();
    a.q_in = reinterpret_cast<float *>(q);
    a.q_thr = reinterpret_cast<float4 *>(q + cap * 11 * 4);
    a.q_wl = reinterpret_cast<float *>(q + cap * 15 * 4);
    a.q_id = reinterpret_cast<uint32_t *>(q + cap * 16 * 4);
    a.q_out = reinterpret_cast<float *>(q + cap * 17 * 4);
    a.q_count = counters + 3;
    a.q_stride = (uint32_t)cap;
    char *rr = r->nrc_records.as<char>();
    a.r_in = reinterpret_cast<float *>(rr);
    a.r_thr = reinterpret_cast<float4 *>(rr + rcap * 11 * 4);
    a.r_snap = reinterpret_cast<float4 *>(rr + rcap * 15 * 4);
    a.r_id = reinterpret_cast<uint32_t *>(rr + rcap * 19 * 4);
    a.r_target = reinterpret_cast<float *>(rr + rcap * 20 * 4);
    a.r_count = counters + 4;
    a.r_cap = (uint32_t)rcap;
    a.sacc = r->nrc_sacc.as<float4>();
    a.unit = 1.0f;
    return a;
}

void normalize3(float *v) {
    float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) v[k] /= l;
}

}  // namespace

namespace vt {
void camera_frame(const float eye[3], const float target[3], float fov_degrees, uint32_t width, uint32_t height,
                  FrameArgs &f) {
    f.width = width;
    f.height = height;
    f.npix = width * height;
    float fw[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
    normalize3(fw);
    float rt[3] = {fw[1] * 0.0f - fw[2] * 1.0f, fw[2] * 0.0f - fw[0] * 0.0f, fw[0] * 1.0f - fw[1] * 0.0f};
    normalize3(rt);
    float up[3] = {rt[1] * fw[2] - rt[2] * fw[1], rt[2] * fw[0] - rt[0] * fw[2], rt[0] * fw[1] - rt[1] * fw[0]};
    for (int k = 0; k < 3; ++k) {
        f.eye[k] = eye[k];
        f.forward[k] = fw[k];
        f.right[k] = rt[k];
        f.up[k] = up[k];
    }
    f.fov_scale = std::tan((fov_degrees * 0.5f) * 3.14159265358979323846f / 180.0f);
    f.aspect = (float)width / (float)height;
}
}  // namespace vt

extern "C" {

vt_status vt_scene_set_materials(vt_scene *scene, const vt_material *materials, uint32_t count) {
    return vt::guarded_call([&]() -> vt_status {
        if (!scene || (count && !materials)) return fail(VT_ERROR_INVALID_ARGUMENT, "scene and materials must not be NULL");
        check(hipSetDevice(scene->ctx->device), "hipSetDevice");
        if (!scene->spectral) {
            scene->spectral = std::make_unique<vt::SpectralTables>();
            vt::build_spectral_tables(*scene->spectral);
            const vt::SpectralTables &t = *scene->spectral;
            scene->icdf.upload(t.icdf, sizeof t.icdf);
            scene->cmf.upload(t.cmf, sizeof t.cmf);
            scene->basis.upload(t.basis, sizeof t.basis);
        }
        const vt::SpectralTables &t = *scene->spectral;
        // Slot records (render_kernels.h); an emitter's selection weight is its luminance
        // (main.swift:9301-9309).
        std::vector<float> rec((size_t)std::max<uint32_t>(count, 1) * vt::kMaterialFloats, 0.0f);
        std::vector<double> weight(count, 0.0);
        bool layered = false, delta = false, lean = false;
        for (uint32_t k = 0; k < count; ++k) {
            const vt_material &m = materials[k];
            if (m.type > 4) return fail(VT_ERROR_INVALID_ARGUMENT, "material type must be 0..4");
            for (int c = 0; c < 3; ++c)
                if (!std::isfinite(m.albedo[c]) || !std::isfinite(m.emission[c]) || m.emission[c] < 0.0f)
                    return fail(VT_ERROR_INVALID_ARGUMENT, "material colours must be finite, emission >= 0");
            if (!(m.roughness >= 0.0f && m.roughness <= 1.0f) || !(m.ior >= 1.0f && m.ior <= 4.0f) ||
                !(m.metalness >= 0.0f && m.metalness <= 1.0f) || !(m.transmission >= 0.0f && m.transmission <= 1.0f) ||
                !(m.coat >= 0.0f && m.coat <= 1.0f) || !(m.anisotropy >= 0.0f && m.anisotropy <= 1.0f) ||
                !(m.fuzz >= 0.0f && m.fuzz <= 1.0f))
                return fail(VT_ERROR_INVALID_ARGUMENT,
                            "roughness, metalness, transmission, coat, anisotropy, fuzz must be in [0, 1], ior in [1, 4]");
            if (m.type == 2 && !(m.ior > 1.0f)) return fail(VT_ERROR_INVALID_ARGUMENT, "a dielectric needs ior > 1");
            float *r = rec.data() + vt::kMaterialFloats * (size_t)k;
            uint32_t type = m.type;
            std::memcpy(&r[0], &type, 4);
            float alb[3];
            for (int c = 0; c < 3; ++c) alb[c] = std::min(std::max(m.albedo[c], 0.0f), 1.0f);
            r[1] = alb[0]; r[2] = alb[1]; r[3] = alb[2];
            vt::reflectance_lagrange(t, alb, r + 4);
            r[8] = m.emission[0]; r[9] = m.emission[1]; r[10] = m.emission[2];
            if (m.emission[0] > 0 || m.emission[1] > 0 || m.emission[2] > 0)
                weight[k] = std::max(1e-8, 0.2126 * m.emission[0] + 0.7152 * m.emission[1] + 0.0722 * m.emission[2]);
            r[11] = (float)weight[k];
            r[12] = m.roughness; r[13] = m.ior; r[14] = m.metalness; r[15] = m.transmission;
            r[16] = m.coat;
            r[17] = m.anisotropy;
            r[18] = m.fuzz;
            if (m.type == 1 || m.type == 4) layered = true;
            if (m.type == 2 || (m.type == 1 && m.roughness < 0.02f)) delta = true;
            if (m.type == 0 || m.type == 2 || (m.type == 1 && m.roughness < 0.02f)) lean = true;
        }
        if (layered) vt::ensure_energy_tables(scene->ctx);
        // Power-sampling emitter list: every face of an emitting slot, weighted by world area x
        // weight, in document order; CDF normalised with the last entry exactly 1 (main.swift:9312-9346).
        std::vector<uint32_t> ids;
        std::vector<double> power;
        const size_t nf = scene->host_faces.size() / 4;
        for (size_t f = 0; f < nf; ++f) {
            uint32_t slot = scene->host_faces[4 * f + 3];
            if (slot >= count || weight[slot] <= 0.0) continue;
            const float *a = &scene->host_vertices[8 * (size_t)scene->host_faces[4 * f]];
            const float *b = &scene->host_vertices[8 * (size_t)scene->host_faces[4 * f + 1]];
            const float *c = &scene->host_vertices[8 * (size_t)scene->host_faces[4 * f + 2]];
            double e1[3], e2[3];
            for (int q = 0; q < 3; ++q) { e1[q] = (double)(b[q] - a[q]); e2[q] = (double)(c[q] - a[q]); }
            double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
            ids.push_back((uint32_t)f);
            power.push_back(0.5 * std::sqrt(cx * cx + cy * cy + cz * cz) * (double)(float)weight[slot]);
        }
        // Light BVH over the same list (MaterialLibrary.lightTreeList, main.swift:9374-9401): kept when
        // there are two or more emitters, more than one node, and a walkable depth.
        scene->tree_node_count = 0;
        if (scene->light_tree && ids.size() > 1) {
            std::vector<vt::LightEmitter> em(ids.size());
            for (size_t k = 0; k < ids.size(); ++k) {
                const uint32_t *fc = &scene->host_faces[4 * (size_t)ids[k]];
                for (int q = 0; q < 3; ++q) {
                    em[k].a[q] = scene->host_vertices[8 * (size_t)fc[0] + q];
                    em[k].b[q] = scene->host_vertices[8 * (size_t)fc[1] + q];
                    em[k].c[q] = scene->host_vertices[8 * (size_t)fc[2] + q];
                }
                em[k].power = power[k];
                em[k].cos_emission = 0.0;
            }
            vt::LightTreeBuilt tree = vt::build_light_tree(em);
            if (tree.node_count > 1 && tree.depth <= vt::kLightTreeMaxDepth && tree.node_count < vt::kLightTreeMaxNodes) {
                std::vector<uint32_t> leaves(3 * ids.size()), pos(ids.size());
                for (size_t leaf = 0; leaf < tree.order.size(); ++leaf) {
                    uint32_t input = tree.order[leaf];
                    pos[input] = (uint32_t)leaf;
                    float pw = (float)power[input];
                    uint32_t bits;
                    std::memcpy(&bits, &pw, 4);
                    leaves[3 * leaf] = ids[input];
                    leaves[3 * leaf + 1] = bits;
                    leaves[3 * leaf + 2] = tree.leaf_node[leaf];
                }
                scene->tree_nodes.upload(tree.words.data(), tree.words.size() * 4);
                scene->tree_leaves.upload(leaves.data(), leaves.size() * 4);
                scene->tree_ids.upload(ids.data(), ids.size() * 4);
                scene->tree_pos.upload(pos.data(), pos.size() * 4);
                scene->tree_node_count = tree.node_count;
                for (int k = 0; k < 3; ++k) {
                    scene->tree_lo[k] = tree.grid_lo[k];
                    scene->tree_scale[k] = (tree.grid_hi[k] - tree.grid_lo[k]) / 65535.0f;
                }
                scene->tree_dec.alloc((size_t)tree.node_count * 4 * 16);
                check(vt::launch_decode_light_tree(scene->render_view(), scene->tree_dec.as<float4>(), nullptr),
                      "decode light tree");
                check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            }
        }
        double total = 0;
        std::vector<float> cdf(ids.size());
        for (double p : power) total += p;
        if (!(total > 0) || !std::isfinite(total)) { ids.clear(); cdf.clear(); total = 0; }
        double run = 0;
        for (size_t k = 0; k < cdf.size(); ++k) {
            run += power[k];
            cdf[k] = k + 1 == cdf.size() ? 1.0f : (float)(run / total);
        }
        scene->materials.upload(rec.data(), rec.size() * 4);
        scene->emitter_face.upload(ids.data(), ids.size() * 4);
        scene->emitter_cdf.upload(cdf.data(), cdf.size() * 4);
        scene->material_count = count;
        scene->emitter_count = (uint32_t)ids.size();
        scene->emitter_total = (float)total;
        scene->shaded = true;
        scene->has_delta = delta;
        scene->has_layered = layered;
        scene->mixed = layered && lean;
        scene->generation = vt_scene::next_generation();
        return VT_OK;
    });
}

vt_status vt_scene_set_surface_settings(vt_scene *scene, const vt_surface_settings *settings, uint32_t slot_count,
                                        const vt_emission *emissions, uint32_t emission_count) {
    return vt::guarded_call([&]() -> vt_status {
        if (!scene || (slot_count && !settings) || (emission_count && !emissions))
            return fail(VT_ERROR_INVALID_ARGUMENT, "scene, settings and emissions must not be NULL");
        uint32_t slots = slot_count;
        for (uint32_t k = 0; k < emission_count; ++k) {
            if (emissions[k].slot >= (1u << 24)) return fail(VT_ERROR_INVALID_ARGUMENT, "emission slot out of range");
            slots = std::max(slots, emissions[k].slot + 1);
        }
        std::vector<vt_material> mats(slots);
        for (uint32_t k = 0; k < slots; ++k) {
            vt_material &m = mats[k];
            m = vt_material{};
            m.type = 0;  // the imported-mesh default (main.swift:2009)
            m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.7f;
            m.roughness = 0.5f;
            m.ior = 1.5f;
            m.specular_weight = 1.0f;
            if (k >= slot_count) continue;
            const vt_surface_settings &s = settings[k];
            if (s.map_mask != 0) return fail(VT_ERROR_UNSUPPORTED, "texture maps are outside this path (map_mask must be 0)");
            if (!s.enabled) continue;
            m.type = 4;
            for (int c = 0; c < 3; ++c) m.albedo[c] = s.color[c];
            m.roughness = s.surface[0];
            m.metalness = s.surface[1];
            // weights are fractions: clamped as the reference clamps its graph inputs (main.swift:2354-2355)
            m.coat = std::min(std::max(s.surface[2], 0.0f), 1.0f);
            m.anisotropy = std::min(std::max(s.surface[3], 0.0f), 1.0f);
            m.fuzz = std::min(std::max(s.detail[0], 0.0f), 1.0f);
            m.ior = s.detail[1];
            m.transmission = s.detail[2];
        }
        for (uint32_t k = 0; k < emission_count; ++k) {
            const vt_emission &e = emissions[k];
            if (!(e.rgb[0] > 0.0f || e.rgb[1] > 0.0f || e.rgb[2] > 0.0f)) continue;
            vt_material &m = mats[e.slot];
            m.type = 3;
            for (int c = 0; c < 3; ++c) m.emission[c] = e.rgb[c];
        }
        return vt_scene_set_materials(scene, mats.data(), slots);
    });
}

static const vt::SpectralTables &host_tables() {
    static vt::SpectralTables *t = [] {
        auto *p = new vt::SpectralTables();
        vt::build_spectral_tables(*p);
        return p;
    }();
    return *t;
}

vt_status vt_spectral_tables(float *icdf, float *cmf, float *basis) {
    return vt::guarded_call([&]() -> vt_status {
        const vt::SpectralTables &t = host_tables();
        if (icdf) std::memcpy(icdf, t.icdf, sizeof t.icdf);
        if (cmf) std::memcpy(cmf, t.cmf, sizeof t.cmf);
        if (basis) std::memcpy(basis, t.basis, sizeof t.basis);
        return VT_OK;
    });
}

vt_status vt_spectral_reflectance(const float rgb[3], float out[4]) {
    return vt::guarded_call([&]() -> vt_status {
        if (!rgb || !out) return fail(VT_ERROR_INVALID_ARGUMENT, "rgb and out must not be NULL");
        vt::reflectance_lagrange(host_tables(), rgb, out);
        return VT_OK;
    });
}

vt_status vt_uplift_grid_build(vt_context *ctx, float *coarse, uint8_t *flags, uint32_t capacity, uint32_t *cells,
                               float *refined, uint32_t *count) {
    return vt::guarded_call([&]() -> vt_status {
        if (!ctx) return fail(VT_ERROR_INVALID_ARGUMENT, "context must not be NULL");
        if (capacity && (!cells || !refined))
            return fail(VT_ERROR_INVALID_ARGUMENT, "cells and refined must not be NULL when capacity > 0");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        constexpr size_t kS = vt::kSpectralSamples, kNodes = 86, kCells = 85;
        std::vector<double> w(3 * kS), c1(kS), c2(kS);
        std::vector<float> wf(3 * kS), c1f(kS), c2f(kS), nodes(kNodes);
        vt::uplift_solve_tables(host_tables(), w.data(), c1.data(), c2.data(), nodes.data());
        for (size_t i = 0; i < 3 * kS; ++i) wf[i] = (float)w[i];
        for (size_t i = 0; i < kS; ++i) {
            c1f[i] = (float)c1[i];
            c2f[i] = (float)c2[i];
        }
        vt::DeviceBuffer dw, dc1, dc2, dwf, dc1f, dc2f, dnodes, dgrid, dflags, dcells, drefined;
        dw.upload(w.data(), w.size() * 8);
        dc1.upload(c1.data(), kS * 8);
        dc2.upload(c2.data(), kS * 8);
        dwf.upload(wf.data(), wf.size() * 4);
        dc1f.upload(c1f.data(), kS * 4);
        dc2f.upload(c2f.data(), kS * 4);
        dnodes.upload(nodes.data(), kNodes * 4);
        const vt::UpliftTables tb{dw.as<const double>(), dc1.as<const double>(), dc2.as<const double>(),
                                  dwf.as<const float>(), dc1f.as<const float>(), dc2f.as<const float>()};
        const size_t ncells = kCells * kCells * kCells;
        dgrid.alloc(kNodes * kNodes * kNodes * 16);
        dflags.alloc(ncells);
        check(hipMemset(dflags.ptr, 0, ncells), "hipMemset");
        check(vt::launch_uplift_coarse(tb, dgrid.as<float4>(), nullptr), "uplift coarse nodes");
        check(vt::launch_uplift_flags(tb, dgrid.as<const float4>(), dnodes.as<const float>(), dflags.as<uint8_t>(), nullptr),
              "uplift refinement flags");
        std::vector<uint8_t> hflags(ncells);
        check(hipMemcpy(hflags.data(), dflags.ptr, ncells, hipMemcpyDeviceToHost), "hipMemcpy");
        std::vector<uint32_t> list;
        for (size_t c = 0; c < ncells; ++c)
            if (hflags[c]) list.push_back((uint32_t)c);
        const uint32_t n = (uint32_t)list.size();
        if (n) {
            dcells.upload(list.data(), list.size() * 4);
            drefined.alloc((size_t)n * 64 * 16);
            check(vt::launch_uplift_refine(tb, dgrid.as<const float4>(), dnodes.as<const float>(), dcells.as<const uint32_t>(),
                                           n, drefined.as<float4>(), nullptr),
                  "uplift refined cells");
        }
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (coarse) {
            std::vector<float> g(kNodes * kNodes * kNodes * 4);
            check(hipMemcpy(g.data(), dgrid.ptr, g.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
            for (size_t i = 0; i < kNodes * kNodes * kNodes; ++i)
                for (int k = 0; k < 3; ++k) coarse[3 * i + k] = g[4 * i + k];
        }
        if (flags) std::memcpy(flags, hflags.data(), ncells);
        const uint32_t keep = std::min(capacity, n);
        if (keep) {
            std::vector<float> r((size_t)keep * 64 * 4);
            check(hipMemcpy(r.data(), drefined.ptr, r.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
            std::memcpy(cells, list.data(), keep * 4);
            for (size_t i = 0; i < (size_t)keep * 64; ++i)
                for (int k = 0; k < 3; ++k) refined[3 * i + k] = r[4 * i + k];
        }
        if (count) *count = n;
        return VT_OK;
    });
}

vt_status vt_energy_tables(vt_context *ctx, float *metal, float *luts) {
    return vt::guarded_call([&]() -> vt_status {
        if (!ctx) return fail(VT_ERROR_INVALID_ARGUMENT, "context must not be NULL");
        vt::ensure_energy_tables(ctx);
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (metal) check(hipMemcpy(metal, ctx->energy_metal.ptr, ctx->energy_metal.bytes, hipMemcpyDeviceToHost), "hipMemcpy");
        if (luts) check(hipMemcpy(luts, ctx->energy_luts.ptr, (size_t)vt::kLtcOffset * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        return VT_OK;
    });
}

vt_status vt_scene_set_light_tree(vt_scene *scene, int enabled) {
    if (!scene) return fail(VT_ERROR_INVALID_ARGUMENT, "scene must not be NULL");
    scene->light_tree = enabled != 0;
    scene->generation = vt_scene::next_generation();
    return VT_OK;
}

vt_status vt_light_tree_build_host(const float *tris, const double *power, uint32_t count, uint32_t **words,
                                   uint32_t *node_count, uint32_t **order, uint32_t **leaf_node, uint32_t *depth,
                                   float grid_lo[3], float grid_hi[3]) {
    return vt::guarded_call([&]() -> vt_status {
        if ((count && (!tris || !power)) || !words || !node_count || !order || !leaf_node || !depth || !grid_lo || !grid_hi)
            return fail(VT_ERROR_INVALID_ARGUMENT, "arguments must not be NULL");
        std::vector<vt::LightEmitter> em(count);
        for (uint32_t k = 0; k < count; ++k) {
            std::memcpy(em[k].a, tris + 9 * (size_t)k, 12);
            std::memcpy(em[k].b, tris + 9 * (size_t)k + 3, 12);
            std::memcpy(em[k].c, tris + 9 * (size_t)k + 6, 12);
            em[k].power = power[k];
            em[k].cos_emission = 0.0;
        }
        vt::LightTreeBuilt t = vt::build_light_tree(em);
        auto copy = [](const std::vector<uint32_t> &v) {
            uint32_t *p = (uint32_t *)std::malloc(std::max<size_t>(1, v.size()) * 4);
            if (!p) throw std::bad_alloc();
            if (!v.empty()) std::memcpy(p, v.data(), v.size() * 4);
            return p;
        };
        *words = copy(t.words);
        *order = copy(t.order);
        *leaf_node = copy(t.leaf_node);
        *node_count = t.node_count;
        *depth = t.depth;
        for (int k = 0; k < 3; ++k) {
            grid_lo[k] = t.grid_lo[k];
            grid_hi[k] = t.grid_hi[k];
        }
        return VT_OK;
    });
}

vt_status vt_renderer_create(vt_context *ctx, vt_renderer **out) {
    return vt::guarded_call([&]() -> vt_status {
        if (!ctx || !out) return fail(VT_ERROR_INVALID_ARGUMENT, "context and out must not be NULL");
        auto r = std::make_unique<vt_renderer>();
        r->ctx = ctx;
        r->counters.alloc(16 * sizeof(uint32_t));
        check(hipHostMalloc((void **)&r->host_count, sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
        *out = r.release();
        return VT_OK;
    });
}

void vt_renderer_destroy(vt_renderer *r) { delete r; }

vt_status vt_reset_accumulation(vt_renderer *r) {
    if (!r) return fail(VT_ERROR_INVALID_ARGUMENT, "renderer must not be NULL");
    r->frame_index = 0;
    r->restir_history = 0;
    return VT_OK;
}

uint32_t vt_accum_frames(const vt_renderer *r) { return r ? r->frame_index : 0; }

vt_status vt_renderer_set_cache(vt_renderer *r, vt_nrc *cache) {
    if (!r) return fail(VT_ERROR_INVALID_ARGUMENT, "renderer must not be NULL");
    r->cache = cache;
    return VT_OK;
}

vt_status vt_renderer_cache_records(vt_renderer *r, float *inputs, float *targets, uint32_t max_records,
                                    uint32_t *count) {
    return vt::guarded_call([&]() -> vt_status {
        if (!r || !count) return fail(VT_ERROR_INVALID_ARGUMENT, "renderer and count must not be NULL");
        *count = 0;
        if (r->nrc_rcap == 0 || !r->counters.ptr) return VT_OK;
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        uint32_t made = 0;
        check(hipMemcpy(&made, r->counters.as<uint32_t>() + 4, 4, hipMemcpyDeviceToHost), "hipMemcpy");
        *count = made;
        const uint32_t n = std::min({made, r->nrc_rcap, max_records});
        const vt::NrcArgs a = nrc_view(r, r->counters.as<uint32_t>());
        if (inputs && n) {
            std::vector<float> soa((size_t)11 * n);
            for (int k = 0; k < 11; ++k)
                check(hipMemcpy(soa.data() + (size_t)k * n, a.r_in + (size_t)k * a.r_cap, (size_t)n * 4, hipMemcpyDeviceToHost),
                      "hipMemcpy");
            for (uint32_t q = 0; q < n; ++q)
                for (int k = 0; k < 11; ++k) inputs[(size_t)q * 11 + k] = soa[(size_t)k * n + q];
        }
        if (targets && n) check(hipMemcpy(targets, a.r_target, (size_t)n * 4 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        return VT_OK;
    });
}

static vt_status render_frame(vt_renderer *r, vt_scene *scene, const vt_frame_params *p, uint32_t samples,
                              void *stream);

namespace {
struct FrameDone {
    vt_frame_fn fn;
    void *user;
    uint32_t frames;
};
void frame_done(void *payload) {
    FrameDone *d = static_cast<FrameDone *>(payload);
    d->fn(d->user, d->frames);
    delete d;
}
}  // namespace

vt_status vt_renderer_set_callbacks(vt_renderer *r, vt_frame_fn on_frame, vt_error_fn on_error, void *user) {
    if (!r) return fail(VT_ERROR_INVALID_ARGUMENT, "renderer must not be NULL");
    r->on_frame = on_frame;
    r->on_error = on_error;
    r->callback_user = user;
    return VT_OK;
}

vt_status vt_render_frame(vt_renderer *r, vt_scene *scene, const vt_frame_params *p, uint32_t samples, void *stream) {
    vt_status st = render_frame(r, scene, p, samples, stream);
    if (st == VT_OK && r->on_frame && samples) {
        st = vt::guarded_call([&]() -> vt_status {
            auto *d = new FrameDone{r->on_frame, r->callback_user, r->frame_index};
            const hipError_t e = hipLaunchHostFunc((hipStream_t)stream, frame_done, d);
            if (e != hipSuccess) {
                delete d;
                return fail(VT_ERROR_DEVICE, std::string("hipLaunchHostFunc: ") + hipGetErrorString(e));
            }
            return VT_OK;
        });
    }
    if (st != VT_OK && r && r->on_error) r->on_error(r->callback_user, vt_last_error());
    return st;
}

static vt_status render_frame(vt_renderer *r, vt_scene *scene, const vt_frame_params *p, uint32_t samples,
                              void *stream) {
    return vt::guarded_call([&]() -> vt_status {
        if (!r || !scene || !p) return fail(VT_ERROR_INVALID_ARGUMENT, "renderer, scene and params must not be NULL");
        if (!scene->shaded) return fail(VT_ERROR_INVALID_ARGUMENT, "scene has no materials (vt_scene_set_materials)");
        if (p->width == 0 || p->height == 0 || (uint64_t)p->width * p->height > (1u << 26))
            return fail(VT_ERROR_INVALID_ARGUMENT, "render size out of range");
        if (p->max_depth < 1 || p->max_depth > 64) return fail(VT_ERROR_INVALID_ARGUMENT, "max_depth must be 1..64");
        if (p->sampling_mode > 3) return fail(VT_ERROR_INVALID_ARGUMENT, "sampling_mode must be 0..3");
        const uint32_t candidates = p->restir_candidates ? p->restir_candidates : 4u;
        if (candidates > 256) return fail(VT_ERROR_INVALID_ARGUMENT, "restir_candidates must be at most 256");
        if (!(p->fov_degrees > 0.0f && p->fov_degrees < 180.0f)) return fail(VT_ERROR_INVALID_ARGUMENT, "fov must be in (0, 180)");
        if (p->sample_index == 0) return fail(VT_ERROR_INVALID_ARGUMENT, "sample_index starts at 1");
        if (p->material_sort > 1) return fail(VT_ERROR_INVALID_ARGUMENT, "material_sort must be 0 or 1");
        if (p->sampler > 1 || p->z_temporal > 3)
            return fail(VT_ERROR_INVALID_ARGUMENT, "sampler must be 0 or 1 and z_temporal 0..3");
        if (samples == 0) return VT_OK;
        check(hipSetDevice(r->ctx->device), "hipSetDevice");
        hipStream_t s = (hipStream_t)stream;
        const uint32_t npix = p->width * p->height;
        if (p->width != r->width || p->height != r->height) {
            r->accum.alloc((size_t)npix * 16);
            check(hipMemsetAsync(r->accum.ptr, 0, r->accum.bytes, s), "hipMemsetAsync");
            r->width = p->width;
            r->height = p->height;
            r->frame_index = 0;
            r->restir_history = 0;
            r->restir[0].reset();
            r->restir[1].reset();
            r->gi_res[0].reset();
            r->gi_res[1].reset();
            r->gi_work.reset();
            r->gi_history = false;
        }
        if (p->indirect_reuse > 2) return fail(VT_ERROR_INVALID_ARGUMENT, "indirect_reuse must be 0, 1 or 2");
        if (!(std::isfinite(p->environment_strength) && p->environment_strength >= 0.0f) ||
            !std::isfinite(p->environment_rotation))
            return fail(VT_ERROR_INVALID_ARGUMENT, "environment strength must be finite and >= 0, rotation finite");
        if (p->environment_strength > 0.0f && !scene->env_image.ptr)
            return fail(VT_ERROR_UNSUPPORTED,
                        "environment_strength needs an environment map (vt_scene_set_environment); the procedural sky is not built");
        const bool restir = p->sampling_mode == 0;
        if (restir && !r->restir[0].ptr) {
            r->restir[0].alloc((size_t)72 * npix);
            r->restir[1].alloc((size_t)72 * npix);
            r->restir_history = 0;
        }
        if (!restir || r->restir_generation != scene->generation) r->restir_history = 0;
        r->restir_generation = scene->generation;
        // Paths per wavefront: whole images, up to ~4M paths; ReSTIR frames run one sample at a time
        // (each reads the previous sample's reservoirs).
        const uint32_t per = restir ? 1u : std::max<uint32_t>(1, std::min<uint32_t>(samples, (1u << 22) / npix));
        ensure_capacity(r, per * npix);

        vt::FrameArgs f{};
        vt::camera_frame(p->eye, p->target, p->fov_degrees, p->width, p->height, f);
        f.scattering_limit = std::max(0, (int)p->max_depth - 1);
        f.sampling_mode = p->sampling_mode;
        f.env_strength = p->environment_strength;
        f.env_rotation = p->environment_rotation;
        if (p->radiance_cache > 2) return fail(VT_ERROR_INVALID_ARGUMENT, "radiance_cache must be 0, 1 or 2");
        const bool use_cache = p->radiance_cache != 0, train_cache = p->radiance_cache == 1;
        if (use_cache && !r->cache) return fail(VT_ERROR_INVALID_ARGUMENT, "radiance_cache needs vt_renderer_set_cache");
        const uint32_t period = p->cache_train_period ? p->cache_train_period : 16u;
        const float lr = p->cache_learning_rate > 0.0f ? p->cache_learning_rate : 1e-3f;
        if (use_cache && (period < 2 || period > 4096)) return fail(VT_ERROR_INVALID_ARGUMENT, "cache_train_period must be 2..4096");
        f.nrc = use_cache ? 1u : 0u;
        f.train_period = period;
        const bool gi = restir && p->indirect_reuse == 0 && f.scattering_limit >= 1 && !use_cache;
        if (gi && !r->gi_res[0].ptr) {
            r->gi_res[0].alloc((size_t)56 * npix);
            r->gi_res[1].alloc((size_t)56 * npix);
            r->gi_work.alloc(kGiWorkBytes * npix);
            r->gi_history = false;
        }
        if (use_cache) {
            if (!scene->box_valid) {
                float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
                for (size_t v = 0; v < scene->host_vertices.size() / 8; ++v)
                    for (int k = 0; k < 3; ++k) {
                        lo[k] = std::min(lo[k], scene->host_vertices[8 * v + k]);
                        hi[k] = std::max(hi[k], scene->host_vertices[8 * v + k]);
                    }
                for (int k = 0; k < 3; ++k) {
                    scene->box_lo[k] = lo[k];
                    scene->box_hi[k] = hi[k];
                }
                scene->box_valid = true;
            }
            for (int k = 0; k < 3; ++k) {
                f.box_lo[k] = scene->box_lo[k];
                const float ext = scene->box_hi[k] - scene->box_lo[k];
                f.box_scale[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
            }
            ensure_nrc_buffers(r, period);
        }

        const vt::RenderScene sc = scene->render_view();
        uint32_t *cnt = r->counters.as<uint32_t>();
        const vt::ShadowQueue sh = shadow_view(r);
        char *hb = r->hits.as<char>();
        const size_t cap = r->capacity;
        vt::HitArgs hits{reinterpret_cast<float *>(hb), reinterpret_cast<float *>(hb + 4 * cap),
                         reinterpret_cast<float *>(hb + 8 * cap), reinterpret_cast<int32_t *>(hb + 12 * cap), nullptr};
        for (uint32_t done = 0; done < samples; done += per) {
            f.samples = std::min(per, samples - done);
            f.sample_index0 = p->sample_index + done;
            f.sampler = p->sampler;
            f.z_temporal = p->z_temporal;
            f.frame0 = r->frame_index;
            const uint32_t n = f.samples * npix;
            check(vt::launch_reset_counters(cnt, 16, s), "reset");
            int cur = 0;
            check(vt::launch_generate(f, r->q[0].view(r->capacity), r->radiance.as<float>(), cnt + 0, s), "generate");
            vt::NrcArgs na{};
            if (use_cache) {
                na = nrc_view(r, cnt);
                double mean = 0.0;  // of the D65 column of the emission basis (float, as the device reads it)
                for (int i = 0; i < vt::kSpectralSamples; ++i) mean += scene->spectral->basis[i][0][3];
                mean /= (double)vt::kSpectralSamples;
                na.unit = (float)(1.0 / mean);
                check(vt::launch_zero_f4(na.sacc, n, s), "zero");
            }
            vt::RestirArgs ra{};
            if (restir) {
                ra.cur = restir_view(r->restir[r->restir_cur], npix);
                ra.hist = restir_view(r->restir[1 - r->restir_cur], npix);
                ra.candidates = candidates;
                ra.temporal = r->restir_history >= 1 ? 1u : 0u;
                std::memcpy(ra.prev_eye, r->prev_camera, 3 * sizeof(float));
                std::memcpy(ra.prev_forward, r->prev_camera + 3, 3 * sizeof(float));
                std::memcpy(ra.prev_right, r->prev_camera + 6, 3 * sizeof(float));
                std::memcpy(ra.prev_up, r->prev_camera + 9, 3 * sizeof(float));
                ra.prev_fov_scale = r->prev_camera[12];
                ra.prev_aspect = r->prev_camera[13];
                if (gi) {
                    ra.gi = 1u;
                    ra.gcur = gi_view(r->gi_res[r->restir_cur], npix);
                    ra.ghist = gi_view(r->gi_res[1 - r->restir_cur], npix);
                    ra.gw = gi_work_view(r->gi_work, npix);
                    ra.gi_temporal = ra.temporal && r->gi_history ? 1u : 0u;
                }
            }
            // Vertices: the scattering budget bounds non-delta events only; with delta materials in the
            // scene, keep extending (checking the live count every 8 vertices) until no path is left.
            const int base = f.scattering_limit + 1 + (f.sampling_mode == 3u ? 1 : 0);
            const int cap = scene->has_delta ? 255 : base;
            for (int b = 0; b < cap; ++b) {
                if (b >= base && (b - base) % 8 == 0) {
                    check(hipMemcpyAsync(r->host_count, cnt + cur, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
                    check(hipStreamSynchronize(s), "hipStreamSynchronize");
                    if (*r->host_count == 0) break;
                }
                uint32_t *cin = cnt + cur, *cout = cnt + (1 - cur), *csh = cnt + 2;
                check(vt::launch_reset_counters(cout, 1, s), "reset");
                check(vt::launch_reset_counters(csh, 1, s), "reset");
                const vt::PathQueue qin = r->q[cur].view(r->capacity), qout = r->q[1 - cur].view(r->capacity);
                vt::RayArgs rays{qin.ox, qin.oy, qin.oz, qin.dx, qin.dy, qin.dz, qin.tmin, qin.tmax};
                check(vt::launch_trace(scene->view, rays, hits, nullptr, n, false, nullptr, s, cin), "extend");
                vt::Counters c{cin, cout, csh};
                if (restir && b == 0) {
                    check(vt::launch_restir_initial(sc, f, qin, hits, ra, s), "restir pass 1");
                    if (gi) {
                        // ReSTIR GI's x1 -> x2 rays, the light samples at x2 and their shadow rays, then
                        // the reservoirs.
                        const vt::GiWork &w = ra.gw;
                        check(vt::launch_restir_gi_rays(sc, f, qin, hits, ra, s), "restir gi directions");
                        const vt::RayArgs grays{w.ox, w.oy, w.oz, w.dx, w.dy, w.dz, w.tmin, w.tmax};
                        check(vt::launch_trace(scene->view, grays, vt::HitArgs{w.t, w.b1, w.b2, w.face, nullptr}, nullptr,
                                               npix, false, nullptr, s), "restir gi rays");
                        check(vt::launch_restir_gi_secondary(sc, f, ra, s), "restir gi secondary");
                        const vt::RayArgs vis{w.sox, w.soy, w.soz, w.sdx, w.sdy, w.sdz, w.stmin, w.stmax};
                        check(vt::launch_trace(scene->view, vis, vt::HitArgs{}, w.occluded, npix, true, nullptr, s),
                              "restir gi shadow");
                        check(vt::launch_restir_gi_finish(sc, f, qin, hits, ra, s), "restir gi reservoirs");
                    }
                    check(vt::launch_shade_restir(sc, f, qin, hits, qout, sh, c, n, r->radiance.as<float>(), ra, na, s),
                          "shade");
                } else if (p->material_sort && scene->mixed) {
                    // Material sorting (classes counted in counters 5 and 6).
                    check(vt::launch_reset_counters(cnt + 5, 2, s), "reset");
                    check(vt::launch_shade_sorted(sc, f, qin, hits, qout, sh, c, n, r->radiance.as<float>(), na,
                                                  r->perm.as<uint32_t>(), cnt + 5, s), "shade");
                } else {
                    check(vt::launch_shade(sc, f, qin, hits, qout, sh, c, n, r->radiance.as<float>(), na, s), "shade");
                }
                vt::RayArgs srays{sh.ox, sh.oy, sh.oz, sh.dx, sh.dy, sh.dz, sh.tmin, sh.tmax};
                check(vt::launch_trace(scene->view, srays, vt::HitArgs{}, r->occluded.as<uint8_t>(), n, true, nullptr, s, csh),
                      "shadow");
                check(vt::launch_resolve(f, sh, r->occluded.as<uint8_t>(), csh, n, r->radiance.as<float>(), na, s), "resolve");
                if (restir && b == 0 && gi) {
                    // The spatial reuse's x1 -> x2 connections.
                    const vt::GiWork &w = ra.gw;
                    const vt::RayArgs vis{w.sox, w.soy, w.soz, w.sdx, w.sdy, w.sdz, w.stmin, w.stmax};
                    check(vt::launch_trace(scene->view, vis, vt::HitArgs{}, w.occluded, npix, true, nullptr, s),
                          "restir gi connections");
                    check(vt::launch_restir_gi_resolve(f, ra, r->radiance.as<float>(), s), "restir gi resolve");
                }
                cur = 1 - cur;
            }
            if (use_cache) {
                // The cache answers the paths that ended at its vertices, then learns from this
                // wavefront's training records: 4 Adam steps over equal slices of them (at most 16384
                // records a step), the slices sized on the device from the record count.
                const vt::NrcDevice nd = r->cache_device();
                check(vt::launch_nrc_infer(nd, na.q_in, n, na.q_out, s, na.q_count, na.q_stride), "cache infer");
                check(vt::launch_cache_resolve(sc, na, n, r->radiance.as<float>(), s), "cache resolve");
                check(vt::launch_build_targets(na, s), "cache targets");
                for (uint32_t part = 0; train_cache && part < 4u; ++part) {
                    const uint32_t batch = std::min<uint32_t>(16384u, na.r_cap);
                    const vt::NrcSlice slice{na.r_count, na.r_cap, 4u, part};
                    check(vt::launch_nrc_backward(nd, na.r_in, na.r_target, batch, 0.0f, nullptr, s, slice, na.r_cap),
                          "cache backward");
                    check(vt::launch_nrc_adam(nd, r->cache_next_step(), lr, s, batch, slice), "cache adam");
                }
            }
            check(vt::launch_accumulate(f, r->radiance.as<float>(), r->accum.as<float>(), r->frame_index + 1, s),
                  "accumulate");
            r->frame_index += f.samples;
            if (restir) {
                r->restir_cur = 1 - r->restir_cur;
                r->restir_history = std::min<uint32_t>(r->restir_history, 1u << 24) + 1u;
                r->gi_history = gi;
                std::memcpy(r->prev_camera, f.eye, 3 * sizeof(float));
                std::memcpy(r->prev_camera + 3, f.forward, 3 * sizeof(float));
                std::memcpy(r->prev_camera + 6, f.right, 3 * sizeof(float));
                std::memcpy(r->prev_camera + 9, f.up, 3 * sizeof(float));
                r->prev_camera[12] = f.fov_scale;
                r->prev_camera[13] = f.aspect;
            }
        }
        return VT_OK;
    });
}

vt_status vt_accum_readback(vt_renderer *r, float *rgba, size_t bytes) {
    return vt::guarded_call([&]() -> vt_status {
        if (!r || !rgba) return fail(VT_ERROR_INVALID_ARGUMENT, "renderer and output must not be NULL");
        const size_t need = (size_t)r->width * r->height * 16;
        if (bytes < need) return fail(VT_ERROR_INVALID_ARGUMENT, "output buffer too small");
        if (need == 0) return VT_OK;
        check(hipSetDevice(r->ctx->device), "hipSetDevice");
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check(hipMemcpy(rgba, r->accum.ptr, need, hipMemcpyDeviceToHost), "hipMemcpy");
        return VT_OK;
    });
}

}  // extern "C"
