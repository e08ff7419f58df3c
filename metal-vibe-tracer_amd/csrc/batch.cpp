// batch.cpp -- batched rendering (vt_batch_*): many images of many scenes advanced together by one
// wavefront of persistent kernels, the whole sample (every chunk of every job, every vertex) captured
// once in a hipGraph and replayed per sample; and the autotuner that picks the wavefront capacity,
// the persistent grid, the traversal work-group size and the BVH leaf size by timing replays.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh8.h"
#include "render_kernels.h"
#include "scene.h"
#include "trace_kernels.h"
#include "vt.h"

using vt::check;
using vt::DeviceBuffer;
using vt::fail;

namespace vt {
void rebuild_bvh(vt_scene *s, const BuildOptions &opt);
}

namespace {

constexpr int kQueueArrays = 8 + 12 + 1;       // rays, path state, job
constexpr int kShadowArrays = 8 + 1 + 3 + 4 + 1;  // rays, id, rgb, lane contributions (float4), job

}  // namespace

struct vt_batch {
    vt_context *ctx = nullptr;
    std::vector<vt_batch_job> jobs;
    std::vector<vt_scene *> scenes;       // distinct scenes, in first-use order
    std::vector<uint32_t> job_scene, first;
    uint32_t max_depth = 6, sampling_mode = 1;
    uint32_t total = 0;                   // pixels of one sample over all jobs
    uint32_t frames = 0;                  // samples accumulated (host mirror of the device counter)
    vt_batch_tuning tuning{};
    int cus = 256;
    // Device tables.
    DeviceBuffer scene_table, frame_table, job_scene_d, first_d, job_trace, frame_count;
    std::vector<vt::FrameArgs> frame_host;
    // Wavefront storage, carved for `tuning.capacity` paths out of buffers sized for `allocated`.
    DeviceBuffer queues, shadow, hits, occluded, radiance, counters, accum;
    uint32_t allocated = 0;
    // The captured sample.
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool layered = false;
    uint32_t tree_depth = 1;
    // vt_scene::generation of each scene when its views went into the tables: a rebuilt hierarchy or
    // new materials replace the buffers the captured graph reads, so a change re-uploads and recaptures.
    std::vector<uint64_t> scene_generation;

    ~vt_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
    vt::BatchTables tables() const {
        vt::BatchTables t{};
        t.scenes = scene_table.as<const vt::RenderScene>();
        t.frames = frame_table.as<vt::FrameArgs>();
        t.job_scene = job_scene_d.as<const uint32_t>();
        t.first = first_d.as<const uint32_t>();
        t.jobs = (uint32_t)jobs.size();
        t.frame_count = frame_count.as<uint32_t>();
        return t;
    }
};

namespace {

// Uploads the per-scene and per-job tables (scene views change when a hierarchy is rebuilt).
void upload_tables(vt_batch *b) {
    std::vector<vt::RenderScene> rs;
    b->layered = false;
    b->tree_depth = 1;
    b->scene_generation.clear();
    for (vt_scene *s : b->scenes) {
        b->scene_generation.push_back(s->generation);
        if (s->has_layered) vt::ensure_energy_tables(b->ctx);
        rs.push_back(s->render_view());
        b->layered = b->layered || s->has_layered;
        b->tree_depth = std::max(b->tree_depth, s->view.depth);
    }
    b->scene_table.upload(rs.data(), rs.size() * sizeof(vt::RenderScene));
    std::vector<vt::TraceScene> jt;
    for (size_t j = 0; j < b->jobs.size(); ++j) jt.push_back(b->scenes[b->job_scene[j]]->view);
    b->job_trace.upload(jt.data(), jt.size() * sizeof(vt::TraceScene));
}

// Per-job frames at sample index 1 and an empty accumulation.
void reset_frames(vt_batch *b) {
    for (auto &f : b->frame_host) f.sample_index0 = 1;
    check(hipMemcpy(b->frame_table.ptr, b->frame_host.data(), b->frame_host.size() * sizeof(vt::FrameArgs),
                    hipMemcpyHostToDevice), "hipMemcpy");
    const uint32_t zero = 0;
    check(hipMemcpy(b->frame_count.ptr, &zero, 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemset(b->accum.ptr, 0, b->accum.bytes), "hipMemset");
    b->frames = 0;
}

void ensure_storage(vt_batch *b, uint32_t cap) {
    if (b->allocated >= cap) return;
    b->queues.alloc((size_t)2 * kQueueArrays * cap * 4);
    b->shadow.alloc((size_t)kShadowArrays * cap * 4);
    b->hits.alloc((size_t)4 * cap * 4);
    b->occluded.alloc(cap);
    b->radiance.alloc((size_t)3 * cap * 4);
    b->allocated = cap;
}

vt::PathQueue queue_view(const vt_batch *b, int which, uint32_t cap) {
    char *base = b->queues.as<char>() + (size_t)which * kQueueArrays * cap * 4;
    auto f = [&](int k) { return reinterpret_cast<float *>(base + (size_t)k * cap * 4); };
    auto u = [&](int k) { return reinterpret_cast<uint32_t *>(base + (size_t)k * cap * 4); };
    vt::PathQueue q;
    q.ox = f(0); q.oy = f(1); q.oz = f(2); q.dx = f(3); q.dy = f(4); q.dz = f(5); q.tmin = f(6); q.tmax = f(7);
    q.id = u(8); q.seed = u(9); q.wl = f(10);
    q.t0 = f(11); q.t1 = f(12); q.t2 = f(13); q.t3 = f(14);
    q.state = u(15); q.pdf = f(16); q.px = f(17); q.py = f(18); q.pz = f(19);
    q.job = u(20);
    return q;
}

vt::ShadowQueue shadow_view(const vt_batch *b, uint32_t cap) {
    char *base = b->shadow.as<char>();
    auto f = [&](int k) { return reinterpret_cast<float *>(base + (size_t)k * cap * 4); };
    vt::ShadowQueue s;
    s.ox = f(0); s.oy = f(1); s.oz = f(2); s.dx = f(3); s.dy = f(4); s.dz = f(5); s.tmin = f(6); s.tmax = f(7);
    s.id = reinterpret_cast<uint32_t *>(f(8));
    s.r = f(9); s.g = f(10); s.b = f(11);
    s.s = reinterpret_cast<float4 *>(f(12));  // 4 arrays' room, 16-byte aligned (cap is a multiple of 64)
    s.job = reinterpret_cast<uint32_t *>(f(16));
    return s;
}

uint32_t persistent_grid(const vt_batch *b, uint32_t block) {
    // waves per SIMD x 4 SIMDs per CU, in blocks of `block` lanes.
    const uint32_t waves = b->tuning.grid_waves * 4u * (uint32_t)b->cus;
    return std::max<uint32_t>(1, waves / (block / 64u));
}

// Records one sample of every job: for each chunk of at most `capacity` pixels, the camera paths,
// max_depth vertices (extend, shade, shadow, resolve) and the running mean; then the step to the
// next sample.  Every count lives on the device, so the recorded launches replay unchanged.
void record_sample(vt_batch *b, hipStream_t s) {
    const uint32_t cap = b->tuning.capacity;
    const vt::BatchTables bt = b->tables();
    const uint32_t grid = persistent_grid(b, 256), tgrid = persistent_grid(b, b->tuning.trace_block);
    uint32_t *cnt = b->counters.as<uint32_t>();
    char *hb = b->hits.as<char>();
    const vt::HitArgs hits{reinterpret_cast<float *>(hb), reinterpret_cast<float *>(hb + 4 * (size_t)cap),
                           reinterpret_cast<float *>(hb + 8 * (size_t)cap), reinterpret_cast<int32_t *>(hb + 12 * (size_t)cap),
                           nullptr};
    const vt::ShadowQueue sh = shadow_view(b, cap);
    const vt::TraceScene *jt = b->job_trace.as<const vt::TraceScene>();
    const int vertices = std::max(0, (int)b->max_depth - 1) + 1 + (b->sampling_mode == 3u ? 1 : 0);
    for (uint32_t base = 0; base < b->total; base += cap) {
        const uint32_t n = std::min(cap, b->total - base);
        check(vt::launch_reset_counters(cnt, 16, s), "reset");
        check(vt::launch_batch_generate(bt, base, n, queue_view(b, 0, cap), b->radiance.as<float>(), cnt, grid, s), "generate");
        int cur = 0;
        for (int v = 0; v < vertices; ++v) {
            uint32_t *cin = cnt + cur, *cout = cnt + (1 - cur), *csh = cnt + 2;
            check(vt::launch_reset_counters(cout, 1, s), "reset");
            check(vt::launch_reset_counters(csh, 1, s), "reset");
            const vt::PathQueue qin = queue_view(b, cur, cap), qout = queue_view(b, 1 - cur, cap);
            const vt::RayArgs rays{qin.ox, qin.oy, qin.oz, qin.dx, qin.dy, qin.dz, qin.tmin, qin.tmax};
            check(vt::launch_trace_batch(jt, qin.job, rays, hits, nullptr, n, false, cin, b->tree_depth,
                                         b->tuning.trace_block, tgrid, s), "extend");
            check(vt::launch_batch_shade(bt, qin, hits, qout, sh, vt::Counters{cin, cout, csh}, n, b->radiance.as<float>(),
                                         b->layered, grid, s), "shade");
            const vt::RayArgs srays{sh.ox, sh.oy, sh.oz, sh.dx, sh.dy, sh.dz, sh.tmin, sh.tmax};
            check(vt::launch_trace_batch(jt, sh.job, srays, vt::HitArgs{}, b->occluded.as<uint8_t>(), n, true, csh,
                                         b->tree_depth, b->tuning.trace_block, tgrid, s), "shadow");
            check(vt::launch_batch_resolve(sh, b->occluded.as<uint8_t>(), csh, n, b->radiance.as<float>(), grid, s),
                  "resolve");
            cur = 1 - cur;
        }
        check(vt::launch_batch_accumulate(bt, base, n, b->radiance.as<float>(), b->accum.as<float>(), grid, s),
              "accumulate");
    }
    check(vt::launch_batch_advance(bt, s), "advance");
}

void capture(vt_batch *b) {
    if (b->exec) check(hipGraphExecDestroy(b->exec), "hipGraphExecDestroy");
    if (b->graph) check(hipGraphDestroy(b->graph), "hipGraphDestroy");
    b->exec = nullptr;
    b->graph = nullptr;
    ensure_storage(b, b->tuning.capacity);
    check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        record_sample(b, b->stream);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    check(hipStreamEndCapture(b->stream, &b->graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

void validate_tuning(const vt_batch *b, vt_batch_tuning &t) {
    if (t.capacity == 0 || t.capacity > b->total) t.capacity = b->total;
    t.capacity = (t.capacity + 63u) & ~63u;
    if (t.grid_waves == 0 || t.grid_waves > 32) throw std::invalid_argument("grid_waves must be 1..32");
    if (t.trace_block != 64 && t.trace_block != 128 && t.trace_block != 256 && t.trace_block != 512)
        throw std::invalid_argument("trace_block must be 64, 128, 256 or 512");
}

// Milliseconds per sample: one warm-up replay, then `reps` timed replays.
double time_replays(vt_batch *b, int reps) {
    check(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    check(hipEventRecord(e0, b->stream), "hipEventRecord");
    for (int k = 0; k < reps; ++k) check(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
    check(hipEventRecord(e1, b->stream), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
}

}  // namespace

extern "C" {

vt_status vt_batch_create(vt_context *ctx, const vt_batch_job *jobs, uint32_t count, uint32_t max_depth,
                          uint32_t sampling_mode, vt_batch **out) {
    return vt::guarded_call([&]() -> vt_status {
        if (!ctx || !out || (count && !jobs)) return fail(VT_ERROR_INVALID_ARGUMENT, "context, jobs and out must not be NULL");
        *out = nullptr;
        if (count == 0) return fail(VT_ERROR_INVALID_ARGUMENT, "a batch needs at least one job");
        if (max_depth < 1 || max_depth > 64) return fail(VT_ERROR_INVALID_ARGUMENT, "max_depth must be 1..64");
        if (sampling_mode < 1 || sampling_mode > 3)
            return fail(VT_ERROR_INVALID_ARGUMENT, "batched sampling_mode must be 1 (MIS), 2 (NEE) or 3 (BSDF)");
        auto b = std::make_unique<vt_batch>();
        b->ctx = ctx;
        b->max_depth = max_depth;
        b->sampling_mode = sampling_mode;
        check(hipSetDevice(ctx->device), "hipSetDevice");
        uint64_t total = 0;
        b->first.push_back(0);
        for (uint32_t j = 0; j < count; ++j) {
            const vt_batch_job &jb = jobs[j];
            if (!jb.scene || !jb.scene->shaded) return fail(VT_ERROR_INVALID_ARGUMENT, "every job needs a scene with materials");
            if (jb.scene->view.kind != 0u)
                return fail(VT_ERROR_UNSUPPORTED, "batched rendering traces 8-wide hierarchies (build option collapse 0 or 1)");
            if (jb.scene->has_delta)
                return fail(VT_ERROR_INVALID_ARGUMENT, "batched rendering bounds paths at max_depth vertices: scenes with "
                                                       "mirrors or smooth glass need vt_render_frame");
            if (jb.width == 0 || jb.height == 0 || (uint64_t)jb.width * jb.height > (1u << 26))
                return fail(VT_ERROR_INVALID_ARGUMENT, "job image size out of range");
            if (!(jb.fov_degrees > 0.0f && jb.fov_degrees < 180.0f)) return fail(VT_ERROR_INVALID_ARGUMENT, "fov must be in (0, 180)");
            auto it = std::find(b->scenes.begin(), b->scenes.end(), jb.scene);
            if (it == b->scenes.end()) {
                b->scenes.push_back(jb.scene);
                it = b->scenes.end() - 1;
            }
            b->job_scene.push_back((uint32_t)(it - b->scenes.begin()));
            vt::FrameArgs f{};
            vt::camera_frame(jb.eye, jb.target, jb.fov_degrees, jb.width, jb.height, f);
            f.samples = 1;
            f.sample_index0 = 1;
            f.scattering_limit = std::max(0, (int)max_depth - 1);
            f.sampling_mode = sampling_mode;
            b->frame_host.push_back(f);
            total += (uint64_t)jb.width * jb.height;
            if (total >= (1ull << 31)) return fail(VT_ERROR_INVALID_ARGUMENT, "a batch holds at most 2^31 pixels");
            b->first.push_back((uint32_t)total);
            b->jobs.push_back(jb);
        }
        b->total = (uint32_t)total;
        int dev = ctx->device;
        check(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        b->frame_table.alloc(b->frame_host.size() * sizeof(vt::FrameArgs));
        b->job_scene_d.upload(b->job_scene.data(), b->job_scene.size() * 4);
        b->first_d.upload(b->first.data(), b->first.size() * 4);
        b->frame_count.alloc(4);
        b->counters.alloc(64 * 4);
        b->accum.alloc((size_t)b->total * 16);
        upload_tables(b.get());
        reset_frames(b.get());
        check(hipStreamCreate(&b->stream), "hipStreamCreate");
        b->tuning.capacity = std::min<uint32_t>(b->total, 1u << 22);
        b->tuning.grid_waves = 4;
        b->tuning.trace_block = 128;
        b->tuning.leaf_size = b->scenes[0]->info.leaf_size;
        validate_tuning(b.get(), b->tuning);
        capture(b.get());
        *out = b.release();
        return VT_OK;
    });
}

void vt_batch_destroy(vt_batch *b) { delete b; }

vt_status vt_batch_get_tuning(const vt_batch *b, vt_batch_tuning *t) {
    if (!b || !t) return fail(VT_ERROR_INVALID_ARGUMENT, "batch and tuning must not be NULL");
    *t = b->tuning;
    return VT_OK;
}

vt_status vt_batch_set_tuning(vt_batch *b, const vt_batch_tuning *t) {
    return vt::guarded_call([&]() -> vt_status {
        if (!b || !t) return fail(VT_ERROR_INVALID_ARGUMENT, "batch and tuning must not be NULL");
        vt_batch_tuning n = *t;
        validate_tuning(b, n);
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        if (n.leaf_size && n.leaf_size != b->tuning.leaf_size) {
            vt::BuildOptions o;
            o.leaf_size = (int)n.leaf_size;
            for (vt_scene *s : b->scenes) vt::rebuild_bvh(s, o);
            upload_tables(b);
        }
        if (!n.leaf_size) n.leaf_size = b->tuning.leaf_size;
        b->tuning = n;
        capture(b);
        return VT_OK;
    });
}

vt_status vt_batch_render(vt_batch *b, uint32_t samples, void *stream) {
    return vt::guarded_call([&]() -> vt_status {
        if (!b) return fail(VT_ERROR_INVALID_ARGUMENT, "batch must not be NULL");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        // A scene rebuilt or given new materials since the capture: its old buffers are gone, and a
        // deeper tree needs a deeper traversal stack, so the tables and the graph are made again.
        bool stale = false;
        for (size_t k = 0; k < b->scenes.size(); ++k) stale = stale || b->scenes[k]->generation != b->scene_generation[k];
        for (vt_scene *s : b->scenes)
            if (s->view.kind != 0u)
                return fail(VT_ERROR_UNSUPPORTED, "a batched scene was rebuilt as the flat binary hierarchy (collapse 2); "
                                                  "batched rendering traces 8-wide hierarchies");
        if (stale) {
            check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
            b->tuning.leaf_size = b->scenes[0]->info.leaf_size;
            upload_tables(b);
            capture(b);
        }
        hipStream_t user = (hipStream_t)stream;
        hipEvent_t ev;
        check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        // Work the caller queued before this call comes first; the caller's stream waits for the samples.
        check(hipEventRecord(ev, user), "hipEventRecord");
        check(hipStreamWaitEvent(b->stream, ev, 0), "hipStreamWaitEvent");
        for (uint32_t k = 0; k < samples; ++k) check(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
        check(hipEventRecord(ev, b->stream), "hipEventRecord");
        check(hipStreamWaitEvent(user, ev, 0), "hipStreamWaitEvent");
        (void)hipEventDestroy(ev);
        b->frames += samples;
        return VT_OK;
    });
}

vt_status vt_batch_reset(vt_batch *b) {
    return vt::guarded_call([&]() -> vt_status {
        if (!b) return fail(VT_ERROR_INVALID_ARGUMENT, "batch must not be NULL");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        reset_frames(b);
        return VT_OK;
    });
}

uint32_t vt_batch_frames(const vt_batch *b) { return b ? b->frames : 0; }

vt_status vt_batch_readback(vt_batch *b, uint32_t job, float *rgba, size_t bytes) {
    return vt::guarded_call([&]() -> vt_status {
        if (!b || !rgba) return fail(VT_ERROR_INVALID_ARGUMENT, "batch and rgba must not be NULL");
        if (job >= b->jobs.size()) return fail(VT_ERROR_INVALID_ARGUMENT, "no such job");
        const size_t need = (size_t)(b->first[job + 1] - b->first[job]) * 16;
        if (bytes < need) return fail(VT_ERROR_INVALID_ARGUMENT, "rgba buffer too small");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        check(hipMemcpy(rgba, b->accum.as<char>() + (size_t)b->first[job] * 16, need, hipMemcpyDeviceToHost), "hipMemcpy");
        return VT_OK;
    });
}

vt_status vt_batch_autotune(vt_batch *b, uint32_t flags, vt_batch_tuning *best, double *best_ms) {
    return vt::guarded_call([&]() -> vt_status {
        if (!b) return fail(VT_ERROR_INVALID_ARGUMENT, "batch must not be NULL");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        // Coordinate descent from the current tuning: each parameter in turn over its candidates,
        // keeping the fastest (milliseconds per sample of the whole batch, 2 timed replays each).
        vt_batch_tuning cur = b->tuning;
        double cur_ms = time_replays(b, 2);
        auto trial = [&](vt_batch_tuning t) {
            validate_tuning(b, t);
            if (t.leaf_size != b->tuning.leaf_size) {
                vt::BuildOptions o;
                o.leaf_size = (int)t.leaf_size;
                for (vt_scene *s : b->scenes) vt::rebuild_bvh(s, o);
                upload_tables(b);
            }
            b->tuning = t;
            capture(b);
            const double ms = time_replays(b, 2);
            if (ms < cur_ms) {
                cur_ms = ms;
                cur = t;
            }
        };
        std::vector<uint32_t> caps;
        for (uint32_t c = 1u << 20; c < b->total; c <<= 1) caps.push_back(c);
        caps.push_back(b->total);
        for (uint32_t c : caps) {
            vt_batch_tuning t = cur;
            t.capacity = c;
            trial(t);
        }
        for (uint32_t w : {1u, 2u, 4u, 8u, 16u}) {
            vt_batch_tuning t = cur;
            t.grid_waves = w;
            trial(t);
        }
        for (uint32_t blk : {64u, 128u, 256u, 512u}) {
            vt_batch_tuning t = cur;
            t.trace_block = blk;
            trial(t);
        }
        if (flags & 1u) {
            for (uint32_t leaf : {1u, 2u, 3u}) {
                vt_batch_tuning t = cur;
                t.leaf_size = leaf;
                trial(t);
            }
        }
        // Settle on the best, with a fresh accumulation.
        if (cur.leaf_size != b->tuning.leaf_size) {
            vt::BuildOptions o;
            o.leaf_size = (int)cur.leaf_size;
            for (vt_scene *s : b->scenes) vt::rebuild_bvh(s, o);
            upload_tables(b);
        }
        b->tuning = cur;
        capture(b);
        check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        reset_frames(b);
        if (best) *best = cur;
        if (best_ms) *best_ms = cur_ms;
        return VT_OK;
    });
}

}  // extern "C"
