// render_kernels.h -- device data of the wavefront path tracer and its launch interface.
//
// The frame runs as a wavefront: generate (camera rays) -> [extend (closest hit) -> shade
// (emission with MIS, next-event sample, Russian roulette, BSDF sample, compaction of surviving
// paths) -> shadow (any hit) -> resolve (add unoccluded next-event contributions)] per bounce ->
// accumulate.  Every queue is structure-of-arrays and compacted with one atomic per wavefront.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "trace_kernels.h"

namespace vt {

// Material slot record: 5 x float4 per slot (indexed by the face tag).
//   [0] type (as float bits), albedo rgb
//   [1] reflectance Lagrange multipliers (L0, L1, L2, w): w = 1 MESE, w = 2 flat grey in x
//   [2] emission rgb, light-selection weight (luminance; 0 = not an emitter)
//   [3] roughness, ior, metalness, transmission (OpenPBR inputs)
//   [4] coat weight, specular anisotropy, fuzz weight, 1 unused
constexpr uint32_t kMaterialFloats = 20;
constexpr uint32_t kLtcOffset = 5 * 32768;  // floats before the LTC sheen table in the energy buffer
enum MaterialType : uint32_t { kDiffuse = 0, kGlossy = 1, kDielectric = 2, kEmissive = 3, kOpenPbr = 4 };

struct RenderScene {
    TraceScene trace;
    const float *vertices;     // 8 floats per vertex (document order)
    const uint32_t *faces;     // 4 u32 per face (document order)
    uint32_t face_count;
    const float *materials;    // kMaterialFloats per slot
    uint32_t material_count;
    const uint32_t *emitter_face;  // emitting faces (power sampling), ascending face index
    const float *emitter_cdf;      // cumulative area x weight / total; last = 1
    uint32_t emitter_count;
    float emitter_total;           // sum of area x weight
    const float *icdf;             // 1025 wavelength inverse-CDF nodes
    const float *cmf;              // 471 x (r, g, b, cos phase)
    const float *basis;            // 471 x 2 x 4 emission basis (w of the first: D65)
    // Light BVH (tree_node_count 0: power sampling through emitter_cdf).
    const uint32_t *tree_nodes;    // 8 words per node (Sources/LightTree.swift:192-205)
    const float4 *tree_dec;        // the same nodes decoded, 4 float4 each (launch_decode_light_tree)
    const uint32_t *tree_leaves;   // 3 words per leaf position: face, power bits, leaf node
    const uint32_t *tree_ids;      // emitting faces, ascending (emitter_count entries)
    const uint32_t *tree_pos;      // leaf position of tree_ids[i]
    uint32_t tree_node_count;
    float tree_lo[3], tree_scale[3];
    // Energy-compensation tables of the layered BSDF (129 x 129 + 129 ideal metal; 5 x 32768 others).
    const float *energy_metal;
    const float *energy_luts;  // 5 x 32768 dielectric tables, then the LTC sheen table at kLtcOffset
    uint32_t layered;  // some slot is GLOSSY or OPENPBR (the shading kernels then carry the layered BSDF)
    // Environment map (null: none): env_w x env_h rgba32f texels, row 0 at the zenith; its sampling
    // tables, rows[2 y + table] (marginal CDFs) and cols[(table env_h + y) env_w + x] (conditional
    // CDFs), table 0 MIS-compensated, 1 original.
    const float4 *env;
    const float *env_rows, *env_cols;
    uint32_t env_w, env_h;
};

// ReSTIR DI state per pixel, the reference's G-buffer and reservoir textures (main.swift:4842-4940):
// gpos = position, camera distance (w = -1: no hit; rgba32f), gnorm = normal, material type as four
// halves (rgba16f), then the reservoir: posdir = sample position + light id bits, emitpdf =
// emission rgb + solid-angle pdf, weights = weight sum, M, W, wavelength number (rgba32f each).
struct RestirBuffers {
    float4 *gpos;
    uint2 *gnorm;
    float4 *posdir, *emitpdf, *weights;
};
// ReSTIR GI reservoir per pixel (main.swift:4625, 4803-4811): pos = secondary point x2 + source pdf
// (area measure), nrm = its normal and a sample flag as four halves (rgba16f), rad = its secondary
// radiance at the sample's four wavelengths, weights = weight sum, M, W, wavelength number.
struct GiBuffers {
    float4 *pos;
    uint2 *nrm;
    float4 *rad, *weights;
};
// Work arrays of ReSTIR GI, per pixel: the x1 -> x2 rays (RayArgs layout) and their hits, the light
// sample drawn at x2 (direction, emission, solid-angle pdf), the stage reached (0 no GI ray, 1 ray
// traced, 2 light sample with a shadow ray, 3 light sample that sees no light), the BSDF pdf of the GI
// direction and the PCG state after the stage's draws.  The visibility rays (s*) and their results
// serve twice: the shadow rays of the x2 light samples in pass 1, then the x1 -> x2 connection rays of
// the spatial reuse in pass 2, whose unoccluded contribution (rgb) is added by the GI resolve.
struct GiWork {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
    float *t, *b1, *b2;
    int32_t *face;
    float *sox, *soy, *soz, *sdx, *sdy, *sdz, *stmin, *stmax;
    uint8_t *occluded;
    float4 *light_wi;    // wi, pdf
    float4 *light_le;    // emission rgb
    float4 *rgb;         // pass 2: the connection's contribution if x1 sees x2
    uint32_t *stage, *seed;
    float *bpdf;
};

struct RestirArgs {
    RestirBuffers cur, hist;  // written by this sample's pass 1 / the previous sample's
    GiBuffers gcur, ghist;    // the same for ReSTIR GI
    GiWork gw;
    uint32_t gi;              // ReSTIR GI at the first indirect diffuse vertex (IndirectReuse.restirGI)
    uint32_t gi_temporal;     // ghist holds the previous sample's GI reservoirs
    uint32_t candidates;      // RIS candidates per pixel (the reference uses 4)
    uint32_t temporal;        // history holds the previous sample of the same image
    // Previous sample's camera for backprojection (prevViewProj, main.swift:4828-4835).
    float prev_eye[3], prev_forward[3], prev_right[3], prev_up[3];
    float prev_fov_scale, prev_aspect;
};

struct FrameArgs {
    uint32_t width, height, npix;
    uint32_t samples;          // samples in this wavefront (path id = sample * npix + pixel)
    uint32_t sample_index0;    // Uniforms.sampleIndex of sample 0
    float eye[3], forward[3], right[3], up[3];
    float fov_scale, aspect;
    int32_t scattering_limit;  // max(0, depth - 1)
    uint32_t sampling_mode;    // 0 ReSTIR DI at the primary vertex, 1 MIS, 2 NEE only, 3 BSDF only
    // Learned radiance cache: 1 = paths end at their first vertex reached by a non-delta scatter and
    // take its outgoing radiance from the cache, except training paths (1 in train_period, by a hash
    // of pixel and sample), which are traced in full and leave one training record per such vertex.
    uint32_t nrc, train_period;
    // Sampler: 0 the PCG stream; 1 the Z++ sampler (pixel key from the accumulation's frame
    // frame0 + sample + 1 and sample index, temporal model z_temporal).
    uint32_t sampler, z_temporal, frame0;
    float box_lo[3], box_scale[3];  // cache inputs: (p - lo) * scale in [0, 1]
    // Uniforms.environment: the map's strength (> 0 with a map: environment lighting on) and rotation.
    float env_strength, env_rotation;
};

// Queues of the learned radiance cache (indexed by queue slot unless stated).
struct NrcArgs {
    float *q_in;         // [11][q_stride] queries: inputs of the vertex where a path ends
    float4 *q_thr;       // throughput reaching it, per lane
    float *q_wl;         // wavelength number
    uint32_t *q_id;      // path id
    uint32_t *q_count;
    uint32_t q_stride;   // query capacity
    float *q_out;        // [cap][4] cache outputs
    float4 *sacc;        // [paths] spectral radiance a training path has gathered so far (by path id)
    float *r_in;         // [11][r_cap] training records: inputs of the cache vertex of a training path
    float4 *r_thr;       // throughput reaching it
    float4 *r_snap;      // sacc of its path when it was reached
    uint32_t *r_id;
    uint32_t *r_count;
    uint32_t r_cap;
    float *r_target;     // [rcap][4] built after the wavefront: (sacc final - snap) / thr * unit
    // The network's unit: its outputs and targets are spectral radiance times `unit` (1 / the mean of
    // the luminance-normalised D65 over the tabulated wavelengths), so that scene radiances of order
    // one, for which the relative loss's 0.01 is made, stay of order one.
    float unit;
};

// One queue of paths: the ray arrays first (RayArgs layout), then the path state.
struct PathQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
    uint32_t *id;       // path id (sample * npix + pixel)
    uint32_t *seed;     // PCG state
    float *wl;          // wavelength number u
    float *t0, *t1, *t2, *t3;  // throughput per wavelength lane
    uint32_t *state;    // bounce | scattering depth << 8 | flags << 16
    float *pdf;         // BSDF pdf of the departure that produced the ray
    float *px, *py, *pz;  // position the ray left
    uint32_t *job = nullptr;  // batched rendering: the path's job (null for a single image)
};

struct ShadowQueue {
    float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
    uint32_t *id;
    float *r, *g, *b;   // contribution if unoccluded
    float4 *s;          // the same contribution per wavelength lane (for cache training)
    uint32_t *job = nullptr;  // batched rendering: the ray's job
};

struct Counters {
    uint32_t *in, *out, *shadow;
};

// Batched rendering: many images (jobs) of many scenes in one wavefront.  A sample's work items are
// the jobs' pixels back to back: item g is pixel g - first[j] of job j, first[j] <= g < first[j + 1],
// and also the pixel's place in the accumulation buffer.
struct BatchTables {
    const RenderScene *scenes;    // per scene
    FrameArgs *frames;            // per job: its camera and image; sample_index0 advances on the device
    const uint32_t *job_scene;    // per job: scene index
    const uint32_t *first;        // per job, plus the total at [jobs]
    uint32_t jobs;
    uint32_t *frame_count;        // samples accumulated (device)
};

hipError_t launch_generate(const FrameArgs &f, const PathQueue &q, float *radiance, uint32_t *count, hipStream_t s);
hipError_t launch_shade(const RenderScene &sc, const FrameArgs &f, const PathQueue &in, const HitArgs &hits,
                        const PathQueue &out, const ShadowQueue &sh, const Counters &c, uint32_t capacity,
                        float *radiance, const NrcArgs &nrc, hipStream_t s);
// launch_shade with material sorting: paths are split by whether their hit needs the layered BSDF
// (perm: capacity slots, classes: 2 zeroed device counters) and each class is shaded by its own kernel.
hipError_t launch_shade_sorted(const RenderScene &sc, const FrameArgs &f, const PathQueue &in, const HitArgs &hits,
                               const PathQueue &out, const ShadowQueue &sh, const Counters &c, uint32_t capacity,
                               float *radiance, const NrcArgs &nrc, uint32_t *perm, uint32_t *classes, hipStream_t s);
// Pass 1 of ReSTIR DI (restir_temporal_kernel): G-buffer, RIS candidates, temporal merge, encode,
// for the wavefront's single sample (q holds the camera paths in pixel order, hits their hits).
hipError_t launch_restir_initial(const RenderScene &sc, const FrameArgs &f, const PathQueue &q, const HitArgs &hits,
                                 const RestirArgs &ra, hipStream_t s);
// ReSTIR GI after pass 1: the x1 -> x2 rays (ra.gw); once they are traced, the light sample at each
// diffuse x2 and its shadow ray (stage 2); then, with the shadow rays traced, the initial reservoir,
// the temporal merge and the encode into ra.gcur.
hipError_t launch_restir_gi_rays(const RenderScene &sc, const FrameArgs &f, const PathQueue &q, const HitArgs &hits,
                                 const RestirArgs &ra, hipStream_t s);
hipError_t launch_restir_gi_secondary(const RenderScene &sc, const FrameArgs &f, const RestirArgs &ra, hipStream_t s);
hipError_t launch_restir_gi_finish(const RenderScene &sc, const FrameArgs &f, const PathQueue &q, const HitArgs &hits,
                                   const RestirArgs &ra, hipStream_t s);
// After pass 2's connection rays (ra.gw s*) are traced into ra.gw.occluded: radiance[pixel] += the
// contribution of each connection that is not occluded.
hipError_t launch_restir_gi_resolve(const FrameArgs &f, const RestirArgs &ra, float *radiance, hipStream_t s);
// launch_shade with ReSTIR DI at the primary vertex (sampling mode 0): spatial reuse of ra.cur,
// shading with one visibility ray, then the path suffix.
hipError_t launch_shade_restir(const RenderScene &sc, const FrameArgs &f, const PathQueue &in, const HitArgs &hits,
                               const PathQueue &out, const ShadowQueue &sh, const Counters &c, uint32_t capacity,
                               float *radiance, const RestirArgs &ra, const NrcArgs &nrc, hipStream_t s);
hipError_t launch_resolve(const FrameArgs &f, const ShadowQueue &sh, const uint8_t *occluded, const uint32_t *count,
                          uint32_t capacity, float *radiance, const NrcArgs &nrc, hipStream_t s);
// Adds each cache query's answer: radiance[id] += rgb(thr * max(y, 0)) at the path's wavelengths.
hipError_t launch_cache_resolve(const RenderScene &sc, const NrcArgs &nrc, uint32_t capacity, float *radiance,
                                hipStream_t s);
// Training targets of the wavefront's records: (sacc[id] - snap) / thr per lane (0 where thr is 0).
hipError_t launch_build_targets(const NrcArgs &nrc, hipStream_t s);
hipError_t launch_zero_f4(float4 *p, uint32_t n, hipStream_t s);
// The light tree's nodes decoded for the walk (5 float4 per node) into out.
hipError_t launch_decode_light_tree(const RenderScene &sc, float4 *out, hipStream_t s);
// Energy-table bake: ideal-metal complements (129 x 129, Hammersley 16384 per node).
hipError_t launch_bake_metal(float *out, hipStream_t s);
hipError_t launch_reset_counters(uint32_t *counters, uint32_t n, hipStream_t s);
hipError_t launch_accumulate(const FrameArgs &f, const float *radiance, float *accum, uint32_t frame_index,
                             hipStream_t s);

// Batched rendering with persistent grids of `grid` blocks of 256 (MIS, no ReSTIR or cache): the
// camera paths of work items [chunk_base, chunk_base + chunk_n) (one sample of each), a vertex's
// shading with per-path scene and camera, the shadow resolve, the running mean of the chunk's pixels,
// and the step to the next sample.
hipError_t launch_batch_generate(const BatchTables &bt, uint32_t chunk_base, uint32_t chunk_n, const PathQueue &q,
                                 float *radiance, uint32_t *count, uint32_t grid, hipStream_t s);
hipError_t launch_batch_shade(const BatchTables &bt, const PathQueue &in, const HitArgs &hits, const PathQueue &out,
                              const ShadowQueue &sh, const Counters &c, uint32_t capacity, float *radiance, bool layered,
                              uint32_t grid, hipStream_t s);
hipError_t launch_batch_resolve(const ShadowQueue &sh, const uint8_t *occluded, const uint32_t *count, uint32_t capacity,
                                float *radiance, uint32_t grid, hipStream_t s);
hipError_t launch_batch_accumulate(const BatchTables &bt, uint32_t chunk_base, uint32_t chunk_n, const float *radiance,
                                   float *accum, uint32_t grid, hipStream_t s);
hipError_t launch_batch_advance(const BatchTables &bt, hipStream_t s);

}  // namespace vt
