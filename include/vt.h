/*
 * vt.h -- C ABI of the MI355X (gfx950) ray-traversal and path-tracing core.
 *
 * This is the drop-in boundary for metal-vibe-tracer's scene-upload / trace / render-frame path.
 * Handles are opaque; the caller owns its host arrays (copied at the call); the library owns
 * device memory.  Ray and hit buffers passed to the trace calls are DEVICE pointers (HIP device
 * memory on the context's GPU), laid out as structure-of-arrays so every lane of a wavefront
 * reads consecutive words.  Errors: a vt_status plus a thread-local message (vt_last_error).
 * Threading: one context per thread; all work of a context is serialised on the stream given
 * to each call (NULL = the default stream).
 */
#ifndef VT_H
#define VT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vt_status {
    VT_OK = 0,
    VT_ERROR_INVALID_ARGUMENT = 1,
    /* The reference throws "The imported meshes exceed the safe GPU memory budget."
     * (Sources/Acceleration.swift:728-733) for uploads over its budget. */
    VT_ERROR_OUT_OF_MEMORY = 2,
    VT_ERROR_DEVICE = 3,
    VT_ERROR_UNSUPPORTED = 4
} vt_status;

typedef struct vt_context vt_context;
typedef struct vt_scene vt_scene;

typedef struct vt_context_desc {
    int32_t device;         /* HIP device ordinal */
    uint64_t memory_budget; /* bytes; 0 = max(256 MiB, 1/4 of device memory), as
                               MaterialLibrary.meshBudget (Sources/Acceleration.swift:610-612) */
} vt_context_desc;

/* Replaces PathTracerRenderer.init(device:shaders:sharing:) (main.swift:10487) and the
 * MaterialLibrary's device binding: one HIP device per context. */
vt_status vt_context_create(const vt_context_desc *desc, vt_context **out);
void vt_context_destroy(vt_context *ctx);
/* Message of the last failed call on this thread ("" if none); replaces the NSError built by
 * MaterialLibrary.error (main.swift:9128). */
const char *vt_last_error(void);
/* Waits for the context's device and reports faults its kernels raised since the last check (then
 * clears them): VT_ERROR_DEVICE when a traversal's node-group stack filled, so that some rays ended
 * before their traversal did.  The stack is sized to the tree's depth, which a traversal never
 * exceeds; only vt_scene_set_stack_cap can make it happen.  No counterpart in the reference, whose
 * fixed 64-entry stack fails silently (main.swift:1788-1812). */
vt_status vt_context_check(vt_context *ctx);

typedef struct vt_build_options {
    uint32_t leaf_size; /* triangles per leaf, 1..3 (0: 3) */
    uint32_t bins;      /* SAH bins per axis, 4..64 (0: 16, Sources/Acceleration.swift:109) */
    uint32_t max_depth; /* tree depth limit before a median rebuild, <= 24 (0: 24) */
    uint32_t collapse;  /* 0: SAH-optimal collapse of a binary SAH tree into 8-wide nodes; 1: greedy
                           widening, splitting the largest child (Sources/Acceleration.swift:247-284);
                           2: no 8-wide tree -- the reference's flat binary median hierarchy
                           (OBJMesh.hierarchy, Sources/StudioModel.swift:701-760: 48-byte MeshNode
                           records, at most 4 triangles per leaf, leaf_size / bins / max_depth
                           ignored), walked as the reference's software trace_scene walks it
                           (main.swift:2058-2080).  Hits are the same; the traversal counters,
                           batched rendering and instancing need 0 or 1 (VT_ERROR_UNSUPPORTED) */
} vt_build_options;

/* Replaces MaterialLibrary.setMesh(_:) for an indexed mesh (Sources/StudioModel.swift:1480-1504,
 * Sources/Acceleration.swift:652-782 publishTwoLevel): vertices are MeshVertex records (8 floats:
 * position, normal, uv; Sources/MeshGeometry.swift:4-13), faces are MeshFace records (4 u32:
 * three vertex indices and the subset/slot tag).  Builds the 8-wide BVH on the host and uploads
 * it.  opts may be NULL.  Hit face IDs returned by the trace calls are DOCUMENT face indices
 * (position in `faces`). */
vt_status vt_scene_upload_indexed(vt_context *ctx, const float *vertices, uint32_t vertex_count,
                                  const uint32_t *faces, uint32_t face_count,
                                  const vt_build_options *opts, vt_scene **out);
/* Replaces MaterialLibrary.setMesh(_ triangles: [MeshTriangle]) (Sources/StudioModel.swift:1480-1494):
 * `count` 128-byte MeshTriangle records (32 floats: a, b, c, na, nb, nc as float4, uvab, uvc with
 * uvc.z the material slot) welded into indexed geometry as MeshGeometry(triangles:) does
 * (vt_mesh_from_triangles), then published as vt_scene_upload_indexed.  Face IDs are the triangles'
 * positions in the array. */
vt_status vt_scene_upload_triangles(vt_context *ctx, const float *triangles, uint32_t count,
                                    const vt_build_options *opts, vt_scene **out);
void vt_scene_destroy(vt_scene *scene);

typedef struct vt_scene_info {
    uint32_t face_count, vertex_count, node_count, depth;
    uint32_t leaf_size, width, node_bytes, tri_bytes; /* per record */
    uint64_t device_bytes;                            /* nodes + triangle records */
    float root_max;     /* max |coordinate| of the root bounds (the traversal enlargement) */
    float build_seconds;
} vt_scene_info;
vt_status vt_scene_get_info(const vt_scene *scene, vt_scene_info *info);

/* Device-resident rays, one array per component (n entries each). */
typedef struct vt_ray_soa {
    const float *ox, *oy, *oz;
    const float *dx, *dy, *dz;
    const float *tmin, *tmax;
} vt_ray_soa;

/* Device-resident closest hits.  face = -1 on a miss (then t = tmax, b1 = b2 = 0).  b1, b2
 * weight vertices b and c, as intersect_mesh_triangle (main.swift:1564-1582).  instance may be
 * NULL (single-instance scenes write 0). */
typedef struct vt_hit_soa {
    float *t, *b1, *b2;
    int32_t *face;
    uint32_t *instance;
} vt_hit_soa;

/* Replaces the software branch of mesh_closest / blas_closest (main.swift:1788-1812,
 * 1878-1900) over a ray buffer, as the reference's ray-query kernels call trace_scene
 * (tests/Fix_accel.swift:70-85).  Closest hit in [tmin, tmax); among hits at exactly equal t the
 * smallest document face index wins.  Asynchronous on `stream` (a hipStream_t, NULL = default). */
vt_status vt_trace_closest(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, const vt_hit_soa *hits,
                           void *stream);

/* Replaces mesh_occluded / blas_any (main.swift:1815-1838, 1968-1983): occluded[i] = 1 when any
 * triangle is hit in [tmin, tmax). */
vt_status vt_trace_any(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, uint8_t *occluded, void *stream);

/* Work counters of the same traversal (not timed): counters[0] = node records fetched,
 * counters[1] = triangle records tested, summed over the n rays; any != 0 counts the any-hit
 * traversal.  counters is a HOST array; the call synchronises the stream. */
vt_status vt_trace_stats(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, int any, uint64_t counters[2],
                         void *stream);
/* vt_trace_stats for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc); no counterpart in the reference.  It allocates nothing: the
 * caller passes the results (hits for the closest-hit count, occluded for any != 0) and `counters`,
 * 2 uint64 in DEVICE memory (node records fetched, triangle records tested), and the 64 partial
 * counter pairs live in a 1 KiB static device array the library holds on each GPU for the process,
 * cleared by a kernel at every call; its size does not depend on n, so no ray count is too large.
 * Asynchronous on `stream` (no synchronisation).  Calls that overlap on one GPU race on that array. */
vt_status vt_trace_stats_noalloc(vt_scene *scene, const vt_ray_soa *rays, uint32_t n, int any,
                                 const vt_hit_soa *hits, uint8_t *occluded, uint64_t *counters, void *stream);

/* ---------------------------------------------------------------- two-level scenes (instances) */

/* One instance of a published scene (the reference's MeshSceneLayout.Instance,
 * Sources/Acceleration.swift:367-376): `mesh` is its asset, whose hierarchy serves as the BLAS, and
 * world[] the rows of its object-to-world matrix (x = world[0..2] . p + world[3], y from world[4..7],
 * z from world[8..11]). */
typedef struct vt_instance {
    vt_scene *mesh;
    float world[12];
} vt_instance;
typedef struct vt_tlas vt_tlas;

/* Replaces MeshSceneLayout.addInstance + sceneBlob (Sources/Acceleration.swift:470-573): checks each
 * transform as addInstance does (VT_ERROR_INVALID_ARGUMENT with its messages: "Combined hierarchy
 * scale exceeds the supported range.", "Singular USD transform.", "Hierarchy transform exceeds the
 * supported scene extent."), builds the instances' world-to-object rows and world bounds, and an
 * 8-wide hierarchy over those bounds, one instance per leaf.  The meshes must outlive the vt_tlas;
 * a mesh rebuilt with vt_scene_rebuild is picked up by the next trace call (which then waits for
 * the device).  An instance of a mesh without faces is kept and never hit. */
vt_status vt_tlas_create(vt_context *ctx, const vt_instance *instances, uint32_t count, vt_tlas **out);
void vt_tlas_destroy(vt_tlas *tlas);

typedef struct vt_tlas_info {
    uint32_t instance_count, node_count, depth; /* the instance hierarchy */
    uint32_t blas_depth;                        /* deepest mesh hierarchy */
    uint64_t rendered_triangles;                /* sum of the instances' face counts */
    uint64_t device_bytes;                      /* instance nodes + instance records */
    float enlarge, reach; /* MeshSceneHeader lo.w and hi.w: per-ray box growth enlarge (|origin|max + reach) */
    float build_seconds;
} vt_tlas_info;
vt_status vt_tlas_get_info(const vt_tlas *tlas, vt_tlas_info *info);

/* Replaces the software branch of mesh_closest (main.swift:1843-1900): closest hit over the
 * instances in [tmin, tmax).  hits->face is the face's index in its mesh (document order),
 * hits->instance the instance's position in the list given to vt_tlas_create (0xffffffff on a
 * miss), b1 / b2 the barycentrics of the object-space hit.  Each instance traces its mesh with the
 * object-space ray origin' = W^-1 (origin), direction' = W^-1 (direction), not renormalised (main.swift:
 * 1634-1637), so t is the world ray's parameter.  Among hits at exactly equal t the smaller instance
 * index wins, then the smaller face index.  Asynchronous on `stream`. */
vt_status vt_tlas_trace_closest(vt_tlas *tlas, const vt_ray_soa *rays, uint32_t n, const vt_hit_soa *hits,
                                void *stream);
/* Replaces the software branch of mesh_occluded (main.swift:1902-1983, without the endpoint slack):
 * occluded[i] = 1 when any instance's mesh is hit in [tmin, tmax). */
vt_status vt_tlas_trace_any(vt_tlas *tlas, const vt_ray_soa *rays, uint32_t n, uint8_t *occluded, void *stream);

/* ---------------------------------------------------------------- materials and rendering */

typedef struct vt_renderer vt_renderer;
typedef struct vt_nrc vt_nrc;

/* One material slot, selected by the face tag (MeshFace.w).  Replaces the per-slot
 * SurfaceSettings / emissions of MaterialLibrary (main.swift:2253-2263, 8926, 9002). */
typedef struct vt_material {
    uint32_t type;          /* MaterialType (main.swift:316-322): 0 DIFFUSE (Lambert), 1 GLOSSY (metal:
                               delta mirror below roughness 0.02, else OpenPBR with metalness 1),
                               2 DIELECTRIC (delta glass, Schlick), 3 EMISSIVE, 4 OPENPBR */
    float albedo[3];        /* linear sRGB reflectance (base colour / metal f0 / glass tint), clamped to [0, 1] */
    float emission[3];      /* linear sRGB radiance under the D65 illuminant; any > 0 makes the slot an emitter */
    float roughness;        /* specular roughness (OpenPBR clamps it to >= 0.03) */
    float ior;              /* specular IOR (>= 1.01 for OpenPBR) */
    float metalness;        /* OPENPBR base_metalness in [0, 1] */
    float specular_weight;  /* MaterialX materials only; ignored (1 for GLOSSY / OPENPBR, main.swift:2525) */
    float transmission;     /* OPENPBR transmission_weight in [0, 1] */
    float coat;             /* GLOSSY / OPENPBR coat_weight in [0, 1] (Material.coat, main.swift:2494): a GGX
                               dielectric coat of roughness 0.12 and IOR 1.6 over the layered BSDF */
    float anisotropy;       /* GLOSSY / OPENPBR specular_roughness_anisotropy in [0, 1] (main.swift:2496): the
                               specular lobe stretched along the hit's tangent dP/du (from the vertices' uv) */
    float fuzz;             /* GLOSSY / OPENPBR fuzz_weight in [0, 1] (main.swift:2497): a white sheen layer of
                               roughness 0.5 (OpenPBR's fuzz lobe) over the coat and base */
    float reserved;
} vt_material;              /* 64 bytes */

/* Replaces MaterialLibrary's settings/emissions and its emitter list (main.swift:9296-9346):
 * emitting faces are chosen by area x luminance for next-event estimation.  Copies the materials;
 * builds the spectral tables (D65 emitters) and the material reflectance spectra. */
vt_status vt_scene_set_materials(vt_scene *scene, const vt_material *materials, uint32_t count);

/* The reference's own material records, for callers that hold them: SurfaceSettings (64 bytes,
 * main.swift:2253-2263; Swift 8911-8919) per slot and the emissions map (main.swift:9002). */
typedef struct vt_surface_settings {
    float color[4];          /* base colour, linear sRGB (w unused); default (1, 1, 1, 1) */
    float surface[4];        /* roughness, metalness, coat, anisotropy; default (0.3, 0, 0, 0) */
    float detail[4];         /* fuzz, IOR, transmission, UV repeat; default (0, 1.5, 0, 1) */
    uint32_t enabled;        /* 0: the slot keeps the imported-mesh default */
    uint32_t map_mask;       /* texture maps (bit per channel) */
    float normal_strength;
    uint32_t padding;
} vt_surface_settings;       /* 64 bytes */
typedef struct vt_emission {
    uint32_t slot;           /* material slot (face tag) */
    float rgb[3];            /* linear sRGB radiance under D65; any > 0 makes the slot's faces emitters */
} vt_emission;
/* Materials of an imported mesh from SurfaceSettings and emissions, as resolve_material and
 * trace_scene assign them (main.swift:2008-2011, 2390-2400): a slot with a positive emission is
 * EMISSIVE; else an enabled slot is OPENPBR with color.rgb, roughness surface.x, metalness surface.y,
 * coat surface.z, anisotropy surface.w, fuzz detail.x (the three clamped to [0, 1]), IOR detail.y and
 * transmission detail.z; else the imported-mesh default, DIFFUSE with albedo 0.7.  Slots beyond
 * slot_count (emitters named only in `emissions`) take the default.  Refused (VT_ERROR_UNSUPPORTED):
 * texture maps (map_mask != 0), which are outside this path.  Then as vt_scene_set_materials. */
vt_status vt_scene_set_surface_settings(vt_scene *scene, const vt_surface_settings *settings, uint32_t slot_count,
                                        const vt_emission *emissions, uint32_t emission_count);

/* Emitter selection for next-event estimation (MaterialLibrary.lightTree, main.swift:9006): 1 (default)
 * builds a light BVH over the emitting faces (Sources/LightTree.swift) when there are at least two and
 * the tree has more than one node; 0 keeps power sampling.  Takes effect at vt_scene_set_materials. */
vt_status vt_scene_set_light_tree(vt_scene *scene, int enabled);

/* Host build of the light BVH over `count` emitter triangles (9 floats each: a, b, c) with powers
 * (area x weight): *words (8 per node), *order (leaf position -> emitter), *leaf_node (leaf position ->
 * node) are malloc'ed (free with vt_host_free).  Mirrors LightTree.build (Sources/LightTree.swift:112). */
vt_status vt_light_tree_build_host(const float *tris, const double *power, uint32_t count, uint32_t **words,
                                   uint32_t *node_count, uint32_t **order, uint32_t **leaf_node, uint32_t *depth,
                                   float grid_lo[3], float grid_hi[3]);

/* Replaces MaterialLibrary.makeEnvironment + publishEnvironment / setEnvironment
 * (Sources/StudioModel.swift:1399-1470): the scene's environment map, `width` x `height` linear rgba32f
 * pixels in latitude-longitude layout (row 0 at the zenith, +y up; column u at azimuth
 * atan2(z, x) = 2 pi (u - 1/2)), copied to the device with its sampling tables (the reference's
 * environmentImportance, see vt_environment_tables).  The reference's checks and messages: at most
 * 16384 x 8192 and 33,554,432 pixels ("Could not decode the environment image (maximum 16384×8192)."),
 * finite pixels ("Environment pixels must be finite."), negative values clamped to 0, and the memory
 * budget (24 bytes per pixel + 8 per row: "Environment image exceeds the safe GPU memory budget.").
 * rgba NULL removes the map.  The map lights a frame whose environment_strength is positive. */
vt_status vt_scene_set_environment(vt_scene *scene, const float *rgba, uint32_t width, uint32_t height);

/* The environment's sampling tables on the host (MaterialLibrary.environmentImportance,
 * Sources/StudioModel.swift:1300-1380): rows (2 x height floats, rows[2 y + table]) holds each table's
 * marginal row CDF, cols (2 x height x width floats, cols[(table height + y) width + x]) the conditional
 * column CDFs; table 0 is the MIS-compensated table (max(0, L - mean L)), 1 the luminance itself, both
 * over the 3 x 3-dilated luminance times sin(theta). */
vt_status vt_environment_tables(const float *rgba, uint32_t width, uint32_t height, float *rows, float *cols);

/* Per-call frame parameters: the subset of Uniforms (main.swift:393-420) this path reads. */
typedef struct vt_frame_params {
    uint32_t width, height;
    float eye[3];            /* cameraPos.xyz */
    float target[3];         /* cameraTarget.xyz (up is +y) */
    float fov_degrees;       /* cameraPos.w, vertical */
    uint32_t max_depth;      /* cameraTarget.w: path vertices, 1 = direct light only */
    uint32_t sampling_mode;  /* samplingMode: 0 = ReSTIR DI at the primary vertex (RIS, temporal and
                                spatial reuse) + MIS suffix, 1 = MIS (NEE + BSDF, power heuristic),
                                2 = NEE only, 3 = BSDF sampling only */
    uint32_t sample_index;   /* sampleIndex of the first sample of this call (>= 1) */
    uint32_t restir_candidates; /* RIS light candidates per pixel in mode 0 (restir_di_initial's
                                   M = 4, main.swift:4649); 0 means 4, at most 256 */
    uint32_t radiance_cache;    /* 1: use the renderer's learned radiance cache (vt_renderer_set_cache):
                                   paths end at their first vertex after a non-delta scatter with the
                                   cache's answer, 1 in cache_train_period is traced in full and trains
                                   the cache on its suffixes (no counterpart in the reference);
                                   2: the same estimator with the cache frozen (no training) */
    uint32_t cache_train_period; /* 0 means 16 */
    float cache_learning_rate;   /* 0 means 1e-3 */
    uint32_t material_sort;      /* 1: at vertices past the first, split the paths by shading class (the
                                    layered BSDF or not) and shade each class with its own kernel; the
                                    image is the same.  0 (default) shades every path in one kernel,
                                    which measured faster on the configurations here */
    uint32_t sampler;            /* 0: every pixel draws from its PCG stream (the reference's PCG mode);
                                    1: the Z++ sampler (main.swift:122-305): the pixel's key and each
                                    draw's sampling event as in the reference's Z mode (ReSTIR's light
                                    candidates are candidates of one event; its resampling
                                    decisions stay on the PCG stream) */
    uint32_t z_temporal;         /* Z++ temporal model 0..3 of the pixel key (main.swift:497-506; the
                                    reference's default is 1) */
    uint32_t indirect_reuse;     /* mode 0, Uniforms.indirectReuse: 0 = ReSTIR GI at the first indirect
                                    diffuse vertex (IndirectReuse.restirGI, main.swift:8419-8434:
                                    initial x1 -> x2 -> light sample, temporal and spatial reuse by
                                    reconnection; the diffuse secondary vertex then takes no NEE);
                                    1 = no indirect reuse, the MIS path suffix of mode 1 (the
                                    reference's values 1 and 2, ReSTIR PT, are not built here).  GI
                                    is off with depth 1 and with the radiance cache, which answers
                                    that vertex's light itself */
    float environment_strength;  /* Uniforms.environment.x: the scene's environment map (set with
                                    vt_scene_set_environment) lights the frame when positive: camera
                                    misses see it, continuations that leave the scene gather it with
                                    MIS, and light samples choose it with probability 1/2 when the scene
                                    also has emitters (imported_light_probability); 0 = black */
    float environment_rotation;  /* Uniforms.environment.y: azimuthal rotation of the map, radians */
} vt_frame_params;

/* Replaces PathTracerRenderer (main.swift:9514) for this path: owns the accumulation buffer. */
vt_status vt_renderer_create(vt_context *ctx, vt_renderer **out);
void vt_renderer_destroy(vt_renderer *r);
/* Replaces renderFrame(output:samples:) (main.swift:10978): renders `samples` samples (sample
 * indices sample_index .. sample_index + samples - 1) and folds them into the running mean
 * accum = prev + (x - prev) / frameIndex (main.swift:7908-7913).  A size change resets the
 * accumulation.  In sampling mode 0 every sample is one ReSTIR frame (encodeSample,
 * main.swift:11214-11708): its reservoirs become the next sample's temporal history, across calls,
 * until a reset, a size change or a sample in another mode.  Asynchronous on `stream`. */
vt_status vt_render_frame(vt_renderer *r, vt_scene *scene, const vt_frame_params *params, uint32_t samples,
                          void *stream);
/* Completion callbacks, standing in for the renderer's onFrameUpdate(UInt32) and onError(String)
 * (main.swift:10292-10293, 11198-11202).  After each vt_render_frame call has queued its work,
 * on_frame(user, frames) is queued on the same stream (hipLaunchHostFunc): it runs on a runtime thread
 * once the call's samples are in the accumulation, with the accumulated sample count, and must not
 * call HIP or this library.  A call that fails calls on_error(user, message) on the calling thread
 * before it returns its status.  Either may be NULL; NULL, NULL detaches. */
typedef void (*vt_frame_fn)(void *user, uint32_t frames);
typedef void (*vt_error_fn)(void *user, const char *message);
vt_status vt_renderer_set_callbacks(vt_renderer *r, vt_frame_fn on_frame, vt_error_fn on_error, void *user);
/* Attaches a learned radiance cache (owned by the caller, which keeps it alive while attached); NULL
 * detaches.  No counterpart in the reference. */
vt_status vt_renderer_set_cache(vt_renderer *r, vt_nrc *cache);
/* The learned cache's training records of the renderer's last wavefront (radiance_cache 1 or 2):
 * *count = records made; up to max_records of them are copied, inputs as [n][11] (the cache's input
 * layout) and targets as [n][4] (spectral radiance per unit throughput at the 4 wavelength lanes);
 * either pointer may be NULL.  Synchronises.  No counterpart in the reference (inspection of the
 * online training data). */
vt_status vt_renderer_cache_records(vt_renderer *r, float *inputs, float *targets, uint32_t max_records,
                                    uint32_t *count);
/* Replaces resetAccumulation() (main.swift:10761); also drops the ReSTIR reservoir history
 * (reservoirHistory = 0). */
vt_status vt_reset_accumulation(vt_renderer *r);
/* Copies the rgba32f running mean (width * height * 4 floats) to host memory; synchronises. */
vt_status vt_accum_readback(vt_renderer *r, float *rgba, size_t bytes);
/* Samples folded into the accumulation so far (frameIndex). */
uint32_t vt_accum_frames(const vt_renderer *r);

/* The spectral tables the integrator reads (D65 scenes), computed on the host: icdf[1025] (wavelength
 * inverse CDF), cmf[471 * 4] (linear-sRGB CMF and cos of the warped phase per nm), basis[471 * 8]
 * (emission basis R G B D65 | C M Y 0 per nm).  Replaces spectral_icdf_kernel's table
 * (main.swift:8117-8185).  Any pointer may be NULL. */
vt_status vt_spectral_tables(float *icdf, float *cmf, float *basis);
/* Reflectance record of a linear-sRGB colour (spectral_lagrange, main.swift:889-896):
 * (L0, L1, L2, 1) bounded-MESE multipliers, or (grey, 0, 0, 2). */
vt_status vt_spectral_reflectance(const float rgb[3], float out[4]);

/* The reference's reflectance-uplift grid built on the context's device, as PathTracerRenderer.
 * refineSpectralGrid builds it once when the grid is loaded (spectral_refine_flags_kernel,
 * spectral_refine_solve_kernel, main.swift:8187-8278): the exact moments of the 86^3 coarse nodes
 * (8-bit codes 0, 3, ..., 255: the generator's Levenberg-Marquardt solve and 1024-node moments,
 * scripts/generate_spectral_tables.py:381-386, in double -- the FourierSRGB86 grid), the cells of the
 * 85^3 whose 8-bit codes the coarse conversion misses by more than 0.35 steps (every code tested, in
 * float as the reference's kernel), and the exact moments of the 4 x 4 x 4 codes of each such cell
 * (float Levenberg-Marquardt from the interpolated start, as the reference's kernel).  Outputs (host;
 * any may be NULL, cells and refined when capacity is 0): coarse = 86^3 x 3 floats, node (r, g, b) at
 * (r 86 + g) 86 + b; flags = 85^3 bytes, cell (i, j, k) at (i 85 + j) 85 + k; cells = the first
 * `capacity` flagged cells in increasing order and refined = their 64 x 3 floats each, code
 * (3i + a, 3j + b, 3k + c) at entry 16 a + 4 b + c (codes clamped to 255); *count = the number of
 * flagged cells.  Allocates its scratch (about 17 MB) and synchronises.  The materials' own uplift
 * (vt_spectral_reflectance) solves the codes it needs on the host in double. */
vt_status vt_uplift_grid_build(vt_context *ctx, float *coarse, uint8_t *flags, uint32_t capacity, uint32_t *cells,
                               float *refined, uint32_t *count);

/* The layered BSDF's energy-compensation tables as the context bakes them on its device (baking them
 * first if needed): metal = 129 x 129 ideal-metal complements then 129 averages (the table
 * scripts/generate_metal_energy.py writes into Shaders/MetalEnergy.metal), luts = 5 x 32 x 32 x 32
 * floats (ideal dielectric complement; its average, ratio and the opaque-dielectric complement and
 * average use the first 32 x 32 of their block), the tables of
 * Vendor/OpenPBR/impl/data/openpbr_energy_arrays.h.  Either pointer may be NULL; synchronises. */
vt_status vt_energy_tables(vt_context *ctx, float *metal, float *luts);

/* ---------------------------------------------------------------- learned radiance cache
 * No counterpart in the reference renderer: a small MLP (BF16 weights on the matrix cores, FP32
 * accumulation) that maps a path vertex -- 11 numbers in [0, 1]: position in the scene box,
 * direction (d + 1) / 2, normal (n + 1) / 2, roughness, wavelength number -- to the radiance leaving it
 * at the path's four wavelengths.  Encoding 12 frequency bands, 4 x 64 ReLU layers, relative-L2 loss,
 * Adam (beta 0.9 / 0.99).  Parameter layout: metal-vibe-tracer_amd/csrc/nrc.h (30720 floats). */
#define VT_NRC_PARAMS 30720
vt_status vt_nrc_create(vt_context *ctx, uint32_t seed, vt_nrc **out);   /* He-uniform initialisation */
void vt_nrc_destroy(vt_nrc *nrc);
vt_status vt_nrc_get_params(vt_nrc *nrc, float *params);                 /* host, VT_NRC_PARAMS floats */
vt_status vt_nrc_set_params(vt_nrc *nrc, const float *params);           /* also resets Adam */
/* outputs[4 q + c] = radiance of query q at lane c; inputs (count x 11) and outputs are device memory. */
vt_status vt_nrc_infer(vt_nrc *nrc, const float *inputs, uint32_t count, float *outputs, void *stream);
/* One Adam step on the batch (inputs count x 11, targets count x 4, device memory). */
vt_status vt_nrc_train(vt_nrc *nrc, const float *inputs, const float *targets, uint32_t count, float learning_rate,
                       void *stream);
/* The batch's loss gradient (host, VT_NRC_PARAMS floats) and mean loss, without updating; synchronises. */
vt_status vt_nrc_gradients(vt_nrc *nrc, const float *inputs, const float *targets, uint32_t count, float *grads,
                           float *loss);

/* Rebuilds a published scene's hierarchy from its own vertex and face streams with other build
 * options (materials and light tree kept).  The reference rebuilds only when the mesh changes
 * (Sources/Acceleration.swift:652-782); this serves the batch autotuner's leaf-size sweep. */
vt_status vt_scene_rebuild(vt_scene *scene, const vt_build_options *opts);
/* Test hook: caps the scene's traversal stack at `groups` node groups (0: the tree's own need), so
 * the overflow report of vt_context_check can be exercised; rays needing more end early. */
vt_status vt_scene_set_stack_cap(vt_scene *scene, uint32_t groups);

/* ---- Batched rendering.  No counterpart in the reference, which renders one view at a time: many
 * images (jobs) of many scenes are advanced together by one wavefront of persistent kernels, and one
 * sample of every job -- all chunks, all vertices -- is captured once in a hipGraph and replayed per
 * sample.  Paths are bounded at max_depth vertices (scenes with mirrors or smooth glass are refused:
 * their delta chains need vt_render_frame's open-ended extension); sampling modes 1-3, no ReSTIR, no
 * cache.  Each job's image is the same estimator, sample for sample, as vt_render_frame with the same
 * camera and sample indices 1, 2, ... */
typedef struct vt_batch vt_batch;
typedef struct vt_batch_job {
    vt_scene *scene;           /* with materials; a scene may serve several jobs */
    uint32_t width, height;
    float eye[3], target[3];   /* as vt_frame_params */
    float fov_degrees;
} vt_batch_job;
typedef struct vt_batch_tuning {
    uint32_t capacity;     /* paths per wavefront chunk (0 or more than a sample's pixels: all of them) */
    uint32_t grid_waves;   /* persistent grids: waves per SIMD, 1..32 */
    uint32_t trace_block;  /* traversal work-group size: 64, 128, 256 or 512 */
    uint32_t leaf_size;    /* BVH leaf size of the batch's scenes, 1..3 (0: keep); changing it rebuilds them */
} vt_batch_tuning;
vt_status vt_batch_create(vt_context *ctx, const vt_batch_job *jobs, uint32_t count, uint32_t max_depth,
                          uint32_t sampling_mode, vt_batch **out);
void vt_batch_destroy(vt_batch *batch);
/* Replays the captured sample `samples` times after the work already queued on `stream`; `stream`
 * waits for them.  A job's scene may be rebuilt (vt_scene_rebuild) or given new materials between
 * calls: the next call sees the scene's new generation, waits for the batch's earlier samples, and
 * uploads the tables and captures the sample again (the accumulation is kept).  Scenes must outlive
 * the batch, and must not be rebuilt while samples of the batch are still running. */
vt_status vt_batch_render(vt_batch *batch, uint32_t samples, void *stream);
/* Sample indices back to 1 and the accumulation cleared; synchronises. */
vt_status vt_batch_reset(vt_batch *batch);
uint32_t vt_batch_frames(const vt_batch *batch);
/* Job `job`'s rgba32f running mean (width * height * 4 floats); synchronises. */
vt_status vt_batch_readback(vt_batch *batch, uint32_t job, float *rgba, size_t bytes);
vt_status vt_batch_get_tuning(const vt_batch *batch, vt_batch_tuning *tuning);
/* Sets the tuning and captures the sample again (rebuilding the scenes for another leaf size). */
vt_status vt_batch_set_tuning(vt_batch *batch, const vt_batch_tuning *tuning);
/* The autotuner: from the current tuning, each parameter in turn (capacity 2^20 .. all, grid waves
 * 1..16, traversal block 64..256, and with flags & 1 the leaf size 1..3) over its candidates, timing
 * replays of the captured sample and keeping the fastest.  Ends with the best tuning set and the
 * accumulation reset; *best_ms = its milliseconds per sample. */
vt_status vt_batch_autotune(vt_batch *batch, uint32_t flags, vt_batch_tuning *best, double *best_ms);

/* Host-only build of the same 8-wide BVH the upload builds (no device needed), for inspection:
 * *nodes receives node_count 80-byte node records, *tris face_count 48-byte triangle records
 * (layout: metal-vibe-tracer_amd/csrc/bvh8.h).  Free both with vt_host_free.  Mirrors the
 * reference's host builder entry WideBVH.build (Sources/Acceleration.swift:296-300). */
vt_status vt_bvh8_build_host(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count,
                             const vt_build_options *opts, void **nodes, uint32_t *node_count, void **tris,
                             uint32_t *depth, float *root_max);
void vt_host_free(void *p);

/* Host-only build of the flat binary hierarchy of build option collapse = 2, replacing OBJMesh.build /
 * hierarchy (Sources/StudioModel.swift:679-760): *order (face_count u32: hierarchy position -> face) and
 * *nodes (node_count 48-byte MeshNode records: lo, hi as 4 floats, links as 4 int32) are malloc'ed
 * (free with vt_host_free); *depth is the number of levels. */
vt_status vt_hierarchy_build_host(const float *vertices, uint32_t vertex_count, const uint32_t *faces,
                                  uint32_t face_count, uint32_t **order, void **nodes, uint32_t *node_count,
                                  uint32_t *depth);

#ifdef __cplusplus
}
#endif
#endif
