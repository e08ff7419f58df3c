/*
 * oracle.h -- CPU restatement of metal-vibe-tracer's ray-traversal hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or calls this code:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as the
 * checker (or as the timed CPU baseline).  Every function names the reference file:line whose
 * algorithm it restates.  Build: oracle/Makefile (gcc, -ffp-contract=off, OpenMP).
 *
 * Pinning (what checks each function against something other than itself):
 *   vo_pcg_hash / vo_rand_f      closed-form hash; tests/GPUChecks.swift:25-30 (seed 890625759:
 *                                second draw is exactly 0) and the published PCG-RXS-M-XS constants.
 *   vo_watertight_edge           golden tie vector tests/Fix_renderer-followups.swift:167-168,234
 *                                (+2^-24 / -2^-24) and exact antisymmetry (:170-176).
 *   vo_intersect_triangle        hand-derived closed-form cases (axis rays through known triangles,
 *                                barycentrics of the hit point), and the brute-force identity
 *                                "every ray aimed at a shared edge/vertex of a closed mesh hits"
 *                                (tests/Fix_accel.swift:302-337).
 *   vo_widebvh_build             structural invariants (every item in exactly one leaf, child
 *                                boxes enclose their items, depth limit), exhaustive SAH check on
 *                                small inputs, and traversal == brute force.
 *   vo_trace_*                   closest hit == brute force (t bit-exact, id by the tie rule),
 *                                any hit == (closest hit exists), tests/Fix_geometry.swift:141-161.
 */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* main.swift:91-101 (pcg_hash, rand_f) */
uint32_t vo_pcg_hash(uint32_t input);
float vo_rand_f(uint32_t *seed);

/* main.swift:1557-1562 (watertight_edge), contraction off, FMA tie-break */
float vo_watertight_edge(float px, float py, float qx, float qy);

/* main.swift:1537-1582 (watertight_ray + intersect_mesh_triangle): two-sided hit in [tmin, tmax).
 * o, d: ray; a, b, c: vertices (3 floats each).  Returns 1 on hit with t, b1, b2 written. */
int vo_intersect_triangle(const float *o, const float *d, const float *a, const float *b, const float *c,
                          float tmin, float tmax, float *t, float *b1, float *b2);

/* Sources/Acceleration.swift:445-458 (MeshSceneLayout.bounds): 6 floats per face. */
void vo_mesh_bounds(const float *vertices /* 8 floats each */, const uint32_t *faces /* 4 u32 each */,
                    uint32_t face_count, float *bounds /* 6 * face_count */);

/* Sources/Acceleration.swift:23-30 (MeshWideNode), 128 bytes. */
typedef struct vo_wide_node {
    float lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4];
    int32_t child[4];
    int32_t count[4];
} vo_wide_node;

/* Sources/Acceleration.swift:106-351 (WideBVH.build): binned-SAH 4-wide build, leaf size
 * `leaf_size`, median rebuild when deeper than `max_depth`.  nodes_out is malloc'ed (free with
 * vo_free); order_out[n] receives the item order.  Returns the node count, depth in *depth. */
int32_t vo_widebvh_build(const float *bounds, int32_t n, int32_t leaf_size, int32_t max_depth,
                         vo_wide_node **nodes_out, int32_t *order_out, int32_t *depth);
void vo_free(void *p);

/* A closest hit; face is the DOCUMENT face index (input order). */
typedef struct vo_hit {
    float t, b1, b2;
    int32_t face; /* -1: miss */
} vo_hit;

/* Scene view used by the traversal functions: faces and vertices in document order plus a BVH
 * over them (order maps BLAS position -> document face). */
typedef struct vo_scene {
    const float *vertices;    /* 8 floats per vertex */
    const uint32_t *faces;    /* 4 u32 per face (a, b, c, tag) */
    uint32_t face_count;
    const vo_wide_node *nodes; /* may be NULL for brute force */
    const int32_t *order;
    float root_max;           /* MeshSceneLayout.Asset.maxAbs */
} vo_scene;

/* main.swift:1788-1812 (blas_closest): reference visit order, first found wins among equal t.
 * Optional counters (may be NULL): node fetches and triangle tests. */
vo_hit vo_blas_closest(const vo_scene *s, const float *o, const float *d, float tmin, float tmax,
                       uint64_t *nodes_visited, uint64_t *tris_tested);
/* main.swift:1815-1838 (blas_any), slack off. */
int vo_blas_any(const vo_scene *s, const float *o, const float *d, float tmin, float tmax,
                uint64_t *nodes_visited, uint64_t *tris_tested);

/* Brute force over every face (tests/Fix_accel.swift:131-141): the closest hit under the
 * canonical rule "smallest t, then smallest document face index". */
vo_hit vo_brute_closest(const vo_scene *s, const float *o, const float *d, float tmin, float tmax);
int vo_brute_any(const vo_scene *s, const float *o, const float *d, float tmin, float tmax);

/* Batched forms over SoA rays, OpenMP-parallel over rays (threads <= 0: OpenMP default).
 * mode 0: reference BVH order (vo_blas_closest), 1: brute force.
 * counters[0] += node fetches, counters[1] += triangle tests (mode 0 only; may be NULL). */
void vo_trace_closest_batch(const vo_scene *s, const float *ox, const float *oy, const float *oz,
                            const float *dx, const float *dy, const float *dz, const float *tmin,
                            const float *tmax, int64_t n, int mode, int threads, float *t, float *b1,
                            float *b2, int32_t *face, uint64_t *counters);
void vo_trace_any_batch(const vo_scene *s, const float *ox, const float *oy, const float *oz,
                        const float *dx, const float *dy, const float *dz, const float *tmin,
                        const float *tmax, int64_t n, int mode, int threads, uint8_t *occluded,
                        uint64_t *counters);

/* ------------------------------------------------------------------ light tree (lighttree.c) */

typedef struct vo_light_tree {
    uint32_t *words;      /* 8 per node */
    uint32_t node_count;
    uint32_t *order;      /* leaf position -> input emitter */
    uint32_t *leaf_node;  /* leaf position -> node */
    uint32_t depth;
    float grid_lo[3], grid_hi[3];
} vo_light_tree;

/* Sources/LightTree.swift:112-190: tris = 9 floats per emitter (a, b, c), power = area x weight. */
int32_t vo_light_tree_build(const float *tris, const double *power, uint32_t count, int max_leaf, double leaf_ratio,
                            vo_light_tree *out);
void vo_light_tree_free(vo_light_tree *t);

/* The tree as the walk reads it (MaterialLibrary.lightTreeList, main.swift:9374-9401). */
typedef struct vo_light_view {
    const uint32_t *nodes;      /* 8 words per node */
    const uint32_t *leaves;     /* 3 words per leaf position: face, power bits, leaf node */
    float lo[3], scale[3];      /* grid lo, (hi - lo) / 65535 */
    uint32_t count;             /* emitters */
    const uint32_t *ids;        /* emitter faces, ascending */
    const uint32_t *positions;  /* leaf position of ids[i] */
} vo_light_view;

float vo_light_tree_importance(const vo_light_view *t, uint32_t node, const float *p);
uint32_t vo_light_tree_sample(const vo_light_view *t, const float *p, float u, float *pmf);
float vo_light_tree_pmf(const vo_light_view *t, const float *p, uint32_t position);

/* ------------------------------------------------------------------ integrator (integrator.c) */

/* Scene as the integrators read it; material records are VO_MATERIAL_FLOATS (20) floats per slot:
 * type (u32 bits), albedo rgb | Lagrange multipliers (L0, L1, L2, w: 1 MESE, 2 flat grey) | emission
 * rgb, selection weight | roughness, ior, metalness, transmission | coat weight, specular anisotropy,
 * fuzz weight, 1 unused. */
#define VO_MATERIAL_FLOATS 20
typedef struct vo_render_scene {
    const vo_scene *geo;
    const float *materials;
    uint32_t material_count;
    const uint32_t *emitter_face;   /* emitting faces, power CDF order */
    const float *emitter_cdf;       /* cumulative area x weight / total, last = 1 */
    uint32_t emitter_count;
    float emitter_total;
    const float *icdf;              /* 1025 */
    const float *cmf;               /* 471 x 4: rgb CMF, cos phase */
    const float *basis;             /* 471 x 2 x 4: R G B D65 | C M Y 0 */
    const vo_light_view *tree;      /* light tree, or NULL for power sampling */
    /* Environment map (lat-long rgba32f, env_w x env_h, row 0 at the zenith) with its sampling tables
     * (vo_env_tables), and the frame's Uniforms.environment: strength (x) and rotation in radians (y).
     * env NULL or strength <= 0: no environment light. */
    const float *env, *env_rows, *env_cols;
    uint32_t env_w, env_h;
    float env_strength, env_rotation;
} vo_render_scene;

/* MaterialLibrary.environmentImportance (Sources/StudioModel.swift:1300-1380): the sampling tables of
 * a width x height rgba32f lat-long map.  rows: 2 * height floats, rows[2 y + table]; cols:
 * 2 * height * width floats, cols[(table * height + y) * width + x]; table 0 is the MIS-compensated
 * table, 1 the original.  Returns 0, or -1 for an empty map or no memory. */
int vo_env_tables(const float *rgba, uint32_t width, uint32_t height, float *rows, float *cols);

typedef struct vo_frame {
    uint32_t width, height;
    float eye[3], target[3], fov;
    int32_t max_depth;
    uint32_t sampling_mode;
    uint32_t sampler;        /* 0: the PCG stream; 1: the Z++ sampler (MIS path loop only) */
    uint32_t z_temporal;     /* Z++ temporal model 0..3 (main.swift:497-506) */
    uint32_t sample_index0;  /* first sample of the accumulation (Z mode: frame = sample - this + 1) */
    uint32_t indirect_reuse; /* sampling mode 0: 0 = ReSTIR GI at the first indirect diffuse vertex (the
                                reference's IndirectReuse.restirGI), 1 = none (plain MIS suffix) */
} vo_frame;

/* Mean linear-sRGB radiance of `samples` samples (sample indices sample_index0...) for each listed
 * pixel (index y * width + x): the reference's MIS path loop (see integrator.c), float arithmetic
 * (_f32: the reference integrator) or double (_f64: the FP64 oracle). */
void vo_render_f32(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                   const uint32_t *pixels, uint32_t npixels, int threads, double *out_rgb);
void vo_render_f64(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                   const uint32_t *pixels, uint32_t npixels, int threads, double *out_rgb);

/* ReSTIR DI (samplingMode 0, see integrator.c): `samples` consecutive samples rendered as frames, each
 * pass 1 (RIS over `candidates` light samples + temporal reuse of the previous sample's reservoirs)
 * then pass 2 (spatial reuse, shading, path suffix); the first sample has no history.  eyes: NULL
 * (fixed camera) or 3 floats per sample.  out_rgb: width * height * 3 per-pixel means. */
/* Environment queries for tests (integrator.c): radiance and light-sample density per direction, and
 * light samples (direction, pdf, emission, light id) from a PCG stream. */
void vo_env_query_f32(const vo_render_scene *sc, const float *dirs, uint32_t n, uint32_t table, float *rgb, float *pdf);
void vo_env_query_f64(const vo_render_scene *sc, const float *dirs, uint32_t n, uint32_t table, float *rgb, float *pdf);
void vo_light_samples_f32(const vo_render_scene *sc, const float *p, uint32_t n, uint32_t seed, uint32_t table,
                          float *wi, float *pdf, float *rgb, uint32_t *light);
void vo_light_samples_f64(const vo_render_scene *sc, const float *p, uint32_t n, uint32_t seed, uint32_t table,
                          float *wi, float *pdf, float *rgb, uint32_t *light);
void vo_render_restir_f32(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                          uint32_t candidates, const float *eyes, int threads, double *out_rgb);
void vo_render_restir_f64(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                          uint32_t candidates, const float *eyes, int threads, double *out_rgb);

/* Energy-compensation tables of the layered BSDF (energy.c).
 * vo_energy_set_dielectric(tables): OpenPBR's five dielectric unorm16 tables back to back, in
 * VO_LUT order (32768, 1024, 1024, 32768, 1024 values); must precede any table read.
 * vo_energy_table(id): unorm16, 32 x 32 x 32 (IOR, sqrt(alpha), cosine) or 32 x 32 (IOR, sqrt(alpha)).
 * vo_metal_energy_table(): 129 x 129 ideal-metal complements (roughness, sqrt cosine), then the
 * 129 average complements. */
enum { VO_LUT_DIELECTRIC = 0, VO_LUT_DIELECTRIC_AVG, VO_LUT_DIELECTRIC_RATIO, VO_LUT_OPAQUE, VO_LUT_OPAQUE_AVG,
       VO_LUT_COUNT };
void vo_energy_set_dielectric(const uint16_t *tables);
/* The fuzz lobe's LTC table (OpenPBR's openpbr_ltc_data.h, committed as data): 32 x 32 entries
 * (a_inv, b_inv, R), row = sheen roughness, column = cosine; copied. */
void vo_ltc_set(const float *table);
const float *vo_ltc_table(void);
const uint16_t *vo_energy_table(int id);
const float *vo_metal_energy_table(void);
float vo_energy_ior_node(int i);

/* The layered BSDF subset (openpbr_impl.h) on its own: params = type (1 GLOSSY, 4 OPENPBR),
 * roughness, ior, metalness, transmission, coat, anisotropy, tangent xyz, fuzz; n shading normal on the arriving side, wo toward the
 * viewer, rho base colour per lane, inside = hit from the back of the surface. */
void vo_bsdf_eval_f32(const float *params, const float *n, const float *wo, const float *wi, const float *rho,
                      int inside, float *fcos, float *pdf);
int vo_bsdf_sample_f32(const float *params, const float *n, const float *wo, const float *rho, int inside,
                       const float *u3, float *dir, float *weight, float *pdf);
void vo_bsdf_eval_f64(const float *params, const float *n, const float *wo, const float *wi, const float *rho,
                      int inside, float *fcos, float *pdf);
int vo_bsdf_sample_f64(const float *params, const float *n, const float *wo, const float *rho, int inside,
                       const float *u3, float *dir, float *weight, float *pdf);

/* The library's cache-terminated estimator with a fixed network (integrator.c): mean radiance of every
 * pixel over `samples` samples, MIS mode.  params: 30720 floats (oracle/nrc.py layout); the cache's
 * position frame is (p - box_lo) * box_scale. */
void vo_render_nrc_f32(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                       uint32_t period, const float *params, const float *box_lo, const float *box_scale, int threads,
                       double *out_rgb);
void vo_render_nrc_f64(const vo_render_scene *sc, const vo_frame *fr, uint32_t sample_index0, uint32_t samples,
                       uint32_t period, const float *params, const float *box_lo, const float *box_scale, int threads,
                       double *out_rgb);

/* For tests that plant a mistake in the oracle (never the product): 0 = faithful,
 * 1 = drop the FMA tie-break of watertight_edge, 2 = swap b1/b2, 3 = ReSTIR spatial reuse leaves
 * the neighbours' confidence out of M, 4 = ReSTIR RIS counts only candidates with pdf > 0 in M,
 * 5 = the GGX reflection lobe drops the shadowing term G1(wi), 10 = ReSTIR GI keeps its source pdf in
 * solid-angle measure, 11 = the environment's cell density drops its 1 / sin(theta), 12 = every
 * environment light sample uses the compensated table (also where no BSDF continuation covers its
 * zeros), 13 = the coat's darkening factor left out (the base under a coat scaled by 1 - in_reflected
 * alone), 14 = the anisotropic alphas swapped (the lobe stretched along the bitangent), 15 = the sheen
 * evaluated in a frame turned the wrong way about the normal.  9 is a test switch, not a mistake: ReSTIR GI without its spatial taps. */
void vo_set_fault(int fault);
int vo_get_fault(void);
/* Nearest half-precision value of x (ties to even): what an rgba16f texel store keeps. */
float vo_half_round(float x);
int vo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
