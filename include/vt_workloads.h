/*
 * vt_workloads.h -- seeded synthetic scenes and ray sets for the benchmark workloads.
 *
 * These have no counterpart in the reference (its scenes are imported files and hard-coded
 * procedural scenes); they generate, deterministically from a seed, the inputs the traversal and
 * integrator are measured on.  All arrays are HOST memory.  Each generator called with NULL
 * outputs only reports the counts.  Vertices are MeshVertex records (8 floats: position, normal,
 * uv), faces MeshFace records (a, b, c, tag).
 */
#ifndef VT_WORKLOADS_H
#define VT_WORKLOADS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A field of `spheres` icospheres, each the icosahedron subdivided `subdivisions` times
 * (3 -> 1280 triangles), radius exp(mu + sigma N(0,1)), centre uniform in [-half_extent,
 * half_extent]^3.  Face tag = sphere index. */
int vt_gen_icosphere_field(uint32_t spheres, uint32_t subdivisions, float mu, float sigma, float half_extent,
                           uint32_t seed, float *vertices, uint32_t *vertex_count, uint32_t *faces,
                           uint32_t *face_count);

/* An fBm value-noise heightfield on a grid x grid vertex lattice over [-half_extent,
 * half_extent]^2 (x, z), `octaves` octaves with amplitude gain `gain`, height scale `height`.
 * 2 (grid-1)^2 faces, row-major, tag 0. */
int vt_gen_heightfield(uint32_t grid, uint32_t octaves, float gain, float half_extent, float height, uint32_t seed,
                       float *vertices, uint32_t *vertex_count, uint32_t *faces, uint32_t *face_count);

/* The Cornell box of the reference's procedural scene 1 as triangles (main.swift:1421-1438): five
 * walls of [-1, 1]^3 (floor, ceiling, back: tag 0 white; left x = -1: tag 1 red; right x = 1: tag 2
 * green), a tall box (centre (-0.33, -0.4, 0.35), half extents (0.3, 0.6, 0.3), yaw 0.32) and a
 * short box (centre (0.33, -0.7, -0.25), half extents 0.3, yaw -0.30), tag 0, and a square ceiling
 * light of side `light_size` at y = 0.999 emitting downwards, tag 3.  Flat-shaded, unwelded:
 * 36 faces, 108 vertices. */
int vt_gen_cornell(float light_size, float *vertices, uint32_t *vertex_count, uint32_t *faces, uint32_t *face_count);

/* Pinhole camera rays through a jittered width x height grid (one per pixel, row-major):
 * eye, target, vertical fov in degrees.  tmin = 0, tmax = 1e20. */
void vt_gen_camera_rays(const float eye[3], const float target[3], float fov_degrees, uint32_t width, uint32_t height,
                        uint32_t seed, float *ox, float *oy, float *oz, float *dx, float *dy, float *dz, float *tmin,
                        float *tmax);

/* Cosine-hemisphere bounce rays from closest hits: ray i starts at hit i, or, when ray i
 * missed, at a hit drawn uniformly from the hits.  Origins are offset along the geometric
 * normal (facing the incoming ray) as ray_origin does for imported meshes (main.swift:1343-1369),
 * tmin = ray_t_min, tmax = 1e20.  Returns the number of input hits (0: no rays written). */
uint32_t vt_gen_bounce_rays(const float *vertices, const uint32_t *faces, uint32_t n, const float *in_dx,
                            const float *in_dy, const float *in_dz, const int32_t *hit_face, const float *hit_b1,
                            const float *hit_b2, uint32_t seed, float *ox, float *oy, float *oz, float *dx, float *dy,
                            float *dz, float *tmin, float *tmax);

/* Uniformly random rays (origin uniform in the box [lo, hi], direction uniform on the sphere),
 * tmin = 0, tmax = 1e20: the incoherent set the reference's own throughput probe uses
 * (scripts/measure_acceleration.swift:176-187). */
void vt_gen_random_rays(const float lo[3], const float hi[3], uint32_t n, uint32_t seed, float *ox, float *oy,
                        float *oz, float *dx, float *dy, float *dz, float *tmin, float *tmax);

#ifdef __cplusplus
}
#endif
#endif
