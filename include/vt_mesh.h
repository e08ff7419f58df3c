/* vt_mesh.h -- the scene-side data formats that feed the path: the indexed mesh payload "VTMESH01"
 * and the OBJ reader, both host-only and exported by libvt.so.  Their output is the vertex / face
 * arrays vt_scene_upload_indexed takes (vertices: 8 floats each -- position, normal, uv; faces: 4 u32
 * each -- three vertex indices and the subset).  Errors return a status and set vt_last_error(). */
#ifndef VT_MESH_H
#define VT_MESH_H

#include <stdint.h>

#include "vt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- VTMESH01 payload
 * Replaces MeshGeometry.payload / MeshGeometry(payload:maximumFaces:) (Sources/MeshGeometry.swift:
 * 61-92): "VTMESH01", vertex and face counts (u32 little-endian), the vertices (8 f32 LE each) and the
 * faces (4 u32 LE each).  16 + 32 V + 16 F bytes. */
uint64_t vt_mesh_payload_bytes(uint32_t vertex_count, uint32_t face_count);
/* Writes the payload into out (capacity bytes, at least vt_mesh_payload_bytes). */
vt_status vt_mesh_payload_write(const float *vertices, uint32_t vertex_count, const uint32_t *faces, uint32_t face_count,
                                uint8_t *out, uint64_t capacity);
/* Reads a payload.  Errors as the reference: "Invalid mesh data: missing header." (shorter than 16
 * bytes or another magic) and "Invalid mesh data: counts do not match the payload." (more faces than
 * maximum_faces, more vertices than 3 x faces, or a size other than 16 + 32 V + 16 F).  Vertex indices
 * are not checked here (vt_scene_upload_indexed does).  *vertices and *faces are malloc'ed: free both
 * with vt_host_free. */
vt_status vt_mesh_payload_read(const uint8_t *data, uint64_t length, uint32_t maximum_faces, float **vertices,
                               uint32_t *vertex_count, uint32_t **faces, uint32_t *face_count);

/* MeshGeometry(triangles:) (Sources/MeshGeometry.swift:21-26, 49-57): the indexed arrays of `count`
 * MeshTriangle records (32 floats each, Sources/StudioModel.swift:13-15: corners a, b, c and normals
 * na, nb, nc as float4 with w ignored, uvab = (uv of a, uv of b), uvc = (uv of c, subset, unused)),
 * welding corners whose position, normal and uv bits are identical (0 and -0 stay distinct).  The
 * subset is uvc.z truncated when finite and in [0, 4e9), else 0.  *vertices and *faces are malloc'ed:
 * free both with vt_host_free. */
vt_status vt_mesh_from_triangles(const float *triangles, uint32_t count, float **vertices, uint32_t *vertex_count,
                                 uint32_t **faces, uint32_t *face_count);

/* ---------------------------------------------------------------- OBJ reader
 * Replaces OBJMesh.parts (Sources/StudioModel.swift:444-684): statements o, g, usemtl, mtllib, v, vn,
 * vt (v defaults to 0 and is stored as 1 - v), f (v, v/vt, v//vn, v/vt/vn; 3-4096 corners, negative
 * indices relative to the end, fan triangulation); physical lines end at LF, CR, CR LF, VT or FF; a
 * trailing backslash outside a comment continues a statement.  A triangle whose corners are collinear
 * or coincident (|cross| <= 1e-7 |e1| |e2|) is skipped and counted.  Faces go to parts keyed by
 * (object, group) in order of first use, each with its materials ("subsets", in order of first use
 * within the part); a part's corners with bit-identical position, normal and uv share one vertex
 * (MeshGeometry.Builder).  Corners without vn take the face normal.  Errors name the statement's
 * first physical line ("OBJ index is out of bounds on line 5.").  node_budget bounds the objects plus
 * parts created (0: no bound; the reference's scene holds 65535 nodes); triangle_limit bounds the
 * triangles kept (0: the reference's 2^30). */
typedef struct vt_obj vt_obj;
vt_status vt_obj_parse(const char *text, uint64_t length, uint32_t node_budget, uint64_t triangle_limit, vt_obj **out);
void vt_obj_destroy(vt_obj *obj);
uint32_t vt_obj_part_count(const vt_obj *obj);
uint32_t vt_obj_skipped(const vt_obj *obj);          /* degenerate triangles left out */
uint32_t vt_obj_library_count(const vt_obj *obj);    /* mtllib statements (their argument text) */
const char *vt_obj_library(const vt_obj *obj, uint32_t index);
/* A part's object and group names, subset count and geometry (owned by obj). */
vt_status vt_obj_part(const vt_obj *obj, uint32_t part, const char **object, const char **group,
                      uint32_t *subset_count, const float **vertices, uint32_t *vertex_count, const uint32_t **faces,
                      uint32_t *face_count);
const char *vt_obj_subset(const vt_obj *obj, uint32_t part, uint32_t subset); /* material name, or NULL */

#ifdef __cplusplus
}
#endif

#endif /* VT_MESH_H */
