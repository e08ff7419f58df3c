"""Python binding of libvt.so (include/vt.h, include/vt_workloads.h) over ctypes.

This is a thin host-side binding used by the tests and the benchmark; the runtime itself is C++
and HIP.  Device buffers are passed as torch tensors on the HIP device (only their data pointers
cross the ABI).  Nothing here computes results: every trace call runs the gfx950 kernels, and a
missing or stale libvt.so raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGE_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PACKAGE_DIR, "libvt.so")

VT_OK = 0
STATUS_NAMES = {0: "VT_OK", 1: "VT_ERROR_INVALID_ARGUMENT", 2: "VT_ERROR_OUT_OF_MEMORY", 3: "VT_ERROR_DEVICE",
                4: "VT_ERROR_UNSUPPORTED"}

# Every symbol include/vt.h, include/vt_mesh.h, include/vt_sampler.h and include/vt_workloads.h declare.
EXPORTED = [
    "vt_context_create", "vt_context_destroy", "vt_last_error", "vt_scene_upload_indexed", "vt_scene_destroy",
    "vt_scene_get_info", "vt_trace_closest", "vt_trace_any", "vt_trace_stats", "vt_bvh8_build_host", "vt_host_free",
    "vt_gen_icosphere_field", "vt_gen_heightfield", "vt_gen_camera_rays", "vt_gen_bounce_rays", "vt_gen_random_rays",
    "vt_scene_set_materials", "vt_renderer_create", "vt_renderer_destroy", "vt_render_frame", "vt_reset_accumulation",
    "vt_accum_readback", "vt_accum_frames", "vt_gen_cornell", "vt_spectral_tables", "vt_spectral_reflectance",
    "vt_scene_set_light_tree", "vt_light_tree_build_host", "vt_energy_tables",
    "vt_nrc_create", "vt_nrc_destroy", "vt_nrc_get_params", "vt_nrc_set_params", "vt_nrc_infer", "vt_nrc_train",
    "vt_nrc_gradients", "vt_renderer_set_cache", "vt_renderer_cache_records", "vt_scene_rebuild",
    "vt_batch_create", "vt_batch_destroy", "vt_batch_render", "vt_batch_reset", "vt_batch_frames",
    "vt_batch_readback", "vt_batch_get_tuning", "vt_batch_set_tuning", "vt_batch_autotune",
    "vt_mesh_payload_bytes", "vt_mesh_payload_write", "vt_mesh_payload_read", "vt_obj_parse", "vt_obj_destroy",
    "vt_obj_part_count", "vt_obj_skipped", "vt_obj_library_count", "vt_obj_library", "vt_obj_part", "vt_obj_subset",
    "vt_z_draws", "vt_context_check", "vt_scene_set_stack_cap", "vt_mesh_from_triangles",
    "vt_scene_upload_triangles", "vt_scene_set_surface_settings", "vt_renderer_set_callbacks",
    "vt_trace_stats_noalloc", "vt_tlas_create", "vt_tlas_destroy", "vt_tlas_get_info", "vt_tlas_trace_closest",
    "vt_tlas_trace_any", "vt_scene_set_environment", "vt_environment_tables", "vt_hierarchy_build_host",
    "vt_uplift_grid_build",
]


class VTError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class _ContextDesc(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("memory_budget", ctypes.c_uint64)]


class _BuildOptions(ctypes.Structure):
    _fields_ = [("leaf_size", ctypes.c_uint32), ("bins", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("collapse", ctypes.c_uint32)]


class _SceneInfo(ctypes.Structure):
    _fields_ = [("face_count", ctypes.c_uint32), ("vertex_count", ctypes.c_uint32), ("node_count", ctypes.c_uint32),
                ("depth", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("node_bytes", ctypes.c_uint32), ("tri_bytes", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64),
                ("root_max", ctypes.c_float), ("build_seconds", ctypes.c_float)]


class Instance(ctypes.Structure):
    """vt_instance: a published scene (its `mesh` handle) and the 12 rows-major floats of its
    object-to-world matrix (3 rows of 4)."""
    _fields_ = [("mesh", ctypes.c_void_p), ("world", ctypes.c_float * 12)]


class _TlasInfo(ctypes.Structure):
    _fields_ = [("instance_count", ctypes.c_uint32), ("node_count", ctypes.c_uint32), ("depth", ctypes.c_uint32),
                ("blas_depth", ctypes.c_uint32), ("rendered_triangles", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64), ("enlarge", ctypes.c_float), ("reach", ctypes.c_float),
                ("build_seconds", ctypes.c_float)]


_fp = ctypes.POINTER(ctypes.c_float)


class Material(ctypes.Structure):
    """vt_material (64 bytes): MaterialType 0 DIFFUSE, 1 GLOSSY, 2 DIELECTRIC, 3 EMISSIVE, 4 OPENPBR."""
    _fields_ = [("type", ctypes.c_uint32), ("albedo", ctypes.c_float * 3), ("emission", ctypes.c_float * 3),
                ("roughness", ctypes.c_float), ("ior", ctypes.c_float), ("metalness", ctypes.c_float),
                ("specular_weight", ctypes.c_float), ("transmission", ctypes.c_float), ("coat", ctypes.c_float),
                ("anisotropy", ctypes.c_float), ("fuzz", ctypes.c_float), ("reserved", ctypes.c_float)]


def material(type=0, albedo=(0.7, 0.7, 0.7), emission=(0.0, 0.0, 0.0), roughness=0.5, ior=1.5, metalness=0.0,
             specular_weight=1.0, transmission=0.0, coat=0.0, anisotropy=0.0, fuzz=0.0) -> Material:
    return Material(type, (ctypes.c_float * 3)(*albedo), (ctypes.c_float * 3)(*emission), roughness, ior, metalness,
                    specular_weight, transmission, coat, anisotropy, fuzz)


class BatchJob(ctypes.Structure):
    """vt_batch_job."""
    _fields_ = [("scene", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("eye", ctypes.c_float * 3), ("target", ctypes.c_float * 3), ("fov_degrees", ctypes.c_float)]


class ZQuery(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("width", "height", "frame_index", "sample_index", "temporal_model",
                                               "path_vertex", "stream", "candidate", "candidate_bits", "draws", "kind")]


def z_draws(width, height, frame_index, sample_index, temporal_model=0, path_vertex=0, stream=0, candidate=0,
            candidate_bits=0, draws=1, kind=1, out=None, stream_handle=None):
    """The Z++ sampler's draws of one event for every pixel on the device (vt_z_draws): a float32 device
    tensor shaped (height, width, draws, kind)."""
    import torch
    if out is None:
        out = torch.empty((height, width, draws, kind), dtype=torch.float32, device="cuda:0")
    elif (out.dtype != torch.float32 or not out.is_cuda or not out.is_contiguous()
          or out.numel() < height * width * draws * kind):
        raise ValueError("out must be a contiguous float32 device tensor of at least height * width * draws * kind "
                         "elements")
    q = ZQuery(width, height, frame_index, sample_index, temporal_model, path_vertex, stream, candidate,
               candidate_bits, draws, kind)
    _check(load().vt_z_draws(ctypes.byref(q), ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream_handle)))
    return out


class BatchTuning(ctypes.Structure):
    """vt_batch_tuning."""
    _fields_ = [("capacity", ctypes.c_uint32), ("grid_waves", ctypes.c_uint32), ("trace_block", ctypes.c_uint32),
                ("leaf_size", ctypes.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class _FrameParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("eye", ctypes.c_float * 3),
                ("target", ctypes.c_float * 3), ("fov_degrees", ctypes.c_float), ("max_depth", ctypes.c_uint32),
                ("sampling_mode", ctypes.c_uint32), ("sample_index", ctypes.c_uint32),
                ("restir_candidates", ctypes.c_uint32), ("radiance_cache", ctypes.c_uint32),
                ("cache_train_period", ctypes.c_uint32), ("cache_learning_rate", ctypes.c_float),
                ("material_sort", ctypes.c_uint32), ("sampler", ctypes.c_uint32), ("z_temporal", ctypes.c_uint32),
                ("indirect_reuse", ctypes.c_uint32), ("environment_strength", ctypes.c_float),
                ("environment_rotation", ctypes.c_float)]


class _RaySoA(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class _HitSoA(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("t", "b1", "b2", "face", "instance")]


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libvt.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    c = ctypes
    vp, u32, i32, f32 = c.c_void_p, c.c_uint32, c.c_int32, c.c_float
    P = c.POINTER
    sig = {
        "vt_context_create": (i32, [P(_ContextDesc), P(vp)]),
        "vt_context_destroy": (None, [vp]),
        "vt_last_error": (c.c_char_p, []),
        "vt_scene_upload_indexed": (i32, [vp, vp, u32, vp, u32, P(_BuildOptions), P(vp)]),
        "vt_scene_destroy": (None, [vp]),
        "vt_scene_get_info": (i32, [vp, P(_SceneInfo)]),
        "vt_trace_closest": (i32, [vp, P(_RaySoA), u32, P(_HitSoA), vp]),
        "vt_trace_any": (i32, [vp, P(_RaySoA), u32, vp, vp]),
        "vt_tlas_create": (i32, [vp, vp, u32, P(vp)]),
        "vt_tlas_destroy": (None, [vp]),
        "vt_tlas_get_info": (i32, [vp, P(_TlasInfo)]),
        "vt_tlas_trace_closest": (i32, [vp, P(_RaySoA), u32, P(_HitSoA), vp]),
        "vt_tlas_trace_any": (i32, [vp, P(_RaySoA), u32, vp, vp]),
        "vt_trace_stats": (i32, [vp, P(_RaySoA), u32, c.c_int, P(c.c_uint64), vp]),
        "vt_bvh8_build_host": (i32, [vp, u32, vp, u32, P(_BuildOptions), P(vp), P(u32), P(vp), P(u32), P(f32)]),
        "vt_host_free": (None, [vp]),
        "vt_gen_icosphere_field": (c.c_int, [u32, u32, f32, f32, f32, u32, vp, P(u32), vp, P(u32)]),
        "vt_gen_heightfield": (c.c_int, [u32, u32, f32, f32, f32, u32, vp, P(u32), vp, P(u32)]),
        "vt_gen_camera_rays": (None, [P(f32), P(f32), f32, u32, u32, u32] + [vp] * 8),
        "vt_gen_bounce_rays": (u32, [vp, vp, u32, vp, vp, vp, vp, vp, vp, u32] + [vp] * 8),
        "vt_gen_random_rays": (None, [P(f32), P(f32), u32, u32] + [vp] * 8),
        "vt_gen_cornell": (c.c_int, [f32, vp, P(u32), vp, P(u32)]),
        "vt_scene_set_materials": (i32, [vp, vp, u32]),
        "vt_renderer_create": (i32, [vp, P(vp)]),
        "vt_renderer_destroy": (None, [vp]),
        "vt_render_frame": (i32, [vp, vp, P(_FrameParams), u32, vp]),
        "vt_reset_accumulation": (i32, [vp]),
        "vt_energy_tables": (i32, [vp, vp, vp]),
        "vt_uplift_grid_build": (i32, [vp, vp, vp, u32, vp, vp, P(u32)]),
        "vt_nrc_create": (i32, [vp, u32, P(vp)]),
        "vt_nrc_destroy": (None, [vp]),
        "vt_nrc_get_params": (i32, [vp, vp]),
        "vt_nrc_set_params": (i32, [vp, vp]),
        "vt_nrc_infer": (i32, [vp, vp, u32, vp, vp]),
        "vt_nrc_train": (i32, [vp, vp, vp, u32, c.c_float, vp]),
        "vt_nrc_gradients": (i32, [vp, vp, vp, u32, vp, vp]),
        "vt_renderer_set_cache": (i32, [vp, vp]),
        "vt_scene_rebuild": (i32, [vp, P(_BuildOptions)]),
        "vt_batch_create": (i32, [vp, P(BatchJob), u32, u32, u32, P(vp)]),
        "vt_batch_destroy": (None, [vp]),
        "vt_batch_render": (i32, [vp, u32, vp]),
        "vt_batch_reset": (i32, [vp]),
        "vt_batch_frames": (u32, [vp]),
        "vt_batch_readback": (i32, [vp, u32, vp, c.c_size_t]),
        "vt_batch_get_tuning": (i32, [vp, P(BatchTuning)]),
        "vt_batch_set_tuning": (i32, [vp, P(BatchTuning)]),
        "vt_batch_autotune": (i32, [vp, u32, P(BatchTuning), P(c.c_double)]),
        "vt_renderer_cache_records": (i32, [vp, vp, vp, u32, P(u32)]),
        "vt_accum_readback": (i32, [vp, vp, c.c_size_t]),
        "vt_accum_frames": (u32, [vp]),
        "vt_spectral_tables": (i32, [vp, vp, vp]),
        "vt_spectral_reflectance": (i32, [P(f32), P(f32)]),
        "vt_scene_set_light_tree": (i32, [vp, c.c_int]),
        "vt_scene_set_environment": (i32, [vp, vp, u32, u32]),
        "vt_environment_tables": (i32, [vp, u32, u32, vp, vp]),
        "vt_hierarchy_build_host": (i32, [vp, u32, vp, u32, P(c.c_void_p), P(c.c_void_p), P(u32), P(u32)]),
        "vt_light_tree_build_host": (i32, [vp, vp, u32, P(vp), P(u32), P(vp), P(vp), P(u32), P(f32), P(f32)]),
        "vt_mesh_payload_bytes": (c.c_uint64, [u32, u32]),
        "vt_mesh_payload_write": (i32, [vp, u32, vp, u32, vp, c.c_uint64]),
        "vt_mesh_payload_read": (i32, [vp, c.c_uint64, u32, P(vp), P(u32), P(vp), P(u32)]),
        "vt_obj_parse": (i32, [c.c_char_p, c.c_uint64, u32, c.c_uint64, P(vp)]),
        "vt_obj_destroy": (None, [vp]),
        "vt_obj_part_count": (u32, [vp]),
        "vt_obj_skipped": (u32, [vp]),
        "vt_obj_library_count": (u32, [vp]),
        "vt_obj_library": (c.c_char_p, [vp, u32]),
        "vt_obj_part": (i32, [vp, u32, P(c.c_char_p), P(c.c_char_p), P(u32), P(vp), P(u32), P(vp), P(u32)]),
        "vt_obj_subset": (c.c_char_p, [vp, u32, u32]),
        "vt_z_draws": (i32, [P(ZQuery), vp, vp]),
        "vt_context_check": (i32, [vp]),
        "vt_mesh_from_triangles": (i32, [vp, u32, P(vp), P(u32), P(vp), P(u32)]),
        "vt_scene_upload_triangles": (i32, [vp, vp, u32, P(_BuildOptions), P(vp)]),
        "vt_scene_set_surface_settings": (i32, [vp, vp, u32, vp, u32]),
        "vt_renderer_set_callbacks": (i32, [vp, FRAME_FN, ERROR_FN, vp]),
        "vt_trace_stats_noalloc": (i32, [vp, P(_RaySoA), u32, c.c_int, P(_HitSoA), vp, vp, vp]),
        "vt_scene_set_stack_cap": (i32, [vp, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int):
    if status != VT_OK:
        raise VTError(status, load().vt_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------------------------
# Workloads (host arrays)
# --------------------------------------------------------------------------------------------

@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 8) float32: position, normal, uv
    faces: np.ndarray     # (F, 4) uint32: a, b, c, tag


def hierarchy_build_host(vertices, faces):
    """vt_hierarchy_build_host: the flat binary median hierarchy (build option collapse 2) ->
    (order (F,) uint32, nodes (N, 12) with lo[4], hi[4] as float32 and links[4] as int32, depth)."""
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(faces, np.uint32)
    order, nodes = ctypes.c_void_p(), ctypes.c_void_p()
    count, depth = ctypes.c_uint32(), ctypes.c_uint32()
    _check(load().vt_hierarchy_build_host(_ptr(v), len(v), _ptr(f), len(f), ctypes.byref(order), ctypes.byref(nodes),
                                          ctypes.byref(count), ctypes.byref(depth)))
    try:
        o = np.ctypeslib.as_array(ctypes.cast(order, ctypes.POINTER(ctypes.c_uint32)), (len(f),)).copy() if len(f) else \
            np.zeros(0, np.uint32)
        raw = np.ctypeslib.as_array(ctypes.cast(nodes, ctypes.POINTER(ctypes.c_uint32)), (count.value * 12,)).copy() \
            if count.value else np.zeros(0, np.uint32)
    finally:
        load().vt_host_free(order)
        load().vt_host_free(nodes)
    raw = raw.reshape(-1, 12)
    return o, {"lo": raw[:, 0:4].view(np.float32), "hi": raw[:, 4:8].view(np.float32),
               "links": raw[:, 8:12].view(np.int32)}, depth.value


def environment_tables(rgba):
    """vt_environment_tables: the library's sampling tables of an (H, W, 4) float32 map -> (rows (H, 2),
    cols (2, H, W)): per table (0 MIS-compensated, 1 original) the marginal and conditional CDFs."""
    img = np.ascontiguousarray(rgba, np.float32)
    h, w = img.shape[:2]
    rows = np.zeros((h, 2), np.float32)
    cols = np.zeros((2, h, w), np.float32)
    _check(load().vt_environment_tables(_ptr(img), w, h, _ptr(rows), _ptr(cols)))
    return rows, cols


def icosphere_field(spheres=1024, subdivisions=3, mu=0.0, sigma=0.5, half_extent=50.0, seed=7) -> Mesh:
    lib = load()
    nv, nf = ctypes.c_uint32(), ctypes.c_uint32()
    lib.vt_gen_icosphere_field(spheres, subdivisions, mu, sigma, half_extent, seed, None, ctypes.byref(nv), None,
                               ctypes.byref(nf))
    v = np.empty((nv.value, 8), np.float32)
    f = np.empty((nf.value, 4), np.uint32)
    rc = lib.vt_gen_icosphere_field(spheres, subdivisions, mu, sigma, half_extent, seed, _ptr(v), ctypes.byref(nv),
                                    _ptr(f), ctypes.byref(nf))
    if rc != 0:
        raise ValueError("invalid icosphere field parameters")
    return Mesh(v, f)


def heightfield(grid=2237, octaves=8, gain=0.5, half_extent=100.0, height=12.0, seed=3) -> Mesh:
    lib = load()
    nv, nf = ctypes.c_uint32(), ctypes.c_uint32()
    if lib.vt_gen_heightfield(grid, octaves, gain, half_extent, height, seed, None, ctypes.byref(nv), None,
                              ctypes.byref(nf)) != 0:
        raise ValueError("invalid heightfield parameters")
    v = np.empty((nv.value, 8), np.float32)
    f = np.empty((nf.value, 4), np.uint32)
    lib.vt_gen_heightfield(grid, octaves, gain, half_extent, height, seed, _ptr(v), ctypes.byref(nv), _ptr(f),
                           ctypes.byref(nf))
    return Mesh(v, f)


def cornell(light_size=0.25) -> Mesh:
    """Config-1 Cornell box (tags: 0 white, 1 red, 2 green, 3 light)."""
    lib = load()
    nv, nf = ctypes.c_uint32(), ctypes.c_uint32()
    lib.vt_gen_cornell(light_size, None, ctypes.byref(nv), None, ctypes.byref(nf))
    v = np.empty((nv.value, 8), np.float32)
    f = np.empty((nf.value, 4), np.uint32)
    lib.vt_gen_cornell(light_size, _ptr(v), ctypes.byref(nv), _ptr(f), ctypes.byref(nf))
    return Mesh(v, f)


def cornell_materials(light_scale=17.0):
    """White 0.73, red (0.65, 0.05, 0.05), green (0.12, 0.45, 0.15) (main.swift:1422-1424) and a D65 light of
    RGB radiance light_scale."""
    return [material(0, (0.73, 0.73, 0.73)), material(0, (0.65, 0.05, 0.05)), material(0, (0.12, 0.45, 0.15)),
            material(3, (0.0, 0.0, 0.0), (light_scale, light_scale, light_scale))]


def orbit_eye(yaw=0.45, pitch=0.25, distance=3.4, target=(0.0, -0.1, 0.0)):
    """The reference's orbit camera position (main.swift:11303-11305), Cornell default view (:10755)."""
    import math
    return (target[0] + distance * math.cos(pitch) * math.sin(yaw), target[1] + distance * math.sin(pitch),
            target[2] - distance * math.cos(pitch) * math.cos(yaw))


def light_tree_build_host(tris: np.ndarray, power: np.ndarray):
    """(n, 9) float32 emitter corners, (n,) float64 powers -> dict(words (m, 8), order, leaf_node, depth, grid)."""
    lib = load()
    t = np.ascontiguousarray(tris, np.float32)
    pw = np.ascontiguousarray(power, np.float64)
    n = len(t)
    words, order, leaf = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    count, depth = ctypes.c_uint32(), ctypes.c_uint32()
    lo, hi = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
    _check(lib.vt_light_tree_build_host(_ptr(t), _ptr(pw), n, ctypes.byref(words), ctypes.byref(count),
                                        ctypes.byref(order), ctypes.byref(leaf), ctypes.byref(depth), lo, hi))
    try:
        w = np.frombuffer(ctypes.string_at(words, 32 * count.value), np.uint32).reshape(-1, 8).copy()
        o = np.frombuffer(ctypes.string_at(order, 4 * n), np.uint32).copy() if n else np.zeros(0, np.uint32)
        ln = np.frombuffer(ctypes.string_at(leaf, 4 * n), np.uint32).copy() if n else np.zeros(0, np.uint32)
    finally:
        for ptr in (words, order, leaf):
            lib.vt_host_free(ptr)
    return {"words": w, "order": o, "leaf_node": ln, "depth": depth.value, "grid_lo": np.array(lo[:], np.float32),
            "grid_hi": np.array(hi[:], np.float32)}


def spectral_tables():
    """(icdf (1025,), cmf (471, 4), basis (471, 2, 4)) float32, as the integrator reads them."""
    icdf = np.empty(1025, np.float32)
    cmf = np.empty((471, 4), np.float32)
    basis = np.empty((471, 2, 4), np.float32)
    _check(load().vt_spectral_tables(_ptr(icdf), _ptr(cmf), _ptr(basis)))
    return icdf, cmf, basis


def spectral_reflectance(rgb):
    out = (ctypes.c_float * 4)()
    _check(load().vt_spectral_reflectance((ctypes.c_float * 3)(*rgb), out))
    return np.array(out[:], np.float32)


RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def _ray_args(r: np.ndarray):
    return [_ptr(r[k]) for k in range(8)]


def camera_rays(eye, target, fov, width, height, seed) -> np.ndarray:
    """(8, width*height) float32 SoA rays."""
    r = np.empty((8, width * height), np.float32)
    e = (ctypes.c_float * 3)(*eye)
    t = (ctypes.c_float * 3)(*target)
    load().vt_gen_camera_rays(e, t, fov, width, height, seed, *_ray_args(r))
    return r


def bounce_rays(mesh: Mesh, rays: np.ndarray, face: np.ndarray, b1: np.ndarray, b2: np.ndarray, seed: int):
    n = rays.shape[1]
    out = np.empty((8, n), np.float32)
    face = np.ascontiguousarray(face, np.int32)
    b1 = np.ascontiguousarray(b1, np.float32)
    b2 = np.ascontiguousarray(b2, np.float32)
    d = [np.ascontiguousarray(rays[k]) for k in (3, 4, 5)]
    hits = load().vt_gen_bounce_rays(_ptr(mesh.vertices), _ptr(mesh.faces), n, _ptr(d[0]), _ptr(d[1]), _ptr(d[2]),
                                     _ptr(face), _ptr(b1), _ptr(b2), seed, *_ray_args(out))
    if hits == 0:
        raise ValueError("no hits to bounce from")
    return out


def random_rays(lo, hi, n, seed) -> np.ndarray:
    r = np.empty((8, n), np.float32)
    load().vt_gen_random_rays((ctypes.c_float * 3)(*lo), (ctypes.c_float * 3)(*hi), n, seed, *_ray_args(r))
    return r


# --------------------------------------------------------------------------------------------
# Mesh formats (include/vt_mesh.h): the VTMESH01 payload and the OBJ reader
# --------------------------------------------------------------------------------------------

def mesh_payload(mesh: Mesh) -> bytes:
    """The VTMESH01 payload of an indexed mesh."""
    lib = load()
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    f = np.ascontiguousarray(mesh.faces, np.uint32)
    n = lib.vt_mesh_payload_bytes(len(v), len(f))
    out = np.empty(n, np.uint8)
    _check(lib.vt_mesh_payload_write(_ptr(v), len(v), _ptr(f), len(f), _ptr(out), n))
    return out.tobytes()


def read_mesh_payload(data: bytes, maximum_faces=1 << 30) -> Mesh:
    lib = load()
    buf = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
    vp, fp = ctypes.c_void_p(), ctypes.c_void_p()
    nv, nf = ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib.vt_mesh_payload_read(_ptr(buf), len(data), maximum_faces, ctypes.byref(vp), ctypes.byref(nv),
                                    ctypes.byref(fp), ctypes.byref(nf)))
    try:
        v = np.frombuffer(ctypes.string_at(vp, 32 * nv.value), np.float32).reshape(-1, 8).copy()
        f = np.frombuffer(ctypes.string_at(fp, 16 * nf.value), np.uint32).reshape(-1, 4).copy()
    finally:
        lib.vt_host_free(vp)
        lib.vt_host_free(fp)
    return Mesh(v, f)


def mesh_from_triangles(triangles) -> Mesh:
    """MeshGeometry(triangles:): (n, 32) float32 MeshTriangle records -> welded indexed Mesh."""
    lib = load()
    t = np.ascontiguousarray(triangles, np.float32).reshape(-1, 32)
    vp, fp = ctypes.c_void_p(), ctypes.c_void_p()
    nv, nf = ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib.vt_mesh_from_triangles(_ptr(t) if len(t) else None, len(t), ctypes.byref(vp), ctypes.byref(nv),
                                      ctypes.byref(fp), ctypes.byref(nf)))
    try:
        v = np.frombuffer(ctypes.string_at(vp, 32 * nv.value), np.float32).reshape(-1, 8).copy()
        f = np.frombuffer(ctypes.string_at(fp, 16 * nf.value), np.uint32).reshape(-1, 4).copy()
    finally:
        lib.vt_host_free(vp)
        lib.vt_host_free(fp)
    return Mesh(v, f)


def mesh_triangles(mesh: Mesh) -> np.ndarray:
    """The MeshTriangle records (n, 32) of an indexed mesh (MeshTriangle(face:vertices:)): corners w = 1,
    normals w = 0, uvc = (uv of c, subset, 0)."""
    v, f = mesh.vertices, mesh.faces
    t = np.zeros((len(f), 32), np.float32)
    for k in range(3):
        c = v[f[:, k]]
        t[:, 4 * k:4 * k + 3] = c[:, :3]
        t[:, 4 * k + 3] = 1.0
        t[:, 12 + 4 * k:15 + 4 * k] = c[:, 3:6]
    t[:, 24:26] = v[f[:, 0], 6:8]
    t[:, 26:28] = v[f[:, 1], 6:8]
    t[:, 28:30] = v[f[:, 2], 6:8]
    t[:, 30] = f[:, 3].astype(np.float32)
    return t


class SurfaceSettings(ctypes.Structure):
    """SurfaceSettings (main.swift:2253-2263), 64 bytes, with the reference's host defaults."""
    _fields_ = [("color", ctypes.c_float * 4), ("surface", ctypes.c_float * 4), ("detail", ctypes.c_float * 4),
                ("enabled", ctypes.c_uint32), ("map_mask", ctypes.c_uint32), ("normal_strength", ctypes.c_float),
                ("padding", ctypes.c_uint32)]

    def __init__(self, color=(1.0, 1.0, 1.0), roughness=0.3, metalness=0.0, coat=0.0, anisotropy=0.0, fuzz=0.0,
                 ior=1.5, transmission=0.0, uv_repeat=1.0, enabled=1, map_mask=0, normal_strength=1.0):
        super().__init__((ctypes.c_float * 4)(*color[:3], 1.0), (ctypes.c_float * 4)(roughness, metalness, coat, anisotropy),
                         (ctypes.c_float * 4)(fuzz, ior, transmission, uv_repeat), enabled, map_mask, normal_strength, 0)


assert ctypes.sizeof(SurfaceSettings) == 64 and ctypes.sizeof(Material) == 64


class Emission(ctypes.Structure):
    _fields_ = [("slot", ctypes.c_uint32), ("rgb", ctypes.c_float * 3)]


def read_obj(text, node_budget=0, triangle_limit=0):
    """Parses OBJ text (str or bytes).  Returns (parts, skipped, libraries); each part is a dict with
    object, group, subsets (material names) and mesh (welded vertices, faces tagged by subset)."""
    lib = load()
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    _check(lib.vt_obj_parse(data, len(data), node_budget, triangle_limit, ctypes.byref(h)))
    try:
        parts = []
        for k in range(lib.vt_obj_part_count(h)):
            obj, grp = ctypes.c_char_p(), ctypes.c_char_p()
            ns, nv, nf = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
            vp, fp = ctypes.c_void_p(), ctypes.c_void_p()
            _check(lib.vt_obj_part(h, k, ctypes.byref(obj), ctypes.byref(grp), ctypes.byref(ns), ctypes.byref(vp),
                                   ctypes.byref(nv), ctypes.byref(fp), ctypes.byref(nf)))
            v = np.frombuffer(ctypes.string_at(vp, 32 * nv.value), np.float32).reshape(-1, 8).copy()
            f = np.frombuffer(ctypes.string_at(fp, 16 * nf.value), np.uint32).reshape(-1, 4).copy()
            parts.append(dict(object=obj.value.decode(), group=grp.value.decode(), mesh=Mesh(v, f),
                              subsets=[lib.vt_obj_subset(h, k, i).decode() for i in range(ns.value)]))
        libraries = [lib.vt_obj_library(h, i).decode() for i in range(lib.vt_obj_library_count(h))]
        return parts, int(lib.vt_obj_skipped(h)), libraries
    finally:
        lib.vt_obj_destroy(h)


def merge_parts(parts) -> Mesh:
    """One indexed mesh of OBJ parts (faces keep their part-local subset)."""
    vs, fs, base = [], [], 0
    for p in parts:
        m = p["mesh"]
        f = m.faces.copy()
        f[:, :3] += base
        vs.append(m.vertices)
        fs.append(f)
        base += len(m.vertices)
    return Mesh(np.concatenate(vs), np.concatenate(fs))


# --------------------------------------------------------------------------------------------
# Host-only BVH inspection
# --------------------------------------------------------------------------------------------

NODE_DTYPE = np.dtype([("p", "<f4", 3), ("e", "u1", 3), ("imask", "u1"), ("child_base", "<u4"), ("tri_base", "<u4"),
                       ("meta", "u1", 8), ("qlo", "u1", (3, 8)), ("qhi", "u1", (3, 8))])
TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("face", "<u4"), ("v1", "<f4", 3), ("pad1", "<u4"), ("v2", "<f4", 3),
                      ("pad2", "<u4")])
assert NODE_DTYPE.itemsize == 80 and TRI_DTYPE.itemsize == 48


def _opts(leaf_size=0, bins=0, max_depth=0, collapse=0):
    return _BuildOptions(leaf_size, bins, max_depth, collapse)


def build_bvh8_host(mesh: Mesh, leaf_size=0, bins=0, max_depth=0, collapse=0):
    """Returns (nodes structured array, tris structured array, depth, root_max)."""
    lib = load()
    nodes, tris = ctypes.c_void_p(), ctypes.c_void_p()
    count, depth, root_max = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_float()
    o = _opts(leaf_size, bins, max_depth, collapse)
    _check(lib.vt_bvh8_build_host(_ptr(mesh.vertices), len(mesh.vertices), _ptr(mesh.faces), len(mesh.faces),
                                  ctypes.byref(o), ctypes.byref(nodes), ctypes.byref(count), ctypes.byref(tris),
                                  ctypes.byref(depth), ctypes.byref(root_max)))
    try:
        nb = count.value * 80
        tb = len(mesh.faces) * 48
        n = np.frombuffer(ctypes.string_at(nodes, nb), NODE_DTYPE).copy() if nb else np.zeros(0, NODE_DTYPE)
        t = np.frombuffer(ctypes.string_at(tris, tb), TRI_DTYPE).copy() if tb else np.zeros(0, TRI_DTYPE)
    finally:
        lib.vt_host_free(nodes)
        lib.vt_host_free(tris)
    return n, t, depth.value, root_max.value


# --------------------------------------------------------------------------------------------
# Device context, scenes and traces
# --------------------------------------------------------------------------------------------

class Context:
    """One HIP device (vt_context)."""

    def __init__(self, device: int = 0, memory_budget: int = 0):
        lib = load()
        h = ctypes.c_void_p()
        desc = _ContextDesc(device, memory_budget)
        _check(lib.vt_context_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load().vt_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self):
        """vt_context_check: waits for the device; raises VTError for a fault its kernels reported."""
        _check(load().vt_context_check(self._h))

    def energy_tables(self):
        """(metal (129, 129), metal averages (129,), luts (5, 32, 32, 32)) as baked on the device."""
        metal = np.empty(129 * 129 + 129, np.float32)
        luts = np.empty((5, 32, 32, 32), np.float32)
        _check(load().vt_energy_tables(self._h, _ptr(metal), _ptr(luts)))
        return metal[:129 * 129].reshape(129, 129), metal[129 * 129:], luts

    def uplift_grid(self, capacity=8192):
        """The reflectance-uplift grid built on the device (vt_uplift_grid_build): coarse moments
        (86, 86, 86, 3), refinement flags (85, 85, 85) bool, flagged cell indices (n,) and their codes'
        moments (n, 4, 4, 4, 3), n = min(count, capacity), and the flagged-cell count."""
        coarse = np.empty((86, 86, 86, 3), np.float32)
        flags = np.empty((85, 85, 85), np.uint8)
        cells = np.empty(capacity, np.uint32)
        refined = np.empty((capacity, 4, 4, 4, 3), np.float32)
        count = ctypes.c_uint32(0)
        _check(load().vt_uplift_grid_build(self._h, _ptr(coarse), _ptr(flags), capacity, _ptr(cells), _ptr(refined),
                                           ctypes.byref(count)))
        n = min(count.value, capacity)
        return coarse, flags.astype(bool), cells[:n].copy(), refined[:n].copy(), count.value

    def set_mesh(self, mesh: Mesh, leaf_size=0, bins=0, max_depth=0, collapse=0) -> "Scene":
        """MaterialLibrary.setMesh for an indexed mesh: builds and uploads the BVH."""
        lib = load()
        v = np.ascontiguousarray(mesh.vertices, np.float32)
        f = np.ascontiguousarray(mesh.faces, np.uint32)
        if v.ndim != 2 or v.shape[1] != 8 or f.ndim != 2 or f.shape[1] != 4:
            raise ValueError("vertices must be (V, 8) float32 and faces (F, 4) uint32")
        h = ctypes.c_void_p()
        o = _opts(leaf_size, bins, max_depth, collapse)
        _check(lib.vt_scene_upload_indexed(self._h, _ptr(v), len(v), _ptr(f), len(f), ctypes.byref(o),
                                           ctypes.byref(h)))
        return Scene(self, h)

    def set_triangles(self, triangles, leaf_size=0, bins=0, max_depth=0, collapse=0) -> "Scene":
        """MaterialLibrary.setMesh(_ triangles: [MeshTriangle]): (n, 32) float32 records."""
        t = np.ascontiguousarray(triangles, np.float32).reshape(-1, 32)
        h = ctypes.c_void_p()
        o = _opts(leaf_size, bins, max_depth, collapse)
        _check(load().vt_scene_upload_triangles(self._h, _ptr(t) if len(t) else None, len(t), ctypes.byref(o),
                                                ctypes.byref(h)))
        return Scene(self, h)


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


def _rays_struct(rays):
    import torch
    if not (isinstance(rays, torch.Tensor) and rays.is_cuda and rays.dtype == torch.float32 and rays.dim() == 2
            and rays.shape[0] == 8 and rays.is_contiguous()):
        raise ValueError("rays must be a contiguous (8, n) float32 device tensor")
    n = rays.shape[1]
    base = rays.data_ptr()
    return _RaySoA(*[base + 4 * n * k for k in range(8)]), n


class Scene:
    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle

    def close(self):
        if self._h:
            load().vt_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def info(self) -> dict:
        i = _SceneInfo()
        _check(load().vt_scene_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _SceneInfo._fields_}

    def trace_closest(self, rays, out=None, stream=None):
        """rays: (8, n) float32 device tensor.  Returns dict of device tensors t, b1, b2, face."""
        import torch
        rs, n = _rays_struct(rays)
        if out is None:
            out = {"t": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "b1": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "b2": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "face": torch.empty(n, dtype=torch.int32, device=rays.device)}
        hs = _HitSoA(out["t"].data_ptr(), out["b1"].data_ptr(), out["b2"].data_ptr(), out["face"].data_ptr(), None)
        _check(load().vt_trace_closest(self._h, ctypes.byref(rs), n, ctypes.byref(hs), _stream_ptr(stream)))
        return out

    def trace_any(self, rays, out=None, stream=None):
        import torch
        rs, n = _rays_struct(rays)
        if out is None:
            out = torch.empty(n, dtype=torch.uint8, device=rays.device)
        _check(load().vt_trace_any(self._h, ctypes.byref(rs), n, ctypes.c_void_p(out.data_ptr()),
                                   _stream_ptr(stream)))
        return out

    def set_light_tree(self, enabled: bool):
        _check(load().vt_scene_set_light_tree(self._h, 1 if enabled else 0))

    def set_environment(self, rgba):
        """vt_scene_set_environment: (H, W, 4) float32 linear rgba, row 0 at the zenith; None removes it."""
        if rgba is None:
            _check(load().vt_scene_set_environment(self._h, None, 0, 0))
            return
        img = np.ascontiguousarray(rgba, np.float32)
        if img.ndim != 3 or img.shape[2] != 4:
            raise ValueError("environment map must be (height, width, 4)")
        _check(load().vt_scene_set_environment(self._h, _ptr(img), img.shape[1], img.shape[0]))

    def set_stack_cap(self, groups: int):
        """Test hook (vt_scene_set_stack_cap): caps the traversal stack at `groups` node groups (0: none)."""
        _check(load().vt_scene_set_stack_cap(self._h, groups))

    def rebuild(self, leaf_size=0, bins=0, max_depth=0, collapse=0):
        """vt_scene_rebuild: the hierarchy again from the scene's own streams with other options."""
        _check(load().vt_scene_rebuild(self._h, ctypes.byref(_opts(leaf_size, bins, max_depth, collapse))))

    def set_materials(self, materials):
        arr = (Material * len(materials))(*materials)
        _check(load().vt_scene_set_materials(self._h, arr, len(materials)))

    def set_surface_settings(self, settings, emissions=None):
        """settings: list of SurfaceSettings per slot; emissions: {slot: (r, g, b)}."""
        s = (SurfaceSettings * max(1, len(settings)))(*settings)
        em = [Emission(k, (ctypes.c_float * 3)(*rgb)) for k, rgb in (emissions or {}).items()]
        e = (Emission * max(1, len(em)))(*em)
        _check(load().vt_scene_set_surface_settings(self._h, s, len(settings), e, len(em)))

    def trace_stats_noalloc(self, rays, counters, out=None, any_hit=False, stream=None):
        """vt_trace_stats_noalloc: counters (2 int64 device tensor) <- (node fetches, triangle tests);
        results into `out` (closest: dict t, b1, b2, face; any: uint8 tensor), allocated here when None."""
        import torch
        rs, n = _rays_struct(rays)
        if counters.dtype != torch.int64 or not counters.is_cuda or not counters.is_contiguous() or counters.numel() < 2:
            raise ValueError("counters must be a contiguous int64 device tensor of 2 elements")
        if any_hit:
            if out is None:
                out = torch.empty(n, dtype=torch.uint8, device=rays.device)
            _check(load().vt_trace_stats_noalloc(self._h, ctypes.byref(rs), n, 1, None, ctypes.c_void_p(out.data_ptr()),
                                                 ctypes.c_void_p(counters.data_ptr()), _stream_ptr(stream)))
            return out
        if out is None:
            out = {"t": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "b1": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "b2": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "face": torch.empty(n, dtype=torch.int32, device=rays.device)}
        hs = _HitSoA(out["t"].data_ptr(), out["b1"].data_ptr(), out["b2"].data_ptr(), out["face"].data_ptr(), None)
        _check(load().vt_trace_stats_noalloc(self._h, ctypes.byref(rs), n, 0, ctypes.byref(hs), None,
                                             ctypes.c_void_p(counters.data_ptr()), _stream_ptr(stream)))
        return out

    def trace_stats(self, rays, any_hit=False, stream=None):
        rs, n = _rays_struct(rays)
        c = (ctypes.c_uint64 * 2)()
        _check(load().vt_trace_stats(self._h, ctypes.byref(rs), n, 1 if any_hit else 0, c, _stream_ptr(stream)))
        return int(c[0]), int(c[1])


FRAME_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint32)
ERROR_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p)


class Renderer:
    """vt_renderer: the wavefront path tracer and its running-mean accumulation."""

    def __init__(self, ctx: Context):
        h = ctypes.c_void_p()
        _check(load().vt_renderer_create(ctx._h, ctypes.byref(h)))
        self.ctx = ctx
        self._h = h
        self.size = (0, 0)

    def close(self):
        if self._h:
            load().vt_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, scene: Scene, width, height, eye, target, fov, max_depth=8, samples=1, sample_index=1,
               sampling_mode=1, candidates=4, cache=False, train_period=16, learning_rate=1e-3, material_sort=False,
               sampler=0, z_temporal=1, indirect_reuse=0, env_strength=0.0, env_rotation=0.0, stream=None):
        p = _FrameParams(width, height, (ctypes.c_float * 3)(*eye), (ctypes.c_float * 3)(*target), fov, max_depth,
                         sampling_mode, sample_index, candidates, int(cache), train_period, learning_rate,
                         int(material_sort), sampler, z_temporal, indirect_reuse, env_strength, env_rotation)
        _check(load().vt_render_frame(self._h, scene._h, ctypes.byref(p), samples, _stream_ptr(stream)))
        self.size = (width, height)

    def reset(self):
        _check(load().vt_reset_accumulation(self._h))

    def set_callbacks(self, on_frame=None, on_error=None):
        """vt_renderer_set_callbacks: on_frame(frames) after each render call's samples (on a runtime
        thread), on_error(message) when a call fails."""
        f = FRAME_FN((lambda user, frames: on_frame(frames)) if on_frame else 0)
        e = ERROR_FN((lambda user, msg: on_error(msg.decode())) if on_error else 0)
        self._callbacks = (f, e)          # kept alive while attached
        _check(load().vt_renderer_set_callbacks(self._h, f, e, None))

    def cache_records(self, max_records=1 << 22):
        """(inputs (n, 11), targets (n, 4), made) of the last wavefront's cache training records."""
        x = np.empty((max_records, 11), np.float32)
        t = np.empty((max_records, 4), np.float32)
        made = ctypes.c_uint32()
        _check(load().vt_renderer_cache_records(self._h, _ptr(x), _ptr(t), max_records, ctypes.byref(made)))
        n = min(made.value, max_records)
        return x[:n], t[:n], made.value

    def set_cache(self, cache):
        """Attach a vt.NRC (None detaches); keep the cache alive while attached."""
        self._cache = cache
        _check(load().vt_renderer_set_cache(self._h, cache._h if cache is not None else None))

    @property
    def frames(self) -> int:
        return int(load().vt_accum_frames(self._h))

    def readback(self) -> np.ndarray:
        w, h = self.size
        out = np.empty((h, w, 4), np.float32)
        _check(load().vt_accum_readback(self._h, _ptr(out), out.nbytes))
        return out


NRC_PARAMS = 30720


class Batch:
    """vt_batch: many images of many scenes rendered together by one captured persistent wavefront.
    jobs: [(scene, width, height, eye, target, fov)]."""

    def __init__(self, ctx: Context, jobs, max_depth=6, sampling_mode=1):
        arr = (BatchJob * len(jobs))()
        self.sizes = []
        for k, (scene, w, h, eye, target, fov) in enumerate(jobs):
            arr[k] = BatchJob(scene._h.value if hasattr(scene._h, "value") else scene._h, w, h,
                              (ctypes.c_float * 3)(*eye), (ctypes.c_float * 3)(*target), fov)
            self.sizes.append((w, h))
        self._scenes = [j[0] for j in jobs]  # kept alive with the batch
        h = ctypes.c_void_p()
        _check(load().vt_batch_create(ctx._h, arr, len(jobs), max_depth, sampling_mode, ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load().vt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, samples=1, stream=None):
        _check(load().vt_batch_render(self._h, samples, _stream_ptr(stream)))

    def reset(self):
        _check(load().vt_batch_reset(self._h))

    @property
    def frames(self) -> int:
        return int(load().vt_batch_frames(self._h))

    def readback(self, job) -> np.ndarray:
        w, h = self.sizes[job]
        out = np.empty((h, w, 4), np.float32)
        _check(load().vt_batch_readback(self._h, job, _ptr(out), out.nbytes))
        return out

    @property
    def tuning(self) -> dict:
        t = BatchTuning()
        _check(load().vt_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    @tuning.setter
    def tuning(self, d):
        t = BatchTuning(d.get("capacity", 0), d.get("grid_waves", 4), d.get("trace_block", 128), d.get("leaf_size", 0))
        _check(load().vt_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, leaf_size=False):
        """Runs the autotuner; returns (best tuning, milliseconds per sample)."""
        t = BatchTuning()
        ms = ctypes.c_double()
        _check(load().vt_batch_autotune(self._h, 1 if leaf_size else 0, ctypes.byref(t), ctypes.byref(ms)))
        return t.as_dict(), ms.value


class Tlas:
    """A two-level scene (vt_tlas): instances = [(scene, world)], world a (3, 4) object-to-world
    matrix.  The scenes must stay open while the Tlas is used."""

    def __init__(self, ctx: "Context", instances):
        arr = (Instance * max(1, len(instances)))()
        self._scenes = []
        for k, (scene, world) in enumerate(instances):
            w = np.ascontiguousarray(world, np.float32).reshape(12)
            arr[k].mesh = scene._h.value
            arr[k].world[:] = w.tolist()
            self._scenes.append(scene)
        h = ctypes.c_void_p()
        _check(load().vt_tlas_create(ctx._h, arr, len(instances), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load().vt_tlas_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def info(self) -> dict:
        i = _TlasInfo()
        _check(load().vt_tlas_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _TlasInfo._fields_}

    def trace_closest(self, rays, out=None, stream=None):
        """rays: (8, n) float32 device tensor.  Returns device tensors t, b1, b2, face, instance."""
        import torch
        rs, n = _rays_struct(rays)
        if out is None:
            out = {"t": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "b1": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "b2": torch.empty(n, dtype=torch.float32, device=rays.device),
                   "face": torch.empty(n, dtype=torch.int32, device=rays.device),
                   "instance": torch.empty(n, dtype=torch.int32, device=rays.device)}
        hs = _HitSoA(out["t"].data_ptr(), out["b1"].data_ptr(), out["b2"].data_ptr(), out["face"].data_ptr(),
                     out["instance"].data_ptr())
        _check(load().vt_tlas_trace_closest(self._h, ctypes.byref(rs), n, ctypes.byref(hs), _stream_ptr(stream)))
        return out

    def trace_any(self, rays, out=None, stream=None):
        import torch
        rs, n = _rays_struct(rays)
        if out is None:
            out = torch.empty(n, dtype=torch.uint8, device=rays.device)
        _check(load().vt_tlas_trace_any(self._h, ctypes.byref(rs), n, ctypes.c_void_p(out.data_ptr()),
                                        _stream_ptr(stream)))
        return out


class NRC:
    """The learned radiance cache (vt_nrc_*): inputs (n, 11) and outputs (n, 4) are float32 device
    tensors; parameters move as host numpy arrays."""

    def __init__(self, ctx: "Context", seed: int = 1):
        h = ctypes.c_void_p()
        _check(load().vt_nrc_create(ctx._h, seed, ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            load().vt_nrc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.empty(NRC_PARAMS, np.float32)
        _check(load().vt_nrc_get_params(self._h, _ptr(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (NRC_PARAMS,)
        _check(load().vt_nrc_set_params(self._h, _ptr(p)))

    @staticmethod
    def _dev(t, cols):
        import torch
        assert isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
        assert t.dim() == 2 and t.shape[1] == cols
        return ctypes.c_void_p(t.data_ptr()), t.shape[0]

    def infer(self, inputs, out=None, stream=None):
        import torch
        ip, n = self._dev(inputs, 11)
        if out is None:
            out = torch.empty((n, 4), dtype=torch.float32, device=inputs.device)
        op, _ = self._dev(out, 4)
        _check(load().vt_nrc_infer(self._h, ip, n, op, _stream_ptr(stream)))
        return out

    def train(self, inputs, targets, lr=1e-3, stream=None):
        ip, n = self._dev(inputs, 11)
        tp, m = self._dev(targets, 4)
        assert m == n
        _check(load().vt_nrc_train(self._h, ip, tp, n, lr, _stream_ptr(stream)))

    def gradients(self, inputs, targets):
        ip, n = self._dev(inputs, 11)
        tp, _ = self._dev(targets, 4)
        g = np.empty(NRC_PARAMS, np.float32)
        loss = ctypes.c_float()
        _check(load().vt_nrc_gradients(self._h, ip, tp, n, _ptr(g), ctypes.byref(loss)))
        return g, loss.value
