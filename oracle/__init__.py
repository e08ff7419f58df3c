"""ctypes wrapper of the oracle (oracle/liboracle.so): TEST INFRASTRUCTURE ONLY.

The oracle restates the reference's traversal arithmetic on the CPU (see oracle.h for the
reference file:line of each function and what pins it).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker or the timed CPU
baseline; the shipped library never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
_lib_path = None


class WideNode(ctypes.Structure):
    _fields_ = [("lox", ctypes.c_float * 4), ("loy", ctypes.c_float * 4), ("loz", ctypes.c_float * 4),
                ("hix", ctypes.c_float * 4), ("hiy", ctypes.c_float * 4), ("hiz", ctypes.c_float * 4),
                ("child", ctypes.c_int32 * 4), ("count", ctypes.c_int32 * 4)]


WIDE_NODE_DTYPE = np.dtype([("lo", "<f4", (3, 4)), ("hi", "<f4", (3, 4)), ("child", "<i4", 4), ("count", "<i4", 4)])
assert ctypes.sizeof(WideNode) == 128 and WIDE_NODE_DTYPE.itemsize == 128


class Hit(ctypes.Structure):
    _fields_ = [("t", ctypes.c_float), ("b1", ctypes.c_float), ("b2", ctypes.c_float), ("face", ctypes.c_int32)]


class SceneView(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("faces", ctypes.c_void_p), ("face_count", ctypes.c_uint32),
                ("nodes", ctypes.c_void_p), ("order", ctypes.c_void_p), ("root_max", ctypes.c_float)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def native_library() -> str:
    """Builds the oracle with -O3 -march=native for this host (oracle/_native/liboracle.so, a few
    seconds) and returns its path: the build of the CPU baseline."""
    subprocess.run(["make", "-s", "-C", HERE, "native"], check=True)
    return os.path.join(HERE, "_native", "liboracle.so")


def load(path=None):
    """The oracle library: oracle/liboracle.so, or the library at `path` from then on (every later
    call uses it)."""
    global _lib, _lib_path
    path = path or _lib_path or LIB_PATH
    if _lib is not None and path == _lib_path:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run make -C oracle")
    lib = ctypes.CDLL(path)
    c = ctypes
    vp, f32, u32, i32 = c.c_void_p, c.c_float, c.c_uint32, c.c_int32
    P = c.POINTER
    sig = {
        "vo_pcg_hash": (u32, [u32]),
        "vo_rand_f": (f32, [P(u32)]),
        "vo_watertight_edge": (f32, [f32, f32, f32, f32]),
        "vo_intersect_triangle": (c.c_int, [P(f32), P(f32), P(f32), P(f32), P(f32), f32, f32, P(f32), P(f32), P(f32)]),
        "vo_mesh_bounds": (None, [vp, vp, u32, vp]),
        "vo_widebvh_build": (i32, [vp, i32, i32, i32, P(vp), vp, P(i32)]),
        "vo_free": (None, [vp]),
        "vo_blas_closest": (Hit, [P(SceneView), P(f32), P(f32), f32, f32, P(c.c_uint64), P(c.c_uint64)]),
        "vo_blas_any": (c.c_int, [P(SceneView), P(f32), P(f32), f32, f32, P(c.c_uint64), P(c.c_uint64)]),
        "vo_brute_closest": (Hit, [P(SceneView), P(f32), P(f32), f32, f32]),
        "vo_brute_any": (c.c_int, [P(SceneView), P(f32), P(f32), f32, f32]),
        "vo_trace_closest_batch": (None, [P(SceneView)] + [vp] * 8 + [c.c_int64, c.c_int, c.c_int] + [vp] * 5),
        "vo_trace_any_batch": (None, [P(SceneView)] + [vp] * 8 + [c.c_int64, c.c_int, c.c_int, vp, vp]),
        "vo_set_fault": (None, [c.c_int]),
        "vo_half_round": (c.c_float, [c.c_float]),
        "vo_render_nrc_f32": (None, [vp, vp, u32, u32, u32, vp, vp, vp, c.c_int, vp]),
        "vo_bsdf_eval_f32": (None, [vp, vp, vp, vp, vp, c.c_int, vp, vp]),
        "vo_bsdf_sample_f32": (c.c_int, [vp, vp, vp, vp, c.c_int, vp, vp, vp, vp]),
        "vo_bsdf_eval_f64": (None, [vp, vp, vp, vp, vp, c.c_int, vp, vp]),
        "vo_bsdf_sample_f64": (c.c_int, [vp, vp, vp, vp, c.c_int, vp, vp, vp, vp]),
        "vo_energy_set_dielectric": (None, [vp]),
        "vo_ltc_set": (None, [vp]),
        "vo_energy_table": (c.POINTER(c.c_uint16), [c.c_int]),
        "vo_metal_energy_table": (c.POINTER(c.c_float), []),
        "vo_energy_ior_node": (c.c_float, [c.c_int]),
        "vo_render_f32": (None, [vp, vp, u32, u32, vp, u32, c.c_int, vp]),
        "vo_render_f64": (None, [vp, vp, u32, u32, vp, u32, c.c_int, vp]),
        "vo_render_restir_f32": (None, [vp, vp, u32, u32, u32, vp, c.c_int, vp]),
        "vo_render_restir_f64": (None, [vp, vp, u32, u32, u32, vp, c.c_int, vp]),
        "vo_max_threads": (c.c_int, []),
        "vo_light_tree_build": (i32, [vp, vp, u32, c.c_int, c.c_double, vp]),
        "vo_light_tree_free": (None, [vp]),
        "vo_light_tree_importance": (f32, [vp, u32, P(f32)]),
        "vo_light_tree_sample": (u32, [vp, P(f32), f32, P(f32)]),
        "vo_light_tree_pmf": (f32, [vp, P(f32), u32]),
        "vo_z_mix": (u32, [u32]),
        "vo_z_shuffle": (u32, [u32, u32]),
        "vo_z_owen": (u32, [u32, u32]),
        "vo_z_owen_seed": (u32, [u32, u32]),
        "vo_z_sobol1": (u32, [u32]),
        "vo_z_o2m3": (u32, [u32, c.c_int]),
        "vo_z_morton": (u32, [u32, u32]),
        "vo_z_pixel_key": (u32, [u32, u32, u32, u32, u32]),
        "vo_z_draws": (None, [u32] * 11 + [vp, vp]),
        "vo_env_tables": (i32, [vp, u32, u32, vp, vp]),
        "vo_env_query_f32": (None, [vp, vp, u32, u32, vp, vp]),
        "vo_env_query_f64": (None, [vp, vp, u32, u32, vp, vp]),
        "vo_light_samples_f32": (None, [vp, vp, u32, u32, u32, vp, vp, vp, vp]),
        "vo_light_samples_f64": (None, [vp, vp, u32, u32, u32, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    tables = reference_energy_tables()
    flat = np.concatenate([tables[k].ravel() for k in DIELECTRIC_KEYS]).astype(np.uint16)
    lib.vo_energy_set_dielectric(flat.ctypes.data_as(ctypes.c_void_p))
    ltc = np.ascontiguousarray(tables["ltc_sheen"], np.float32)
    lib.vo_ltc_set(ltc.ctypes.data_as(ctypes.c_void_p))
    _lib, _lib_path = lib, path
    return lib


# OpenPBR's dielectric tables in the oracle's LUT order (VO_LUT_* in oracle.h).
DIELECTRIC_KEYS = ("ideal_complement", "ideal_avg_complement", "reflection_ratio", "opaque_complement",
                   "opaque_avg_complement")
ENERGY_FIXTURE = os.path.join(os.path.dirname(HERE), "tests", "golden", "openpbr_energy_tables.npz")


def reference_energy_tables():
    """The reference's energy tables as committed data (tests/golden/make_energy_golden.py): OpenPBR's
    vendored unorm16 dielectric arrays and the reference's ideal-metal table (MetalEnergy.metal)."""
    with np.load(ENERGY_FIXTURE) as z:
        return {k: z[k] for k in z.files}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f3(x):
    return (ctypes.c_float * 3)(*[float(v) for v in x])


def pcg_hash(x: int) -> int:
    return load().vo_pcg_hash(x & 0xffffffff)


def rand_f(seed: int):
    s = ctypes.c_uint32(seed)
    v = load().vo_rand_f(ctypes.byref(s))
    return v, s.value


def watertight_edge(p, q) -> float:
    return load().vo_watertight_edge(p[0], p[1], q[0], q[1])


def intersect_triangle(o, d, a, b, c, tmin=0.0, tmax=1e20):
    t, b1, b2 = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    hit = load().vo_intersect_triangle(_f3(o), _f3(d), _f3(a), _f3(b), _f3(c), tmin, tmax, ctypes.byref(t),
                                       ctypes.byref(b1), ctypes.byref(b2))
    return (t.value, b1.value, b2.value) if hit else None


def set_fault(f: int):
    load().vo_set_fault(f)


def max_threads() -> int:
    return load().vo_max_threads()


def mesh_bounds(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(faces, np.uint32)
    b = np.empty((len(f), 6), np.float32)
    load().vo_mesh_bounds(_p(v), _p(f), len(f), _p(b))
    return b


def widebvh_build(bounds: np.ndarray, leaf_size=2, max_depth=20):
    """Sources/Acceleration.swift:296-351: returns (nodes structured array, order, depth)."""
    lib = load()
    b = np.ascontiguousarray(bounds, np.float32)
    n = len(b)
    order = np.empty(n, np.int32)
    nodes = ctypes.c_void_p()
    depth = ctypes.c_int32()
    count = lib.vo_widebvh_build(_p(b), n, leaf_size, max_depth, ctypes.byref(nodes), _p(order), ctypes.byref(depth))
    try:
        arr = np.frombuffer(ctypes.string_at(nodes, 128 * count), WIDE_NODE_DTYPE).copy() if count else \
            np.zeros(0, WIDE_NODE_DTYPE)
    finally:
        lib.vo_free(nodes)
    return arr, order, depth.value


class Scene:
    """Document-order mesh plus the reference 4-wide BVH over it."""

    def __init__(self, vertices: np.ndarray, faces: np.ndarray, with_bvh=True, leaf_size=2, max_depth=20):
        self.vertices = np.ascontiguousarray(vertices, np.float32)
        self.faces = np.ascontiguousarray(faces, np.uint32)
        self.nodes = None
        self.order = None
        self.depth = 0
        self.root_max = 0.0
        if with_bvh and len(self.faces):
            self.nodes, self.order, self.depth = widebvh_build(mesh_bounds(self.vertices, self.faces), leaf_size,
                                                               max_depth)
            root = self.nodes[0]
            used = root["child"] != 0
            lo, hi = root["lo"][:, used], root["hi"][:, used]
            self.root_max = float(max(np.abs(lo).max(), np.abs(hi).max()))
        self._view = SceneView(_p(self.vertices), _p(self.faces), len(self.faces),
                               _p(self.nodes) if self.nodes is not None else None,
                               _p(self.order) if self.order is not None else None, self.root_max)

    def closest(self, rays: np.ndarray, mode=0, threads=0, counters=None):
        """rays (8, n) float32 host SoA.  mode 0: reference BVH order; 1: brute force."""
        if mode == 0 and self.nodes is None:
            raise ValueError("scene built without a BVH")
        r = np.ascontiguousarray(rays, np.float32)
        n = r.shape[1]
        t, b1, b2 = (np.empty(n, np.float32) for _ in range(3))
        face = np.empty(n, np.int32)
        c = (ctypes.c_uint64 * 2)()
        load().vo_trace_closest_batch(ctypes.byref(self._view), *[_p(r[k]) for k in range(8)], n, mode, threads,
                                      _p(t), _p(b1), _p(b2), _p(face), c)
        if counters is not None:
            counters[0] += c[0]
            counters[1] += c[1]
        return {"t": t, "b1": b1, "b2": b2, "face": face}

    def any(self, rays: np.ndarray, mode=0, threads=0, counters=None):
        r = np.ascontiguousarray(rays, np.float32)
        n = r.shape[1]
        occ = np.empty(n, np.uint8)
        c = (ctypes.c_uint64 * 2)()
        load().vo_trace_any_batch(ctypes.byref(self._view), *[_p(r[k]) for k in range(8)], n, mode, threads, _p(occ),
                                  c)
        if counters is not None:
            counters[0] += c[0]
            counters[1] += c[1]
        return occ


# ------------------------------------------------------------------------------------------------
# Integrator (integrator.c): the reference's MIS path loop, with tables from oracle/spectral.py
# ------------------------------------------------------------------------------------------------

class _LightTreeC(ctypes.Structure):
    _fields_ = [("words", ctypes.c_void_p), ("node_count", ctypes.c_uint32), ("order", ctypes.c_void_p),
                ("leaf_node", ctypes.c_void_p), ("depth", ctypes.c_uint32), ("grid_lo", ctypes.c_float * 3),
                ("grid_hi", ctypes.c_float * 3)]


class LightView(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_void_p), ("leaves", ctypes.c_void_p), ("lo", ctypes.c_float * 3),
                ("scale", ctypes.c_float * 3), ("count", ctypes.c_uint32), ("ids", ctypes.c_void_p),
                ("positions", ctypes.c_void_p)]


class LightTree:
    """Sources/LightTree.swift build (oracle/lighttree.c) and the GPU walk's float restatement."""

    def __init__(self, tris: np.ndarray, power: np.ndarray, ids: np.ndarray = None):
        lib = load()
        t = np.ascontiguousarray(tris, np.float32)
        self.power = np.ascontiguousarray(power, np.float64)
        n = len(t)
        c = _LightTreeC()
        lib.vo_light_tree_build(_p(t), _p(self.power), n, 8, 0.75, ctypes.byref(c))
        try:
            self.words = np.frombuffer(ctypes.string_at(c.words, 32 * c.node_count), np.uint32).reshape(-1, 8).copy()
            self.order = np.frombuffer(ctypes.string_at(c.order, 4 * n), np.uint32).copy() if n else np.zeros(0, np.uint32)
            self.leaf_node = np.frombuffer(ctypes.string_at(c.leaf_node, 4 * n), np.uint32).copy() if n else \
                np.zeros(0, np.uint32)
            self.depth = c.depth
            self.grid_lo = np.array(c.grid_lo[:], np.float32)
            self.grid_hi = np.array(c.grid_hi[:], np.float32)
        finally:
            lib.vo_light_tree_free(ctypes.byref(c))
        self.ids = np.arange(n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, np.uint32)
        leaves = np.zeros((n, 3), np.uint32)
        self.positions = np.zeros(n, np.uint32)
        for leaf, inp in enumerate(self.order):
            self.positions[inp] = leaf
            leaves[leaf] = (self.ids[inp], np.float32(self.power[inp]).view(np.uint32), self.leaf_node[leaf])
        self.leaves = leaves
        scale = ((self.grid_hi - self.grid_lo) / np.float32(65535.0)).astype(np.float32)
        self.view = LightView(_p(self.words), _p(self.leaves), (ctypes.c_float * 3)(*self.grid_lo),
                              (ctypes.c_float * 3)(*scale), n, _p(self.ids), _p(self.positions))

    def sample(self, p, u):
        pmf = ctypes.c_float()
        pos = load().vo_light_tree_sample(ctypes.byref(self.view), _f3(p), u, ctypes.byref(pmf))
        return (None if pos == 0xffffffff else int(pos)), pmf.value

    def pmf(self, p, position):
        return load().vo_light_tree_pmf(ctypes.byref(self.view), _f3(p), position)


class RenderScene(ctypes.Structure):
    _fields_ = [("geo", ctypes.c_void_p), ("materials", ctypes.c_void_p), ("material_count", ctypes.c_uint32),
                ("emitter_face", ctypes.c_void_p), ("emitter_cdf", ctypes.c_void_p),
                ("emitter_count", ctypes.c_uint32), ("emitter_total", ctypes.c_float), ("icdf", ctypes.c_void_p),
                ("cmf", ctypes.c_void_p), ("basis", ctypes.c_void_p), ("tree", ctypes.c_void_p),
                ("env", ctypes.c_void_p), ("env_rows", ctypes.c_void_p), ("env_cols", ctypes.c_void_p),
                ("env_w", ctypes.c_uint32), ("env_h", ctypes.c_uint32), ("env_strength", ctypes.c_float),
                ("env_rotation", ctypes.c_float)]


def env_tables(rgba):
    """MaterialLibrary.environmentImportance restated (oracle/environment.c): rgba (H, W, 4) float32
    -> (rows (H, 2), cols (2, H, W)) float32: per table (0 compensated, 1 original) the marginal row
    CDF and the conditional column CDFs."""
    img = np.ascontiguousarray(rgba, np.float32)
    h, w = img.shape[:2]
    rows = np.zeros((h, 2), np.float32)
    cols = np.zeros((2, h, w), np.float32)
    if load().vo_env_tables(_p(img), w, h, _p(rows), _p(cols)) != 0:
        raise ValueError("environment map must be non-empty")
    return rows, cols


class Frame(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("eye", ctypes.c_float * 3),
                ("target", ctypes.c_float * 3), ("fov", ctypes.c_float), ("max_depth", ctypes.c_int32),
                ("sampling_mode", ctypes.c_uint32), ("sampler", ctypes.c_uint32), ("z_temporal", ctypes.c_uint32),
                ("sample_index0", ctypes.c_uint32), ("indirect_reuse", ctypes.c_uint32)]


_SPECTRAL = None


def spectral_data():
    global _SPECTRAL
    if _SPECTRAL is None:
        from oracle import spectral
        d = spectral.Data()
        _SPECTRAL = (d, d.tables())
    return _SPECTRAL


def material_records(materials):
    """materials: list of dicts with type, albedo, emission, and for the layered types roughness, ior,
    metalness, transmission, coat, anisotropy, fuzz -> (records (n, 20) float32, weights)."""
    d, _ = spectral_data()
    rec = np.zeros((len(materials), 20), np.float32)
    weights = np.zeros(len(materials))
    for k, m in enumerate(materials):
        rec[k, 0] = np.frombuffer(np.uint32(m["type"]).tobytes(), np.float32)[0]
        alb = np.clip(np.asarray(m["albedo"], np.float32), 0, 1)
        rec[k, 1:4] = alb
        rec[k, 4:8] = d.reflectance_lagrange(alb)
        e = np.asarray(m.get("emission", (0, 0, 0)), np.float32)
        rec[k, 8:11] = e
        if (e > 0).any():
            weights[k] = max(1e-8, 0.2126 * float(e[0]) + 0.7152 * float(e[1]) + 0.0722 * float(e[2]))
        rec[k, 11] = np.float32(weights[k])
        rec[k, 12:16] = (m.get("roughness", 0.5), m.get("ior", 1.5), m.get("metalness", 0.0), m.get("transmission", 0.0))
        rec[k, 16] = m.get("coat", 0.0)
        rec[k, 17] = m.get("anisotropy", 0.0)
        rec[k, 18] = m.get("fuzz", 0.0)
    return rec, weights


def emitter_list(vertices, faces, weights):
    """main.swift:9312-9346: emitting faces in document order, power = area (double) x weight."""
    slot = faces[:, 3].astype(np.int64)
    ok = (slot < len(weights))
    ok[ok] = weights[slot[ok]] > 0
    ids = np.nonzero(ok)[0].astype(np.uint32)
    a, b, c = (vertices[faces[ids, k], :3].astype(np.float32) for k in range(3))
    e1 = (b - a).astype(np.float64)
    e2 = (c - a).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)
    power = area * weights[slot[ids]].astype(np.float32).astype(np.float64)
    total = power.sum()
    cdf = (np.cumsum(power) / total).astype(np.float32) if total > 0 else np.zeros(0, np.float32)
    if len(cdf):
        cdf[-1] = 1.0
    return ids, cdf, float(total)


def emitter_power(vertices, faces, ids, weights):
    slot = faces[ids, 3].astype(np.int64)
    a, b, c = (vertices[faces[ids, k], :3].astype(np.float32) for k in range(3))
    area = 0.5 * np.linalg.norm(np.cross((b - a).astype(np.float64), (c - a).astype(np.float64)), axis=1)
    return area * weights[slot].astype(np.float32).astype(np.float64)


def build_scene_tree(vertices, faces, ids, weights):
    """MaterialLibrary.lightTreeList (main.swift:9374-9401): a tree when >= 2 emitters and > 1 node."""
    if len(ids) < 2:
        return None
    tris = np.concatenate([vertices[faces[ids, k], :3] for k in range(3)], axis=1).astype(np.float32)
    tree = LightTree(tris, emitter_power(vertices, faces, ids, weights), ids=ids)
    if len(tree.words) <= 1 or tree.depth > 64:
        return None
    return tree


class BSDF:
    """The oracle's layered BSDF subset for one material: type 1 (GLOSSY: metal) or 4 (OPENPBR), with a
    coat of weight `coat` (roughness 0.12, IOR 1.6), a specular lobe of anisotropy `anisotropy` along
    `tangent`, and a fuzz (sheen) layer of weight `fuzz` (roughness 0.5, white)."""

    def __init__(self, type, roughness, ior=1.5, metalness=0.0, transmission=0.0, rho=(0.8, 0.8, 0.8, 0.8),
                 precision="f32", coat=0.0, anisotropy=0.0, tangent=(1.0, 0.0, 0.0), fuzz=0.0):
        self.params = np.array([type, roughness, ior, metalness, transmission, coat, anisotropy, *tangent, fuzz],
                               np.float32)
        self.rho = np.asarray(rho, np.float32)
        self.sfx = precision

    def eval(self, n, wo, wi, inside=False):
        f = np.zeros(4, np.float32)
        pdf = np.zeros(1, np.float32)
        getattr(load(), "vo_bsdf_eval_" + self.sfx)(_p(self.params), _p(np.asarray(n, np.float32)),
                                                     _p(np.asarray(wo, np.float32)), _p(np.asarray(wi, np.float32)),
                                                     _p(self.rho), int(inside), _p(f), _p(pdf))
        return f.astype(np.float64), float(pdf[0])

    def sample(self, n, wo, u3, inside=False):
        d = np.zeros(3, np.float32)
        w = np.zeros(4, np.float32)
        pdf = np.zeros(1, np.float32)
        ok = getattr(load(), "vo_bsdf_sample_" + self.sfx)(_p(self.params), _p(np.asarray(n, np.float32)),
                                                            _p(np.asarray(wo, np.float32)), _p(self.rho), int(inside),
                                                            _p(np.asarray(u3, np.float32)), _p(d), _p(w), _p(pdf))
        return (d.astype(np.float64), w.astype(np.float64), float(pdf[0])) if ok else None


def energy_tables():
    """(metal complements (129, 129), metal averages (129,), dict of the 32-node unorm tables as floats)."""
    lib = load()
    m = np.ctypeslib.as_array(lib.vo_metal_energy_table(), shape=(129 * 129 + 129,)).copy()
    names = ["dielectric", "dielectric_avg", "dielectric_ratio", "opaque", "opaque_avg"]
    shapes = [(32, 32, 32), (32, 32), (32, 32), (32, 32, 32), (32, 32)]
    luts = {}
    for i, (nm, sh) in enumerate(zip(names, shapes)):
        a = np.ctypeslib.as_array(lib.vo_energy_table(i), shape=(int(np.prod(sh)),)).copy()
        luts[nm] = a.reshape(sh).astype(np.float64) / 65535.0
    iors = np.array([lib.vo_energy_ior_node(i) for i in range(32)], np.float64)
    return m[:129 * 129].reshape(129, 129), m[129 * 129:], luts, iors


def _render_inputs(scene, materials, frame, light_tree):
    _, (icdf, cmf, basis) = spectral_data()
    rec, weights = material_records(materials)
    ids, cdf, total = emitter_list(scene.vertices, scene.faces, weights)
    tree = build_scene_tree(scene.vertices, scene.faces, ids, weights) if light_tree else None
    keep = (rec, ids, cdf, icdf, cmf, basis, tree)
    rs = RenderScene(ctypes.cast(ctypes.pointer(scene._view), ctypes.c_void_p), _p(rec), len(rec), _p(ids), _p(cdf),
                     len(ids), total, _p(icdf), _p(cmf), _p(basis),
                     ctypes.cast(ctypes.pointer(tree.view), ctypes.c_void_p) if tree else None)
    fr = Frame(frame["width"], frame["height"], (ctypes.c_float * 3)(*frame["eye"]),
               (ctypes.c_float * 3)(*frame["target"]), frame["fov"], frame["max_depth"], frame.get("sampling_mode", 1))
    fr.sampler = frame.get("sampler", 0)
    fr.z_temporal = frame.get("z_temporal", 1)
    fr.indirect_reuse = frame.get("indirect_reuse", 0)  # 0: ReSTIR GI, as the library's default
    env = frame.get("environment")
    if env is not None and frame.get("env_strength", 1.0) > 0:
        # frame["environment"]: (H, W, 4) rgba; env_strength (default 1) and env_rotation (radians).
        img = np.ascontiguousarray(env, np.float32)
        rows, cols = env_tables(img)
        keep = keep + (img, rows, cols)
        rs.env, rs.env_rows, rs.env_cols = _p(img), _p(rows), _p(cols)
        rs.env_h, rs.env_w = img.shape[0], img.shape[1]
        rs.env_strength = float(frame.get("env_strength", 1.0))
        rs.env_rotation = float(frame.get("env_rotation", 0.0))
    return rs, fr, keep


def render(scene: "Scene", materials, frame: dict, sample_index0: int, samples: int, pixels=None, precision="f32",
           threads=0, light_tree=True):
    """Mean radiance per pixel (pixels: flat indices, default all) -> (len(pixels), 3) float64."""
    lib = load()
    rs, fr, keep = _render_inputs(scene, materials, frame, light_tree)
    w, h = frame["width"], frame["height"]
    pix = np.arange(w * h, dtype=np.uint32) if pixels is None else np.ascontiguousarray(pixels, np.uint32)
    out = np.zeros((len(pix), 3), np.float64)
    fn = lib.vo_render_f64 if precision == "f64" else lib.vo_render_f32
    fn(ctypes.byref(rs), ctypes.byref(fr), sample_index0, samples, _p(pix), len(pix), threads, _p(out))
    del keep
    return out


def env_query(scene: "Scene", materials, frame: dict, dirs, table=0, precision="f32", light_tree=True):
    """The environment of frame["environment"] (with env_strength / env_rotation) seen by the oracle:
    (radiance (n, 3), light-sample density (n,) of `table`, importedProbability applied)."""
    rs, _, keep = _render_inputs(scene, materials, frame, light_tree)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    rgb = np.zeros((len(d), 3), np.float32)
    pdf = np.zeros(len(d), np.float32)
    fn = load().vo_env_query_f64 if precision == "f64" else load().vo_env_query_f32
    fn(ctypes.byref(rs), _p(d), len(d), table, _p(rgb), _p(pdf))
    del keep
    return rgb, pdf


def light_samples(scene: "Scene", materials, frame: dict, p, n, seed=1, table=0, precision="f32", light_tree=True):
    """n light samples at point p (sample_direct_light on the PCG stream `seed`) -> dict wi (n, 3),
    pdf, rgb (n, 3), light (0 none, 1 environment, face + 2)."""
    rs, _, keep = _render_inputs(scene, materials, frame, light_tree)
    out = {"wi": np.zeros((n, 3), np.float32), "pdf": np.zeros(n, np.float32), "rgb": np.zeros((n, 3), np.float32),
           "light": np.zeros(n, np.uint32)}
    fn = load().vo_light_samples_f64 if precision == "f64" else load().vo_light_samples_f32
    fn(ctypes.byref(rs), _f3(p), n, seed, table, _p(out["wi"]), _p(out["pdf"]), _p(out["rgb"]), _p(out["light"]))
    del keep
    return out


def render_restir(scene: "Scene", materials, frame: dict, sample_index0: int, samples: int, candidates=4, eyes=None,
                  precision="f32", threads=0, light_tree=True):
    """ReSTIR frames (sampling mode 0: ReSTIR DI, and ReSTIR GI unless frame["indirect_reuse"] is 1):
    per-pixel mean over `samples` consecutive samples, the first without history -> (height, width, 3)
    float64.  eyes: None or (samples, 3) cameras."""
    lib = load()
    rs, fr, keep = _render_inputs(scene, materials, dict(frame, sampling_mode=0), light_tree)
    w, h = frame["width"], frame["height"]
    ep = None
    if eyes is not None:
        eyes = np.ascontiguousarray(eyes, np.float32).reshape(samples, 3)
        ep = _p(eyes)
    out = np.zeros((h * w, 3), np.float64)
    fn = lib.vo_render_restir_f64 if precision == "f64" else lib.vo_render_restir_f32
    fn(ctypes.byref(rs), ctypes.byref(fr), sample_index0, samples, candidates, ep, threads, _p(out))
    del keep
    return out.reshape(h, w, 3)


def render_nrc(scene: "Scene", materials, frame: dict, sample_index0: int, samples: int, params, period=16,
               threads=0, light_tree=True):
    """The library's cache-terminated estimator with fixed network parameters -> (h, w, 3)."""
    lib = load()
    rs, fr, keep = _render_inputs(scene, materials, frame, light_tree)
    v = np.asarray(scene.vertices, np.float32)[:, :3]
    lo = v.min(0).astype(np.float32)
    ext = (v.max(0) - lo).astype(np.float32)
    scale = np.where(ext > 0, np.float32(1) / np.where(ext > 0, ext, 1), 0).astype(np.float32)
    p = np.ascontiguousarray(params, np.float32)
    w, h = frame["width"], frame["height"]
    out = np.zeros((h * w, 3), np.float64)
    lib.vo_render_nrc_f32(ctypes.byref(rs), ctypes.byref(fr), sample_index0, samples, period, _p(p), _p(lo), _p(scale),
                          threads, _p(out))
    del keep
    return out.reshape(h, w, 3)


def z_draws(width, height, frame_index, sample_index, temporal_model=0, path_vertex=0, stream=0, candidate=0,
            candidate_bits=0, draws=1, kind=1):
    """The Z++ sampler's draws for every pixel (oracle/zsampler.c): (floats, raw 0.32 words), each
    shaped (height, width, draws, kind)."""
    n = width * height * draws * kind
    out = np.empty(n, np.float32)
    raw = np.empty(n, np.uint32)
    load().vo_z_draws(width, height, frame_index, sample_index, temporal_model, path_vertex, stream, candidate,
                      candidate_bits, draws, kind, out.ctypes.data_as(ctypes.c_void_p), raw.ctypes.data_as(ctypes.c_void_p))
    shape = (height, width, draws, kind)
    return out.reshape(shape), raw.reshape(shape)
