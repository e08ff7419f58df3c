"""CPU tests of libvt.so: exported symbols, the host BVH builder's layout invariants, error paths
and the workload generators.  No compute on a GPU."""
import ctypes
import glob
import os
import re
import struct

import numpy as np
import pytest

import scenes
import vt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(vt_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = vt.load()
    declared = _declared_functions()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing
    assert declared == set(vt.EXPORTED)


def _dequant(q, e, p):
    # fma(q, 2^(e-127), p): the product is exact, so float64 sum then one rounding to float32.
    s = np.ldexp(1.0, e.astype(np.int64) - 127)
    return (p.astype(np.float64) + q.astype(np.float64) * s).astype(np.float32)


def _check_bvh8(mesh, nodes, tris, depth, leaf_size):
    f = len(mesh.faces)
    assert len(tris) == f
    assert sorted(tris["face"].tolist()) == list(range(f))
    # Triangle records are exact copies of the document's corner positions.
    fc = mesh.faces[tris["face"]]
    for k, key in enumerate(("v0", "v1", "v2")):
        assert np.array_equal(tris[key], mesh.vertices[fc[:, k], :3])
    assert depth <= 24
    seen_tris = np.zeros(f, np.int64)
    stack = [(0, 1)]
    visited = 0
    while stack:
        i, level = stack.pop()
        visited += 1
        assert level <= depth
        nd = nodes[i]
        rank = 0
        for s in range(8):
            meta = int(nd["meta"][s])
            if meta == 0:
                continue
            lo = _dequant(nd["qlo"][:, s], nd["e"], nd["p"])
            hi = _dequant(nd["qhi"][:, s], nd["e"], nd["p"])
            if meta & 0x80:
                assert nd["imask"] >> s & 1
                child = int(nd["child_base"]) + rank
                rank += 1
                stack.append((child, level + 1))
                sub = _subtree_tris(nodes, child)
            else:
                assert not (nd["imask"] >> s & 1)
                off, cnt = meta >> 2, meta & 3
                assert 1 <= cnt <= leaf_size
                sub = np.arange(int(nd["tri_base"]) + off, int(nd["tri_base"]) + off + cnt)
                seen_tris[sub] += 1
            pts = np.concatenate([tris["v0"][sub], tris["v1"][sub], tris["v2"][sub]])
            assert np.all(pts >= lo) and np.all(pts <= hi), "child box must enclose its triangles"
        assert rank == bin(int(nd["imask"])).count("1")
    assert visited == len(nodes)
    assert np.all(seen_tris == 1)


def _subtree_tris(nodes, i):
    out = []
    stack = [i]
    while stack:
        nd = nodes[stack.pop()]
        rank = 0
        for s in range(8):
            meta = int(nd["meta"][s])
            if meta & 0x80:
                stack.append(int(nd["child_base"]) + rank)
                rank += 1
            elif meta:
                out.extend(range(int(nd["tri_base"]) + (meta >> 2), int(nd["tri_base"]) + (meta >> 2) + (meta & 3)))
    return np.array(out, np.int64)


@pytest.mark.parametrize("collapse", [0, 1])
@pytest.mark.parametrize("leaf", [1, 2, 3])
def test_bvh8_layout_invariants(leaf, collapse):
    mesh = vt.icosphere_field(spheres=24, subdivisions=2, seed=3, half_extent=6)
    nodes, tris, depth, root_max = vt.build_bvh8_host(mesh, leaf_size=leaf, collapse=collapse)
    _check_bvh8(mesh, nodes, tris, depth, leaf)
    assert root_max == np.abs(mesh.vertices[:, :3]).max()


def _sah_cost(nodes, c_node=1.0, c_tri=0.1):
    """SAH cost of an 8-wide tree from its (dequantised) child boxes, relative to the root box."""
    def half_area(lo, hi):
        e = np.maximum(hi.astype(np.float64) - lo.astype(np.float64), 0)
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]
    total, stack = 0.0, [0]
    root = nodes[0]
    valid = [s for s in range(8) if root["meta"][s]]
    lo = np.min([_dequant(root["qlo"][:, s], root["e"], root["p"]) for s in valid], 0)
    hi = np.max([_dequant(root["qhi"][:, s], root["e"], root["p"]) for s in valid], 0)
    a_root = half_area(lo, hi)
    total += c_node * a_root
    while stack:
        nd = nodes[stack.pop()]
        rank = 0
        for s in range(8):
            meta = int(nd["meta"][s])
            if not meta:
                continue
            a = half_area(_dequant(nd["qlo"][:, s], nd["e"], nd["p"]), _dequant(nd["qhi"][:, s], nd["e"], nd["p"]))
            if meta & 0x80:
                total += c_node * a
                stack.append(int(nd["child_base"]) + rank)
                rank += 1
            else:
                total += c_tri * (meta & 3) * a
    return total / a_root


def test_optimal_collapse_fills_nodes_and_lowers_sah_cost():
    """Greedy widening splits the same ranges with the same binned-SAH splits, so its tree is one of
    the collapses of the binary tree the dynamic programme chooses among: the programme's tree must
    cost no more (measured on the quantised boxes) and fill its nodes better."""
    for mesh in (vt.icosphere_field(spheres=40, subdivisions=3, seed=5, half_extent=20),
                 vt.heightfield(grid=129, octaves=6, half_extent=30, height=4, seed=2)):
        opt = vt.build_bvh8_host(mesh, collapse=0)
        greedy = vt.build_bvh8_host(mesh, collapse=1)
        _check_bvh8(mesh, opt[0], opt[1], opt[2], 3)
        fill = lambda n: float((n["meta"] != 0).sum(1).mean())
        c_opt, c_greedy = _sah_cost(opt[0]), _sah_cost(greedy[0])
        print(f"{len(mesh.faces)} tris: nodes {len(opt[0])} vs {len(greedy[0])}, children/node {fill(opt[0]):.2f} vs "
              f"{fill(greedy[0]):.2f}, SAH {c_opt:.2f} vs {c_greedy:.2f}")
        assert c_opt < c_greedy and fill(opt[0]) > 6.0 > fill(greedy[0]) and len(opt[0]) < len(greedy[0])


def test_bvh8_degenerate_inputs():
    # Coincident triangles (no SAH split; median fallback), a single triangle, tiny and huge scales.
    tri = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)
    for tris_in in (np.repeat(tri, 50, 0), tri, tri * np.float32(1e-20), tri * np.float32(1e20) + np.float32(3e20)):
        v, f = scenes.mesh_from_triangles(tris_in)
        mesh = vt.Mesh(v, f)
        for collapse in (0, 1):
            nodes, t, depth, _ = vt.build_bvh8_host(mesh, collapse=collapse)
            _check_bvh8(mesh, nodes, t, depth, 3)


@pytest.mark.parametrize("collapse", [0, 1])
def test_bvh8_deterministic(collapse):
    mesh = vt.icosphere_field(spheres=80, subdivisions=3, seed=4, half_extent=30)
    a = vt.build_bvh8_host(mesh, collapse=collapse)
    os.environ["VT_BUILD_THREADS"] = "1"
    try:
        b = vt.build_bvh8_host(mesh, collapse=collapse)
    finally:
        del os.environ["VT_BUILD_THREADS"]
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()


def test_bvh8_depth_limit_falls_back():
    """A tree deeper than max_depth is rebuilt greedily, then with median splits; the reported depth
    bounds the real one (the traversal sizes its stack from it), and too small a limit is an error."""
    mesh = vt.icosphere_field(spheres=24, subdivisions=2, seed=3, half_extent=6)
    free = vt.build_bvh8_host(mesh)
    for limit in (free[2] - 1, free[2] - 2):
        nodes, tris, depth, _ = vt.build_bvh8_host(mesh, max_depth=limit)
        assert depth <= limit
        _check_bvh8(mesh, nodes, tris, depth, 3)
        assert len(nodes) != len(free[0])
    with pytest.raises(vt.VTError):
        vt.build_bvh8_host(mesh, max_depth=3)


def test_bvh8_empty_scene():
    mesh = vt.Mesh(np.zeros((0, 8), np.float32), np.zeros((0, 4), np.uint32))
    nodes, tris, depth, _ = vt.build_bvh8_host(mesh)
    assert len(nodes) == 0 and len(tris) == 0


def test_error_paths():
    mesh = vt.Mesh(np.zeros((3, 8), np.float32), np.array([[0, 1, 5, 0]], np.uint32))
    with pytest.raises(vt.VTError) as e:
        vt.build_bvh8_host(mesh)
    assert e.value.status == 1 and "out of range" in e.value.message
    ok = vt.Mesh(np.zeros((3, 8), np.float32), np.array([[0, 1, 2, 0]], np.uint32))
    with pytest.raises(vt.VTError) as e:
        vt.build_bvh8_host(ok, leaf_size=4)
    assert e.value.status == 1
    bad = ok.vertices.copy()
    bad[0, 0] = np.nan
    with pytest.raises(vt.VTError):
        vt.build_bvh8_host(vt.Mesh(bad, ok.faces))


def test_context_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(vt.VTError) as e:
        vt.Context(0)
    assert e.value.message


def test_workload_counts_and_determinism():
    lib = vt.load()
    nv, nf = ctypes.c_uint32(), ctypes.c_uint32()
    lib.vt_gen_icosphere_field(1024, 3, 0.0, 0.5, 50.0, 7, None, ctypes.byref(nv), None, ctypes.byref(nf))
    assert nf.value == 1_310_720 and nv.value == 1024 * 642
    lib.vt_gen_heightfield(2237, 8, 0.5, 100.0, 12.0, 3, None, ctypes.byref(nv), None, ctypes.byref(nf))
    assert nf.value == 9_999_392
    a = vt.icosphere_field(spheres=16, seed=7)
    b = vt.icosphere_field(spheres=16, seed=7)
    c = vt.icosphere_field(spheres=16, seed=8)
    assert a.vertices.tobytes() == b.vertices.tobytes()
    assert a.vertices.tobytes() != c.vertices.tobytes()
    r1 = vt.camera_rays((0, 1, -5), (0, 0, 0), 40, 64, 32, 1)
    r2 = vt.camera_rays((0, 1, -5), (0, 0, 0), 40, 64, 32, 1)
    assert r1.tobytes() == r2.tobytes()
    assert np.allclose(np.linalg.norm(r1[3:6], axis=0), 1, atol=1e-6)


def test_icosphere_radius_distribution():
    m = vt.icosphere_field(spheres=1024, subdivisions=0, seed=7)
    p = m.vertices[:, :3].reshape(1024, 12, 3).astype(np.float64)
    c = p.mean(1)
    r = np.linalg.norm(p - c[:, None], axis=2).mean(1)
    lr = np.log(r)
    assert abs(lr.mean()) < 0.05 and abs(lr.std() - 0.5) < 0.05
    assert np.all(np.abs(c) <= 50.0 + 1e-3)
