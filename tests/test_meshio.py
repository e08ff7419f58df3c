"""CPU tests of the mesh formats (include/vt_mesh.h, metal-vibe-tracer_amd/csrc/meshio.cpp).

VTMESH01 (Sources/MeshGeometry.swift:61-92): the byte layout written out by hand here (magic, u32 LE
counts, f32 / u32 LE records), round trips, and every rejection the reference makes with its messages.
OBJ (Sources/StudioModel.swift:444-684): the reference's own known-answer checks of its reader
(tests/Fix_geometry.swift:5-40 -- 0.1 mm faces kept, a 20 um triangle kept, collinear and coincident
faces skipped and counted, the no-usable-faces message, error line numbers with LF, CR LF and
continuations, continuations at the end of the file, a backslash inside a comment, single-component
vt), plus the welding, part / subset, negative-index and error rules the reader states.  Normals
from vn and face normals are normalised with x / sqrt(x.x); the reference's simd_normalize may round
the last bit differently (parity unpinned at 1 ulp for normals; positions and uvs are exact)."""
import struct

import numpy as np
import pytest

import vt


def _tri_corners(part, f=0):
    m = part["mesh"]
    return m.vertices[m.faces[f, :3]]


# ------------------------------------------------------------------ VTMESH01
def test_payload_layout_and_round_trip():
    mesh = vt.icosphere_field(spheres=6, subdivisions=1, seed=2)
    data = vt.mesh_payload(mesh)
    nv, nf = len(mesh.vertices), len(mesh.faces)
    assert len(data) == 16 + 32 * nv + 16 * nf
    expect = b"VTMESH01" + struct.pack("<II", nv, nf) + mesh.vertices.astype("<f4").tobytes() + \
        mesh.faces.astype("<u4").tobytes()
    assert data == expect
    back = vt.read_mesh_payload(data)
    assert np.array_equal(back.vertices.view(np.uint32), mesh.vertices.view(np.uint32))
    assert np.array_equal(back.faces, mesh.faces)
    # Special float values travel bit for bit.
    odd = vt.Mesh(np.array([[-0.0, np.inf, 1e-45, 3.0, 0, 0, 0, 0]] * 3, np.float32), np.array([[0, 1, 2, 7]], np.uint32))
    again = vt.read_mesh_payload(vt.mesh_payload(odd))
    assert np.array_equal(again.vertices.view(np.uint32), odd.vertices.view(np.uint32))
    empty = vt.read_mesh_payload(b"VTMESH01" + struct.pack("<II", 0, 0))
    assert len(empty.vertices) == 0 and len(empty.faces) == 0


def test_payload_rejections():
    mesh = vt.icosphere_field(spheres=2, subdivisions=1, seed=2)
    data = vt.mesh_payload(mesh)
    nf = len(mesh.faces)

    def message(d, maximum=1 << 30):
        with pytest.raises(vt.VTError) as e:
            vt.read_mesh_payload(d, maximum)
        return e.value.message

    assert message(data[:15]) == "Invalid mesh data: missing header."
    assert message(b"VTMESH02" + data[8:]) == "Invalid mesh data: missing header."
    counts = "Invalid mesh data: counts do not match the payload."
    assert message(data[:-1]) == counts
    assert message(data + b"\0") == counts
    assert message(data, maximum=nf - 1) == counts
    assert vt.read_mesh_payload(data, maximum_faces=nf).faces.shape == (nf, 4)
    # More vertices than three per face.
    three = vt.Mesh(np.zeros((4, 8), np.float32), np.array([[0, 1, 2, 0]], np.uint32))
    assert message(vt.mesh_payload(three)) == counts


# ------------------------------------------------------------------ OBJ: the reference's checks
def test_obj_relative_degeneracy_and_skipped_count():
    cells = 40
    lines = []
    for j in range(cells + 1):
        for i in range(cells + 1):
            lines.append(f"v {np.float32(1 + np.float32(i) * np.float32(0.0001))} 0.25 {np.float32(np.float32(j) * np.float32(0.0001))}")
    for j in range(cells):
        for i in range(cells):
            a = j * (cells + 1) + i + 1
            lines.append(f"f {a} {a + 1} {a + cells + 2} {a + cells + 1}")
    parts, skipped, _ = vt.read_obj("\n".join(lines) + "\n")
    assert sum(len(p["mesh"].faces) for p in parts) == 2 * cells * cells and skipped == 0
    tiny, skipped, _ = vt.read_obj("v 0.25 0.1 0.4\nv 0.25002 0.1 0.4\nv 0.25 0.10002 0.4\nf 1 2 3\n")
    assert len(tiny[0]["mesh"].faces) == 1
    mixed, skipped, _ = vt.read_obj("v 0 0 0\nv 1 0 0\nv 2 0 0\nv 0 1 0\nf 1 2 3\nf 1 2 4\nf 1 1 4\n")
    assert sum(len(p["mesh"].faces) for p in mixed) == 1 and skipped == 2
    with pytest.raises(vt.VTError) as e:
        vt.read_obj("v 0 0 0\nv 1 0 0\nv 2 0 0\nf 1 2 3\n")
    assert e.value.message == "OBJ contains no usable faces (1 degenerate faces skipped)."
    with pytest.raises(vt.VTError) as e:
        vt.read_obj("v 0 0 0\n# nothing else\n")
    assert e.value.message == "OBJ contains no usable faces."


def _failure(text):
    with pytest.raises(vt.VTError) as e:
        vt.read_obj(text)
    return e.value.message


def test_obj_error_line_numbers_and_continuations():
    assert "line 5" in _failure("v 0 0 0\nv 1 0 0\n\n# note\nf 1 2 9\n")
    assert "line 5" in _failure("v 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nf 1 2 3\r\nvn 0 1\r\n")
    assert "line 6" in _failure("v 0 0 0\nv 1 \\\n 0 0\nv 0 1 0\nf 1 2 3\nf 1 2 x\n")
    continued, _, _ = vt.read_obj("v 0 0 0\nv 1 \\\n  0 0\nv 0 1 0\nf 1 \\\n 2 3 \\")
    c = _tri_corners(continued[0])
    assert len(continued[0]["mesh"].faces) == 1 and c[1, 0] == 1.0
    commented, _, _ = vt.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0 # comment \\\nf 1 2 3\n")
    assert len(commented[0]["mesh"].faces) == 1
    # VT and FF also end lines; a lone CR too.
    assert "line 4" in _failure("v 0 0 0\x0bv 1 0 0\x0cv 0 1 0\rf 1 2 7\n")


def test_obj_single_component_uv():
    parts, _, _ = vt.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.25\nvt 0.75\nvt 0.5 0.5\nf 1/1 2/2 3/3\n")
    c = _tri_corners(parts[0])
    # uvab == (0.25, 1, 0.75, 1) and uvc.x == 0.5 in the reference's MeshTriangle.
    assert list(c[0, 6:8]) == [0.25, 1.0] and list(c[1, 6:8]) == [0.75, 1.0] and c[2, 6] == 0.5


# ------------------------------------------------------------------ OBJ: the reader's stated rules
def test_obj_welding_normals_and_fan():
    text = "v 0 0 0\nv 2 0 0\nv 2 2 0\nv 0 2 0\nv 1 1 1\nvn 0 0 3\nvn 0 0 0\nf 1 2 3 4\nf 1//1 2//1 5//2\n"
    parts, skipped, _ = vt.read_obj(text)
    m = parts[0]["mesh"]
    # The quad fans into (1 2 3) and (1 3 4) sharing two welded corners.  In the last face, vn 0 0 3
    # normalises to the quad's face normal (0, 0, 1), so corners 1 and 2 weld to the quad's vertices
    # bit for bit; corner 5 is new, and its zero vn becomes +y.
    assert skipped == 0 and m.faces.tolist() == [[0, 1, 2, 0], [0, 2, 3, 0], [0, 1, 4, 0]]
    assert len(m.vertices) == 5
    assert np.array_equal(m.vertices[:4, 3:6], np.tile([0, 0, 1], (4, 1)))
    assert np.array_equal(m.vertices[4], [1, 1, 1, 0, 1, 0, 0, 0])
    # The same position with another uv is another vertex.
    parts, _, _ = vt.read_obj("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 1\nf 1/1 2/1 3/1\nf 1/2 2/1 3/1\n")
    assert len(parts[0]["mesh"].vertices) == 4
    # A corner without vn takes its triangle's normalised edge cross product: (0, -2, 2) / |.| here.
    parts, _, _ = vt.read_obj("v 0 0 0\nv 2 0 0\nv 1 1 1\nf 1 2 3\n")
    n = parts[0]["mesh"].vertices[:, 3:6]
    assert np.allclose(n, [0, -np.sqrt(0.5), np.sqrt(0.5)], rtol=0, atol=6e-8) and np.all(n == n[0])


def test_obj_parts_subsets_and_indices():
    text = "\n".join([
        "mtllib scene.mtl  extra.mtl",
        "v 0 0 0", "v 1 0 0", "v 0 1 0", "v 0 0 1",
        "f 1 2 3",                      # Object / Mesh, Default
        "usemtl red", "f -4 -3 -1",     # negative indices: 1 2 4
        "o Box", "usemtl blue", "f 1 2 3",
        "g lid", "f 1 3 4",
        "o Object", "usemtl red", "f 2 3 4",  # back to the first part, its red subset
        "usemtl blue", "f 1 2 4",
    ]) + "\n"
    parts, skipped, libs = vt.read_obj(text)
    assert libs == ["scene.mtl extra.mtl"]
    assert [(p["object"], p["group"]) for p in parts] == [("Object", "Mesh"), ("Box", "Mesh"), ("Box", "lid")]
    assert parts[0]["subsets"] == ["Default", "red", "blue"]
    assert parts[0]["mesh"].faces[:, 3].tolist() == [0, 1, 1, 2]
    c = _tri_corners(parts[0], 1)
    assert c[:, :3].tolist() == [[0, 0, 0], [1, 0, 0], [0, 0, 1]]
    assert parts[1]["subsets"] == ["blue"] and parts[2]["subsets"] == ["blue"]
    merged = vt.merge_parts(parts)
    assert len(merged.faces) == 6 and merged.faces[:, :3].max() < len(merged.vertices)


def test_obj_errors_and_limits():
    assert _failure("v 0 0 zero\n") == "OBJ contains an invalid number on line 1."
    assert _failure("v 0 0 1e9\n") == "OBJ contains an invalid number on line 1."
    assert _failure("v 0 0 inf\n") == "OBJ contains an invalid number on line 1."
    assert _failure("v 0 0\n") == "Incomplete OBJ vertex on line 1."
    assert _failure("vt\n") == "Incomplete OBJ UV on line 1."
    assert _failure("v 0 0 0\nv 1 0 0\nf 1 2\n") == "OBJ face must contain 3–4096 vertices on line 3."
    assert _failure("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 0\n") == "Invalid OBJ index on line 4."
    assert _failure("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3/1/1/1\n") == "Invalid OBJ face on line 4."
    assert _failure("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -4\n") == "OBJ index is out of bounds on line 4."
    tri = "v 0 0 0\nv 1 0 0\nv 0 1 0\n"
    with pytest.raises(vt.VTError) as e:
        vt.read_obj(tri + "f 1 2 3\nf 1 2 3\n", triangle_limit=1)
    assert e.value.message == "OBJ exceeds the 1-triangle import limit on line 5."
    with pytest.raises(vt.VTError) as e:
        vt.read_obj(tri + "o A\nf 1 2 3\no B\nf 1 2 3\n", node_budget=3)
    assert e.value.message.startswith("OBJ has too many objects and groups") and "line 7" in e.value.message
    assert len(vt.read_obj(tri + "o A\nf 1 2 3\no B\nf 1 2 3\n", node_budget=4)[0]) == 2


def test_obj_feeds_the_builder():
    """An OBJ written from a generated mesh reads back to the same triangles and builds a valid tree."""
    mesh = vt.icosphere_field(spheres=8, subdivisions=2, seed=5, half_extent=6)
    lines = [f"v {x!r} {y!r} {z!r}" for x, y, z in mesh.vertices[:, :3].astype(np.float32).tolist()]
    lines += [f"vn {x!r} {y!r} {z!r}" for x, y, z in mesh.vertices[:, 3:6].astype(np.float32).tolist()]
    lines += [f"f {a + 1}//{a + 1} {b + 1}//{b + 1} {c + 1}//{c + 1}" for a, b, c in mesh.faces[:, :3].tolist()]
    parts, skipped, _ = vt.read_obj("\n".join(lines) + "\n")
    got = vt.merge_parts(parts)
    assert skipped == 0 and len(got.faces) == len(mesh.faces)
    assert np.array_equal(got.vertices[got.faces[:, :3]][..., :3], mesh.vertices[mesh.faces[:, :3]][..., :3])
    nodes, tris, depth, _ = vt.build_bvh8_host(got)
    assert len(tris) == len(got.faces) and depth >= 1


# ------------------------------------------------------------------ MeshGeometry(triangles:)
def test_triangles_weld_back_to_the_indexed_mesh():
    # MeshTriangle records of an indexed mesh weld back to it: corners with identical position,
    # normal and uv bits share a vertex (Sources/MeshGeometry.swift:35-57), in order of first use.
    mesh = vt.icosphere_field(spheres=5, subdivisions=1, seed=4)
    tris = vt.mesh_triangles(mesh)
    assert tris.shape == (len(mesh.faces), 32)
    back = vt.mesh_from_triangles(tris)
    assert len(back.faces) == len(mesh.faces) and len(back.vertices) <= len(mesh.vertices)
    assert np.array_equal(back.vertices[back.faces[:, :3]].view(np.uint32),
                          mesh.vertices[mesh.faces[:, :3]].view(np.uint32))
    assert np.array_equal(back.faces[:, 3], mesh.faces[:, 3])
    # First use order: the first face's corners are vertices 0, 1, 2.
    assert back.faces[0, :3].tolist() == [0, 1, 2]


def test_triangle_welding_rules():
    t = np.zeros((4, 32), np.float32)
    corners = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    for i in range(4):
        t[i, 0:3], t[i, 4:7], t[i, 8:11] = corners
    t[1, 0] = -0.0                          # 0 and -0 stay distinct
    t[2, 12] = 1.0                          # another normal at corner a: another vertex
    t[:, 30] = [3.0, 7.9, np.nan, -1.0]     # subset: truncated, or 0 when not finite / negative
    m = vt.mesh_from_triangles(t)
    assert m.faces[:, 3].tolist() == [3, 7, 0, 0]
    assert m.faces[0, :3].tolist() == [0, 1, 2] and m.faces[3, :3].tolist() == [0, 1, 2]
    assert m.faces[1, :3].tolist() == [3, 1, 2] and m.faces[2, :3].tolist() == [4, 1, 2]
    assert len(m.vertices) == 5
    empty = vt.mesh_from_triangles(np.zeros((0, 32), np.float32))
    assert len(empty.faces) == 0 and len(empty.vertices) == 0
