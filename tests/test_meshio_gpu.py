"""GPU: scenes fed through the mesh formats.  An OBJ written from a generated scene and its VTMESH01
payload both upload through vt_scene_upload_indexed and trace bit-exact against the oracle's brute
force on the mesh the reader produced."""
import numpy as np
import pytest

import oracle
import vt

pytestmark = pytest.mark.gpu


def _obj_text(mesh):
    lines = [f"v {x!r} {y!r} {z!r}" for x, y, z in mesh.vertices[:, :3].tolist()]
    lines += [f"f {a + 1} {b + 1} {c + 1}" for a, b, c in mesh.faces[:, :3].tolist()]
    return "\n".join(["o field", "usemtl grey"] + lines) + "\n"


def test_obj_and_payload_scenes_trace_like_the_oracle(gpu_ctx):
    import torch
    src = vt.icosphere_field(spheres=40, subdivisions=2, seed=9, half_extent=10)
    parts, skipped, _ = vt.read_obj(_obj_text(src))
    mesh = vt.merge_parts(parts)
    again = vt.read_mesh_payload(vt.mesh_payload(mesh))
    rays = vt.camera_rays((0.0, 5.0, -40.0), (0.0, 0.0, 0.0), 45.0, 160, 96, 5)
    ref = oracle.Scene(mesh.vertices, mesh.faces, with_bvh=False).closest(rays, mode=1)
    print(f"obj: {len(mesh.faces)} faces ({skipped} skipped), {len(mesh.vertices)} vertices, "
          f"hits {np.mean(ref['face'] >= 0):.3f}", flush=True)
    for name, m in (("obj", mesh), ("payload", again)):
        scene = gpu_ctx.set_mesh(m)
        h = scene.trace_closest(torch.from_numpy(rays).to("cuda:0"))
        torch.cuda.synchronize()
        got = {k: v.cpu().numpy() for k, v in h.items()}
        for k in ("t", "b1", "b2", "face"):
            assert np.array_equal(got[k], ref[k]), f"{name}: {k}"
        scene.close()


def test_triangle_upload_traces_like_the_indexed_upload(gpu_ctx):
    # setMesh([MeshTriangle]): the records weld to the same faces, so hits equal the oracle's on the
    # indexed mesh, face IDs being the records' positions.
    import torch
    mesh = vt.icosphere_field(spheres=30, subdivisions=2, seed=3, half_extent=10)
    rays = vt.camera_rays((0.0, 5.0, -40.0), (0.0, 0.0, 0.0), 45.0, 128, 80, 2)
    ref = oracle.Scene(mesh.vertices, mesh.faces, with_bvh=False).closest(rays, mode=1)
    scene = gpu_ctx.set_triangles(vt.mesh_triangles(mesh))
    assert scene.info["face_count"] == len(mesh.faces)
    h = scene.trace_closest(torch.from_numpy(rays).to("cuda:0"))
    torch.cuda.synchronize()
    for k in ("t", "b1", "b2", "face"):
        assert np.array_equal(h[k].cpu().numpy(), ref[k]), k
    scene.close()


def test_surface_settings_render_as_their_materials(gpu_ctx):
    # SurfaceSettings + emissions map onto the materials resolve_material gives them: an enabled slot
    # is OPENPBR with its colour / roughness / metalness / coat / anisotropy / fuzz / IOR / transmission,
    # an emission makes an emitter, a disabled slot is the imported-mesh default (Lambert 0.7).  Same
    # image as the explicit materials; texture maps are refused.
    mesh = vt.cornell()
    settings = [vt.SurfaceSettings(color=(0.73, 0.73, 0.73), roughness=0.4, ior=1.45, coat=0.6, fuzz=0.5),
                vt.SurfaceSettings(color=(0.65, 0.05, 0.05), roughness=0.2, metalness=1.0, anisotropy=0.6),
                vt.SurfaceSettings(enabled=0)]
    explicit = [vt.material(4, (0.73, 0.73, 0.73), roughness=0.4, ior=1.45, coat=0.6, fuzz=0.5),
                vt.material(4, (0.65, 0.05, 0.05), roughness=0.2, ior=1.5, metalness=1.0, anisotropy=0.6),
                vt.material(0, (0.7, 0.7, 0.7)),
                vt.material(3, (0.7, 0.7, 0.7), emission=(17.0, 17.0, 17.0))]
    imgs = []
    for how in ("settings", "materials"):
        scene = gpu_ctx.set_mesh(mesh)
        if how == "settings":
            scene.set_surface_settings(settings, {3: (17.0, 17.0, 17.0)})
        else:
            scene.set_materials(explicit)
        r = vt.Renderer(gpu_ctx)
        r.render(scene, 64, 64, vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0, max_depth=6, samples=4, sample_index=1)
        imgs.append(r.readback())
        r.close()
        scene.close()
    assert imgs[0][..., :3].mean() > 0 and np.array_equal(imgs[0], imgs[1])
    scene = gpu_ctx.set_mesh(mesh)
    for bad in (vt.SurfaceSettings(map_mask=1),):
        with pytest.raises(vt.VTError):
            scene.set_surface_settings([bad])
    scene.close()
