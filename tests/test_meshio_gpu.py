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
