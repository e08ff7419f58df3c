"""GPU parity on the large-scene recipes (metal-vibe-tracer_amd/vt/workloads.py) at a reduced grid:
an fBm terrain with 64 surface slots (Lambert and GGX dielectric, or the OpenPBR glossy / metal /
transmissive mix) lit by 4096 small emitters through the light tree.  The GPU and the oracle draw
the same sample streams, so the images agree pixel by pixel up to float-order differences; bars:
mean radiance within 0.5% (MIS) and most pixels within 0.1%.  ReSTIR DI (sampling mode 0, 32
candidates) is checked against the oracle's restatement of both reuse passes over 4 frames."""
import numpy as np
import pytest

import oracle
import vt
from vt import workloads

pytestmark = pytest.mark.gpu
W, H = 160, 90


def _frame(depth):
    return dict(width=W, height=H, eye=workloads.TERRAIN_EYE, target=workloads.TERRAIN_TARGET,
                fov=workloads.TERRAIN_FOV, max_depth=depth)


@pytest.mark.parametrize("kind", ["restir", "openpbr"])
def test_terrain_matches_oracle(gpu_ctx, kind):
    mesh, mats = workloads.emitter_terrain(kind, grid=129)
    scene = gpu_ctx.set_mesh(mesh)
    scene.set_materials(workloads.as_materials(mats))
    r = vt.Renderer(gpu_ctx)
    spp, depth = 32, 6
    r.render(scene, W, H, workloads.TERRAIN_EYE, workloads.TERRAIN_TARGET, workloads.TERRAIN_FOV, max_depth=depth,
             samples=spp, sample_index=1)
    got = r.readback()[..., :3].astype(np.float64)
    ref = oracle.render(oracle.Scene(mesh.vertices, mesh.faces), mats, _frame(depth), 1, spp).reshape(H, W, 3)
    rel = abs(got.mean() / ref.mean() - 1)
    close = np.isclose(got, ref, rtol=1e-3, atol=1e-4).all(axis=2).mean()
    print(f"terrain {kind}: gpu {got.mean():.6f} oracle {ref.mean():.6f} rel {rel:.2e} close {close:.4f}", flush=True)
    assert rel < 5e-3 and close > 0.9
    r.close()
    scene.close()


def test_terrain_restir_matches_oracle(gpu_ctx):
    mesh, mats = workloads.emitter_terrain("restir", grid=129)
    scene = gpu_ctx.set_mesh(mesh)
    scene.set_materials(workloads.as_materials(mats))
    r = vt.Renderer(gpu_ctx)
    frames, depth = 4, 3
    for k in range(frames):
        r.render(scene, W, H, workloads.TERRAIN_EYE, workloads.TERRAIN_TARGET, workloads.TERRAIN_FOV,
                 max_depth=depth, samples=1, sample_index=1 + k, sampling_mode=0, candidates=32)
    got = r.readback()[..., :3].astype(np.float64)
    ref = oracle.render_restir(oracle.Scene(mesh.vertices, mesh.faces), mats, _frame(depth), 1, frames, candidates=32)
    rel = abs(got.mean() / ref.mean() - 1)
    close = np.isclose(got, ref, rtol=1e-3, atol=1e-4).all(axis=2).mean()
    print(f"terrain restir: gpu {got.mean():.6f} oracle {ref.mean():.6f} rel {rel:.2e} close {close:.4f}", flush=True)
    assert rel < 5e-3 and close > 0.9
    r.close()
    scene.close()
