"""CPU checks of the large-scene recipes (metal-vibe-tracer_amd/vt/workloads.py): deterministic,
emitters that face down with radiance power / (pi area) for powers in [1, 100], material slots in
range; the batched workload's triangle counts and image sizes in their ranges; the committed FP64
reference images match the recipes' pixel grids."""
import os

import numpy as np

from vt import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_emitter_terrain_recipe():
    for kind in ("restir", "openpbr"):
        mesh, mats = workloads.emitter_terrain(kind, grid=65, emitters=256)
        again, mats2 = workloads.emitter_terrain(kind, grid=65, emitters=256)
        assert np.array_equal(mesh.faces, again.faces) and np.array_equal(mesh.vertices, again.vertices)
        assert mats == mats2
        nf = 2 * 64 * 64
        assert len(mesh.faces) == nf + 256 and len(mats) == 64 + 256
        assert mesh.faces[:nf, 3].max() < 64
        assert np.array_equal(mesh.faces[nf:, 3], 64 + np.arange(256))
        p = mesh.vertices[mesh.faces[nf:, :3], :3].astype(np.float64)
        n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        assert np.all(n[:, 1] < 0) and np.allclose(n[:, [0, 2]], 0, atol=1e-5)   # emit downwards
        area = 0.5 * np.linalg.norm(n, axis=1)
        power = np.array([m["emission"][0] for m in mats[64:]]) * np.pi * area
        assert power.min() >= 1.0 - 1e-3 and power.max() <= 100.0 + 1e-3
        # Above the terrain's nearest grid vertex by 2..6.
        grid_y = mesh.vertices[:65 * 65, 1].reshape(65, 65)
        c = p.mean(1)
        step = 200.0 / 64
        i = np.clip(np.rint((c[:, 0] + 100) / step), 0, 64).astype(int)
        j = np.clip(np.rint((c[:, 2] + 100) / step), 0, 64).astype(int)
        lift = c[:, 1] - grid_y[j, i]
        assert lift.min() >= 2.0 - 1e-4 and lift.max() <= 6.0 + 1e-4
        for m in mats[:64]:
            assert m["type"] in (0, 4) and all(0.0 <= a <= 1.0 for a in m["albedo"])
            if m["type"] == 4:
                assert 0.02 <= m["roughness"] <= 1.0 and 1.3 <= m["ior"] <= 1.7


def test_batch_workload_recipe():
    work = workloads.batch_workload(count=6, seed=3, tri_range=(10_000, 60_000))
    for mesh, mats, w, h, eye, target, fov in work:
        assert (w, h) in workloads.BATCH_SIZES
        assert 0.5 * 10_000 <= len(mesh.faces) <= 1.5 * 60_000
        tags = mesh.faces[:, 3]
        assert tags.max() < len(mats) and any(m["type"] == 3 for m in mats)
        # Tags are per sphere (1280 faces each).
        assert np.all(tags.reshape(-1, 1280) == tags.reshape(-1, 1280)[:, :1])


def test_reference_images_match_the_recipes():
    for kind in ("restir", "openpbr"):
        with np.load(os.path.join(ROOT, "tests", "golden", f"terrain_{kind}_ref.npz")) as z:
            pix, rad = z["pixels"], z["radiance"]
            w, h, spp = int(z["width"]), int(z["height"]), int(z["spp"])
        assert spp == 65536 and (w, h) == (1920, 1080)
        assert np.array_equal(pix, workloads.stratified_pixels(w, h, 64, 36))
        assert rad.shape == (len(pix), 3) and np.all(np.isfinite(rad)) and rad.mean() > 0.1
