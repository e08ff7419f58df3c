"""Writes the FP64 oracle reference images of the large-scene workloads (terrain_<kind>_ref.npz):
the emitter terrain of metal-vibe-tracer_amd/vt/workloads.py (2237^2 grid, 4096 emitters) seen at
1920 x 1080, max depth 8, MIS (NEE + BSDF), on a 64 x 36 stratified pixel subset at 65536 samples
per pixel.  Run from the repository root: python tests/golden/make_terrain_reference.py [kind ...]
(8 cores: about 12 minutes per kind)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "metal-vibe-tracer_amd")]
import oracle  # noqa: E402
from vt import workloads  # noqa: E402

WIDTH, HEIGHT, DEPTH, SPP = 1920, 1080, 8, 65536
GRID_X, GRID_Y = 64, 36


def main(kinds):
    for kind in kinds:
        t0 = time.time()
        mesh, mats = workloads.emitter_terrain(kind)
        scene = oracle.Scene(mesh.vertices, mesh.faces)
        frame = dict(width=WIDTH, height=HEIGHT, eye=workloads.TERRAIN_EYE, target=workloads.TERRAIN_TARGET,
                     fov=workloads.TERRAIN_FOV, max_depth=DEPTH)
        pix = workloads.stratified_pixels(WIDTH, HEIGHT, GRID_X, GRID_Y)
        print(f"{kind}: {len(mesh.faces)} triangles, scene ready in {time.time() - t0:.0f} s", flush=True)
        ref = oracle.render(scene, mats, frame, 1, SPP, pixels=pix, precision="f64")
        out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"terrain_{kind}_ref.npz")
        np.savez(out, pixels=pix.astype(np.uint32), radiance=ref, spp=SPP, depth=DEPTH, width=WIDTH, height=HEIGHT)
        print(f"{kind}: mean {ref.mean():.6f}, {time.time() - t0:.0f} s -> {out}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["openpbr", "restir"])
