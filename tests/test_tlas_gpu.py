"""GPU tests of two-level scenes (vt_tlas_*, include/vt.h) against the two-level oracle
(oracle/instances.py): closest hits (t, b1, b2, face, instance) and any hits bit for bit over
instances with general transforms (rotations, non-uniform and mirroring scales, translations), the
tie rule for coincident instances, instances of an empty mesh, a mesh rebuilt under a live vt_tlas,
and the reference's transform errors."""
import numpy as np
import pytest

import oracle
from oracle import instances as oinst
import vt

from scenes import random_rays, scattered_triangles

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def _worlds(count, seed, spread=40.0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        s = rng.uniform(0.4, 3.0, 3) if k % 3 else np.full(3, rng.uniform(0.4, 3.0))
        if k % 5 == 4:
            s[rng.integers(0, 3)] *= -1.0                      # mirroring
        w = np.zeros((3, 4), np.float32)
        w[:, :3] = _rotation(rng) * s
        w[:, 3] = rng.uniform(-spread, spread, 3)
        out.append(w)
    return out


def _meshes():
    ico = vt.icosphere_field(spheres=4, subdivisions=1, seed=3, half_extent=2)
    hf = vt.heightfield(grid=17, octaves=4, half_extent=3, height=1, seed=5)
    sv, sf = scattered_triangles(300, seed=8)
    return [(ico.vertices, ico.faces), (hf.vertices, hf.faces), (sv, sf)]


def _closest(t, rays):
    import torch
    h = t.trace_closest(_dev(rays))
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in h.items()}
    out["instance"] = out["instance"].view(np.uint32)
    return out


def _any(t, rays):
    import torch
    o = t.trace_any(_dev(rays))
    torch.cuda.synchronize()
    return o.cpu().numpy()


def _assert_exact(got, ref, what):
    for k in ("t", "b1", "b2", "face", "instance"):
        bad = np.nonzero(got[k] != ref[k])[0]
        assert len(bad) == 0, f"{what}: {k} differs at {len(bad)} rays, first {bad[:5]}"


def _camera_and_random(seed, spread):
    cam = vt.camera_rays((0.0, 0.6 * spread, -2.5 * spread), (0.0, 0.0, 0.0), 50.0, 160, 120, seed)
    rnd = random_rays(12000, seed, box=1.2 * spread)
    return np.concatenate([cam, rnd], axis=1)


def test_tlas_matches_the_oracle(gpu_ctx):
    meshes = _meshes()
    scenes = [gpu_ctx.set_mesh(vt.Mesh(v, f)) for v, f in meshes]
    assets = [oracle.Scene(v, f, with_bvh=False) for v, f in meshes]
    worlds = _worlds(60, 21)
    inst = [(k % 3, w) for k, w in enumerate(worlds)]
    tl = vt.Tlas(gpu_ctx, [(scenes[a], w) for a, w in inst])
    info = tl.info
    assert info["instance_count"] == 60 and info["rendered_triangles"] == sum(len(meshes[a][1]) for a, _ in inst)
    rays = _camera_and_random(4, 40.0)
    got = _closest(tl, rays)
    ref = oinst.closest(assets, inst, rays)
    print(f"tlas: {info}, hits {np.mean(ref['face'] >= 0):.3f}", flush=True)
    assert np.mean(ref["face"] >= 0) > 0.1
    _assert_exact(got, ref, "closest")
    # Any hit with finite segments (half the closest distance, and past it).
    seg = rays.copy()
    seg[7] = np.where(ref["face"] >= 0, ref["t"] * np.float32(0.5), np.float32(60.0))
    assert np.array_equal(_any(tl, seg), oinst.occluded(assets, inst, seg))
    seg[7] = np.where(ref["face"] >= 0, ref["t"] * np.float32(1.001), np.float32(60.0))
    occ = _any(tl, seg)
    assert np.array_equal(occ, oinst.occluded(assets, inst, seg)) and occ[ref["face"] >= 0].all()
    gpu_ctx.check()
    tl.close()
    for s in scenes:
        s.close()


def test_single_identity_instance_equals_the_mesh(gpu_ctx):
    import torch
    v, f = scattered_triangles(2000, seed=4)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f))
    tl = vt.Tlas(gpu_ctx, [(scene, np.eye(3, 4, dtype=np.float32))])
    rays = random_rays(5000, 3)
    got = _closest(tl, rays)
    h = scene.trace_closest(_dev(rays))
    torch.cuda.synchronize()
    for k in ("t", "b1", "b2", "face"):
        assert np.array_equal(got[k], h[k].cpu().numpy()), k
    assert np.all(got["instance"][got["face"] >= 0] == 0) and np.all(got["instance"][got["face"] < 0] == 0xFFFFFFFF)
    tl.close()
    scene.close()


def test_ties_empty_meshes_and_rebuilds(gpu_ctx):
    meshes = _meshes()
    scenes = [gpu_ctx.set_mesh(vt.Mesh(v, f)) for v, f in meshes]
    empty = gpu_ctx.set_mesh(vt.Mesh(np.zeros((0, 8), np.float32), np.zeros((0, 4), np.uint32)))
    assets = [oracle.Scene(v, f, with_bvh=False) for v, f in meshes]
    worlds = _worlds(12, 5, spread=10.0)
    # Instances 2 and 9 coincide (same mesh, same transform): every hit of theirs is a tie, which
    # the smaller index wins; instance 4 is of a mesh without faces.
    worlds[9] = worlds[2].copy()
    inst = [(k % 3, w) for k, w in enumerate(worlds)]
    inst[9] = (inst[2][0], worlds[9])
    entries = [(scenes[a], w) for a, w in inst]
    entries[4] = (empty, worlds[4])
    tl = vt.Tlas(gpu_ctx, entries)
    ref_inst = [x for k, x in enumerate(inst) if k != 4]
    ref_index = [k for k in range(len(inst)) if k != 4]
    rays = _camera_and_random(7, 10.0)
    ref = oinst.closest(assets, ref_inst, rays)
    ref["instance"] = np.where(ref["face"] >= 0, np.array(ref_index, np.uint32)[np.minimum(ref["instance"], len(ref_index) - 1)],
                               np.uint32(0xFFFFFFFF)).astype(np.uint32)
    got = _closest(tl, rays)
    _assert_exact(got, ref, "ties")
    assert (got["instance"] == 2).any() and not (got["instance"] == 9).any() and not (got["instance"] == 4).any()
    # A mesh rebuilt with another leaf size under the live vt_tlas: same hits.
    scenes[1].rebuild(leaf_size=1)
    _assert_exact(_closest(tl, rays), ref, "after rebuild")
    tl.close()
    # No instances: every ray misses.
    none = vt.Tlas(gpu_ctx, [])
    h = _closest(none, rays[:, :100])
    assert np.all(h["face"] == -1) and np.all(h["instance"] == 0xFFFFFFFF) and np.array_equal(h["t"], rays[7, :100])
    assert not _any(none, rays[:, :100]).any()
    none.close()
    for s in scenes + [empty]:
        s.close()


def test_transform_errors(gpu_ctx):
    v, f = scattered_triangles(50, seed=2)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f))
    singular = np.zeros((3, 4), np.float32)
    singular[0, 0] = singular[1, 0] = 1.0
    singular[2, 2] = 1.0
    tiny = np.eye(3, 4, dtype=np.float32) * np.float32(1e-7)
    far = np.eye(3, 4, dtype=np.float32)
    far[0, 3] = 2e8
    for w, msg in ((singular, "Singular USD transform"), (tiny, "scale exceeds the supported range"),
                   (far, "exceeds the supported scene extent")):
        with pytest.raises(vt.VTError, match=msg):
            vt.Tlas(gpu_ctx, [(scene, np.eye(3, 4, dtype=np.float32)), (scene, w)])
    scene.close()
