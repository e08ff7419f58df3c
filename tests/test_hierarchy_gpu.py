"""GPU parity of the flat binary hierarchy (build option collapse = 2, the reference's OBJMesh.hierarchy
walked as its software trace_scene walks it): closest hits bit-exact in t, b1, b2 and face against the
oracle's brute force (the same closest-hit rule as the 8-wide traversal), any-hit equal, exact ties and
negative tmin; a rendered image equal to the 8-wide scene's bit for bit (hits do not depend on the
hierarchy); the entry points that need an 8-wide tree refuse it."""
import numpy as np
import pytest

import oracle
import scenes
import vt

from test_trace_gpu import _any, _assert_exact, _closest, _small_cases

pytestmark = pytest.mark.gpu


def test_closest_and_any_bit_exact_vs_brute_force(gpu_ctx):
    for name, v, f, rays in _small_cases():
        print(f"case {name} rays {rays.shape[1]}", flush=True)
        scene = gpu_ctx.set_mesh(vt.Mesh(v, f), collapse=2)
        info = scene.info
        assert info["width"] == 2 and info["node_bytes"] == 48
        osc = oracle.Scene(v, f, with_bvh=False)
        _assert_exact(_closest(scene, rays), osc.closest(rays, mode=1), name)
        seg = rays.copy()
        seg[7] = np.random.default_rng(5).uniform(0.01, 30.0, seg.shape[1]).astype(np.float32)
        assert np.array_equal(_any(scene, seg), osc.any(seg, mode=1)), name
        scene.close()


def test_exact_ties_and_negative_tmin(gpu_ctx):
    v, f = scenes.scattered_triangles(1500, seed=8)
    perm = np.random.default_rng(3).permutation(2 * len(f))
    f2 = np.concatenate([f, f])[perm]
    osc = oracle.Scene(v, f2, with_bvh=False)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f2), collapse=2)
    for tmin in (0.0, -0.7):
        rays = scenes.random_rays(6000, 11, tmin=tmin)
        _assert_exact(_closest(scene, rays), osc.closest(rays, mode=1), f"ties tmin={tmin}")
    scene.close()


def test_bounce_rays_of_a_larger_field(gpu_ctx):
    m = vt.icosphere_field(spheres=96, subdivisions=3, seed=7, half_extent=20)
    rays = vt.camera_rays((0, 8, -70), (0, 0, 0), 45, 256, 128, 7)
    osc = oracle.Scene(m.vertices, m.faces)
    prim = osc.closest(rays, mode=1)
    bounce = vt.bounce_rays(m, rays, prim["face"], prim["b1"], prim["b2"], 11)
    scene = gpu_ctx.set_mesh(m, collapse=2)
    print(f"field {len(m.faces)} tris, depth {scene.info['depth']}", flush=True)
    _assert_exact(_closest(scene, rays), prim, "field camera")
    _assert_exact(_closest(scene, bounce), osc.closest(bounce, mode=1), "field bounce")
    scene.close()


def test_render_equals_the_wide_tree_image(gpu_ctx):
    mesh = vt.cornell()
    imgs = []
    for collapse in (0, 2):
        scene = gpu_ctx.set_mesh(mesh, collapse=collapse)
        scene.set_materials(vt.cornell_materials(17.0))
        r = vt.Renderer(gpu_ctx)
        r.render(scene, 64, 64, vt.orbit_eye(), (0.0, -0.1, 0.0), 42.0, max_depth=4, samples=4)
        imgs.append(r.readback())
        r.close()
        scene.close()
    assert np.array_equal(imgs[0], imgs[1])


def test_entry_points_that_need_the_wide_tree(gpu_ctx):
    mesh = vt.cornell()
    scene = gpu_ctx.set_mesh(mesh, collapse=2)
    rays = scenes.random_rays(64, 2)
    with pytest.raises(vt.VTError, match="8-wide"):
        scene.trace_stats(_dev(rays))
    with pytest.raises(vt.VTError, match="8-wide"):
        vt.Tlas(gpu_ctx, [(scene, np.eye(3, 4, dtype=np.float32))])
    scene.rebuild(collapse=0)
    assert scene.info["width"] == 8
    scene.rebuild(collapse=2)
    assert scene.info["width"] == 2
    scene.close()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
