"""GPU parity: the gfx950 closest-hit / any-hit kernels through the C ABI against the oracle.

Bar: bit-exact t, b1, b2 and face index against the oracle's brute force (canonical rule: smallest
t, then smallest document face index) on small scenes; at the full traversal-stress size, exact t
and face against the oracle's restatement of the reference 4-wide traversal except at exact
t-ties (the reference keeps the first triangle it visits), which are counted and reported."""
import sys

import numpy as np
import pytest

import oracle
import scenes
import vt

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _closest(scene, rays):
    import torch
    h = scene.trace_closest(_dev(rays))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in h.items()}


def _any(scene, rays):
    import torch
    o = scene.trace_any(_dev(rays))
    torch.cuda.synchronize()
    return o.cpu().numpy()


def _assert_exact(got, ref, what):
    for k in ("t", "b1", "b2", "face"):
        bad = np.nonzero(~((got[k] == ref[k]) | (np.isnan(got[k]) & np.isnan(ref[k]))))[0] if k != "face" else \
            np.nonzero(got[k] != ref[k])[0]
        assert len(bad) == 0, f"{what}: {k} differs at {len(bad)} rays, first {bad[:5]}"


def _small_cases():
    out = []
    m = vt.icosphere_field(spheres=48, subdivisions=2, seed=7, half_extent=10)
    out.append(("spheres-camera", m.vertices, m.faces, vt.camera_rays((0, 5, -40), (0, 0, 0), 45, 200, 100, 7)))
    hf = vt.heightfield(grid=97, octaves=6, half_extent=20, height=3, seed=3)
    out.append(("heightfield-camera", hf.vertices, hf.faces, vt.camera_rays((0, 15, -30), (0, 0, 0), 50, 160, 90, 3)))
    v, f, grid = scenes.patch(16)
    out.append(("patch-edges", v, f, scenes.aimed_rays(scenes.EYES, scenes.edge_targets(grid))))
    v, f = scenes.scattered_triangles(3000)
    out.append(("scattered-random", v, f, scenes.random_rays(5000, 3)))
    out.append(("scattered-axis", v, f, scenes.axis_rays(3001, 5)))
    return out


@pytest.mark.parametrize("leaf,collapse", [(1, 0), (2, 0), (3, 0), (3, 1)])
def test_closest_bit_exact_vs_brute_force(gpu_ctx, leaf, collapse):
    for name, v, f, rays in _small_cases():
        print(f"case {name} leaf {leaf} collapse {collapse} rays {rays.shape[1]}", flush=True)
        scene = gpu_ctx.set_mesh(vt.Mesh(v, f), leaf_size=leaf, collapse=collapse)
        ref = oracle.Scene(v, f, with_bvh=False).closest(rays, mode=1)
        _assert_exact(_closest(scene, rays), ref, name)
        scene.close()


def test_bounce_rays_and_any_hit(gpu_ctx):
    for name, v, f, rays in _small_cases()[:2]:
        print(f"case {name} bounce", flush=True)
        mesh = vt.Mesh(v, f)
        scene = gpu_ctx.set_mesh(mesh)
        osc = oracle.Scene(v, f, with_bvh=False)
        prim = osc.closest(rays, mode=1)
        bounce = vt.bounce_rays(mesh, rays, prim["face"], prim["b1"], prim["b2"], 11)
        ref = osc.closest(bounce, mode=1)
        _assert_exact(_closest(scene, bounce), ref, name + "-bounce")
        # Shadow-style segments: tmax cuts some hits off.
        seg = bounce.copy()
        seg[7] = np.random.default_rng(2).uniform(0.01, 20.0, seg.shape[1]).astype(np.float32)
        assert np.array_equal(_any(scene, seg), osc.any(seg, mode=1)), name
        scene.close()


def test_watertight_edge_rays_any(gpu_ctx):
    v, f, grid = scenes.patch(16)
    targets = scenes.edge_targets(grid)
    rays = scenes.aimed_rays(scenes.EYES, targets)
    dist = np.linalg.norm(np.tile(targets, (len(scenes.EYES), 1)) - np.repeat(scenes.EYES, len(targets), 0), axis=1)
    rays[7] = (dist * 1.001 + 1e-3).astype(np.float32)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f))
    assert np.all(_any(scene, rays) == 1)          # 0 any-hit misses at shared edges/vertices
    assert np.all(_closest(scene, rays)["face"] >= 0)


@pytest.mark.parametrize("leaf", [1, 3])
def test_exact_ties_and_negative_tmin(gpu_ctx, leaf):
    """Every triangle twice (same vertices, two face indices, the copies interleaved at random): each
    hit is an exact t-tie that the smaller face index must win wherever the copies sit in the tree.
    The lanes of a wave test a step's triangles together and fold hits by a minimum over (t, face),
    so this pins that rule. Rays with tmin < 0 accept hits behind the origin (negative t)."""
    v, f = scenes.scattered_triangles(1500, seed=8)
    perm = np.random.default_rng(3).permutation(2 * len(f))
    f2 = np.concatenate([f, f])[perm]
    osc = oracle.Scene(v, f2, with_bvh=False)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f2), leaf_size=leaf)
    for tmin in (0.0, -0.7, -3.0):
        rays = scenes.random_rays(6000, 11, tmin=tmin)
        got, ref = _closest(scene, rays), osc.closest(rays, mode=1)
        print(f"leaf {leaf} tmin {tmin}: hits {np.mean(ref['face'] >= 0):.3f}, negative t {np.mean(ref['t'] < 0):.3f}",
              flush=True)
        _assert_exact(got, ref, f"ties tmin={tmin}")
        seg = rays.copy()
        seg[7] = np.random.default_rng(4).uniform(-0.5, 1.0, seg.shape[1]).astype(np.float32)
        assert np.array_equal(_any(scene, seg), osc.any(seg, mode=1)), f"any tmin={tmin}"
    if tmin < 0:
        assert np.mean(ref["t"] < 0) > 0.05
    scene.close()


def test_signed_zero_ties_on_coplanar_duplicates(gpu_ctx):
    """Pairs of coincident triangles with opposite winding, rays starting inside them with tmin 0 and
    -1: the pair's two hits sit at t = +0 and t = -0 (the sign of T / det), an exact tie that the
    smaller face index must win, also when the wave's triangle tests fold them together."""
    rng = np.random.default_rng(5)
    n = 200
    tris, rays = [], np.empty((8, 2 * n), np.float32)
    for k in range(n):
        a = rng.uniform(-1, 1, 3).astype(np.float32)
        b, c = a + np.float32([0.25, 0, 0]), a + np.float32([0, 0, 0.25])
        pair = [(a, b, c), (a, c, b)]
        tris += pair if k % 2 else pair[::-1]
        o = (a + np.float32(0.25) * (b - a) + np.float32(0.25) * (c - a)).astype(np.float32)
        for j in range(2):
            d = rng.normal(size=3)
            rays[:3, 2 * k + j] = o
            rays[3:6, 2 * k + j] = d / np.linalg.norm(d)
            rays[6, 2 * k + j] = -float(j)
            rays[7, 2 * k + j] = 1e20
    v, f = scenes.mesh_from_triangles(np.array(tris, np.float32))
    ref = oracle.Scene(v, f, with_bvh=False).closest(rays, mode=1)
    assert np.mean(np.signbit(ref["t"]) & (ref["t"] == 0)) > 0.2     # the case is exercised
    for leaf in (1, 2, 3):
        scene = gpu_ctx.set_mesh(vt.Mesh(v, f), leaf_size=leaf)
        got = _closest(scene, rays)
        _assert_exact(got, ref, f"signed-zero ties leaf {leaf}")
        assert np.array_equal(np.signbit(got["t"]), np.signbit(ref["t"])), "the winner's own t, sign included"
        scene.close()


def test_stack_overflow_is_reported(gpu_ctx):
    """A traversal whose node-group stack fills ends the ray and raises the context's fault word, which
    vt_context_check reports (and clears).  The stack is sized to the tree, so the test caps it with the
    scene's test hook; every kernel still exits normally.  Without the cap the same rays are exact and
    nothing is reported."""
    v, f = scenes.scattered_triangles(3000, seed=6)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f), leaf_size=1)
    assert scene.info["depth"] >= 4
    rays = scenes.random_rays(4000, 12)
    gpu_ctx.check()
    scene.set_stack_cap(1)
    _closest(scene, rays)
    _any(scene, rays)
    with pytest.raises(vt.VTError, match="stack overflow"):
        gpu_ctx.check()
    gpu_ctx.check()                                   # reported once, then cleared
    scene.set_stack_cap(0)
    _assert_exact(_closest(scene, rays), oracle.Scene(v, f, with_bvh=False).closest(rays, mode=1), "uncapped")
    gpu_ctx.check()
    scene.close()


def test_empty_scene_and_ragged_sizes(gpu_ctx):
    import torch
    empty = gpu_ctx.set_mesh(vt.Mesh(np.zeros((0, 8), np.float32), np.zeros((0, 4), np.uint32)))
    rays = scenes.random_rays(77, 1)
    h = _closest(empty, rays)
    assert np.all(h["face"] == -1) and np.array_equal(h["t"], rays[7])
    assert np.all(_any(empty, rays) == 0)
    v, f = scenes.scattered_triangles(500)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f))
    osc = oracle.Scene(v, f, with_bvh=False)
    for n in (1, 63, 64, 65, 127, 129, 1000):
        r = scenes.random_rays(n, n)
        _assert_exact(_closest(scene, r), osc.closest(r, mode=1), f"n={n}")
    # n = 0 is a no-op.
    z = torch.zeros((8, 0), dtype=torch.float32, device="cuda:0")
    scene.trace_closest(z)


def test_stats_counters(gpu_ctx):
    v, f = scenes.scattered_triangles(2000)
    scene = gpu_ctx.set_mesh(vt.Mesh(v, f))
    rays = scenes.random_rays(4096, 9)
    nodes, tris = scene.trace_stats(_dev(rays))
    assert nodes >= 4096 and tris > 0
    # Same kernel body: the counting variant returns the same hits.
    assert scene.trace_stats(_dev(rays)) == (nodes, tris)


@pytest.fixture(scope="module")
def stress():
    """The traversal-stress workload: 1024 icospheres (1,310,720 triangles), 2M camera rays
    (2048 x 1024) and 2M cosine bounce rays."""
    mesh = vt.icosphere_field(spheres=1024, subdivisions=3, mu=0.0, sigma=0.5, half_extent=50.0, seed=7)
    prim = vt.camera_rays((0.0, 30.0, -170.0), (0.0, 0.0, 0.0), 45.0, 2048, 1024, 7)
    return mesh, prim


@pytest.fixture(scope="module")
def heightfield():
    """The large scene: the 2237^2 fBm heightfield (9,999,392 triangles), 2M camera rays framing the
    surface (2048 x 1024), as the benchmark's headline workload."""
    mesh = vt.heightfield(grid=2237, octaves=8, gain=0.5, half_extent=100.0, height=12.0, seed=3)
    prim = vt.camera_rays((0.0, 70.0, -70.0), (0.0, 0.0, 0.0), 40.0, 2048, 1024, 3)
    return mesh, prim


def _parity_vs_reference(gpu_ctx, mesh, prim, seed, what):
    scene = gpu_ctx.set_mesh(mesh)
    print(f"{what}: scene", scene.info, flush=True)
    osc = oracle.Scene(mesh.vertices, mesh.faces)
    gp = _closest(scene, prim)
    bounce = vt.bounce_rays(mesh, prim, gp["face"], gp["b1"], gp["b2"], seed)
    gb = _closest(scene, bounce)
    for name, rays, got in (("primary", prim, gp), ("secondary", bounce, gb)):
        ref = osc.closest(rays, mode=0)
        assert np.array_equal(got["t"], ref["t"]), name
        diff = got["face"] != ref["face"]
        # Faces may differ only where two triangles are hit at exactly the same t.
        ties = int(diff.sum())
        assert np.all(got["face"][diff] >= 0) and np.all(ref["face"][diff] >= 0)
        print(f"{what} {name}: {rays.shape[1]} rays, {(got['face'] >= 0).sum()} hits, {ties} exact-t ties "
              f"resolved differently from the reference visit order", flush=True)
        assert ties <= rays.shape[1] // 1000
        # Where the faces agree, barycentrics are bit-exact too.
        same = ~diff
        assert np.array_equal(got["b1"][same], ref["b1"][same]) and np.array_equal(got["b2"][same], ref["b2"][same])
    # Property at full size: any-hit over the open segment == closest hit exists.
    occ = _any(scene, bounce)
    assert np.array_equal(occ.astype(bool), gb["face"] >= 0)
    scene.close()
    return gp, gb


def test_stress_parity_vs_reference_traversal(gpu_ctx, stress):
    mesh, prim = stress
    _parity_vs_reference(gpu_ctx, mesh, prim, 7, "stress")


def test_heightfield_10m_parity_vs_reference_traversal(gpu_ctx, heightfield):
    # The 10M-triangle scene at full size: t bit-exact for 2M camera + 2M bounce rays, faces equal
    # except exact-t ties, against the reference's 4-wide traversal restated in the oracle.
    mesh, prim = heightfield
    gp, gb = _parity_vs_reference(gpu_ctx, mesh, prim, 3, "heightfield")
    assert (gp["face"] >= 0).mean() > 0.95      # the camera frames the surface
