"""CPU tests that pin the oracle (oracle/) against the reference's golden vectors, closed forms,
brute force and invariants.  No GPU."""
import numpy as np
import pytest

import oracle
import scenes
import vt


# ---------------------------------------------------------------- RNG (main.swift:91-101)

def test_pcg_seed_golden():
    # tests/GPUChecks.swift:25-30: seed 890625759 makes the second rand_f draw exactly 0.
    v1, s = oracle.rand_f(890625759)
    v2, s = oracle.rand_f(s)
    assert v2 == 0.0 and 0.0 <= v1 < 1.0


def test_rand_f_range_and_uniformity():
    s = 12345
    vals = []
    for _ in range(20000):
        v, s = oracle.rand_f(s)
        vals.append(v)
    vals = np.array(vals)
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert abs(vals.mean() - 0.5) < 0.01 and abs(vals.var() - 1 / 12) < 0.005
    # 24-bit output grid (main.swift:99): every draw is k / 2^24.
    assert np.all(vals * 16777216.0 == np.floor(vals * 16777216.0))


def test_pcg_hash_is_a_bijection_on_a_sample():
    xs = np.arange(0, 1 << 16, dtype=np.uint64)
    hs = {oracle.pcg_hash(int(x)) for x in xs}
    assert len(hs) == len(xs)


# ---------------------------------------------------------------- watertight edge (main.swift:1557-1562)

def test_edge_golden_tie_vector():
    # tests/Fix_renderer-followups.swift:167-168, 234: rounded products tie; exact sign wins.
    p, q = (1.000244140625, 1.00048828125), (1.0, 1.000244140625)
    assert oracle.watertight_edge(p, q) == 2.0 ** -24
    assert oracle.watertight_edge(q, p) == -(2.0 ** -24)


def test_edge_antisymmetry_reference_generator():
    # tests/Fix_renderer-followups.swift:170-176, with the reference's own draw sequence.
    failures = 0
    for gid in range(192):
        seed = (gid * 7919 + 3) & 0xffffffff
        for k in range(64):
            def r():
                nonlocal seed
                v, seed = oracle.rand_f(seed)
                return np.float32(v)
            p = np.array([r(), r()], np.float32) * np.float32(4) - np.float32(2)
            s = r() * np.float32(3) - np.float32(1.5)
            if k < 32:
                jx = np.float32(1) + np.float32(1e-7) * (r() - np.float32(0.5))
                jy = np.float32(1) + np.float32(1e-7) * (r() - np.float32(0.5))
                q = p * s * np.array([jx, jy], np.float32)
            else:
                q = np.array([r(), r()], np.float32) * np.float32(4) - np.float32(2)
            if oracle.watertight_edge(p, q) != -oracle.watertight_edge(q, p):
                failures += 1
    assert failures == 0


def test_planted_edge_fault_is_caught():
    oracle.set_fault(1)
    try:
        p, q = (1.000244140625, 1.00048828125), (1.0, 1.000244140625)
        assert oracle.watertight_edge(p, q) != 2.0 ** -24
    finally:
        oracle.set_fault(0)


# ---------------------------------------------------------------- triangle test (main.swift:1537-1582)

TRI = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))


@pytest.mark.parametrize("d", [(0, 0, -1), (0, 0, 1)])
def test_triangle_closed_form_two_sided(d):
    o = (0.25, 0.5, 1.0 if d[2] < 0 else -1.0)
    hit = oracle.intersect_triangle(o, d, *TRI)
    assert hit is not None
    t, b1, b2 = hit
    # Hit point (0.25, 0.5, 0) = b0 a + b1 b + b2 c with b = (1,0,0), c = (0,1,0).
    assert (t, b1, b2) == (1.0, 0.25, 0.5)


def test_triangle_interval_is_half_open():
    o, d = (0.25, 0.25, 1.0), (0.0, 0.0, -1.0)
    assert oracle.intersect_triangle(o, d, *TRI, tmin=1.0, tmax=2.0) is not None
    assert oracle.intersect_triangle(o, d, *TRI, tmin=0.0, tmax=1.0) is None


def test_triangle_misses_and_degenerates():
    assert oracle.intersect_triangle((1.0, 1.0, 1.0), (0, 0, -1), *TRI) is None          # outside
    assert oracle.intersect_triangle((0.2, 0.2, 0.0), (1, 0, 0), *TRI) is None           # in plane: det = 0
    assert oracle.intersect_triangle((0.2, 0.2, 1.0), (0, 0, -1), (0, 0, 0), (1, 1, 0), (2, 2, 0)) is None


def test_triangle_shared_edge_hits_both_sides():
    # A quad split along its diagonal; a ray through the diagonal hits both triangles (edge value
    # 0 counts as inside for both, WOOP2013).
    a, b, c, d = (0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)
    o, dr = (0.5, 0.5, 1.0), (0, 0, -1)
    assert oracle.intersect_triangle(o, dr, a, b, c) is not None
    assert oracle.intersect_triangle(o, dr, a, c, d) is not None


def test_planted_barycentric_fault_is_caught():
    oracle.set_fault(2)
    try:
        t, b1, b2 = oracle.intersect_triangle((0.25, 0.5, 1.0), (0, 0, -1), *TRI)
        assert (b1, b2) != (0.25, 0.5)
    finally:
        oracle.set_fault(0)


def test_triangle_barycentrics_reconstruct_hit_point():
    rng = np.random.default_rng(1)
    for _ in range(200):
        tri = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        w = rng.dirichlet([1, 1, 1])
        target = (w @ tri.astype(np.float64))
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        d = (target - o)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        hit = oracle.intersect_triangle(o, d, *tri)
        if hit is None:  # rounding at an edge of the triangle
            continue
        t, b1, b2 = hit
        p_ray = o.astype(np.float64) + t * d.astype(np.float64)
        p_bar = (1 - b1 - b2) * tri[0] + b1 * tri[1] + b2 * tri[2]
        assert np.allclose(p_ray, p_bar, atol=1e-4)


# ---------------------------------------------------------------- WideBVH (Sources/Acceleration.swift:106-351)

def _leaf_items(nodes, order):
    items = []
    for n in nodes:
        for k in range(4):
            c = n["child"][k]
            if c < 0:
                first = -c - 1
                items.extend(order[first:first + n["count"][k]])
    return items


def _check_bvh(nodes, order, bounds, leaf_size, max_depth, depth):
    n = len(bounds)
    items = _leaf_items(nodes, order)
    assert sorted(items) == list(range(n))   # every item in exactly one leaf
    assert depth <= max_depth
    # Children enclose their items; inner links point forward.
    def subtree_items(i):
        out = []
        nd = nodes[i]
        for k in range(4):
            c = nd["child"][k]
            if c > 0:
                assert c > i
                out.extend(subtree_items(c))
            elif c < 0:
                out.extend(order[-c - 1:-c - 1 + nd["count"][k]])
        return out
    for i, nd in enumerate(nodes[:200]):
        for k in range(4):
            c = nd["child"][k]
            if c == 0:
                continue
            its = subtree_items(c) if c > 0 else order[-c - 1:-c - 1 + nd["count"][k]]
            if c < 0:
                assert 1 <= nd["count"][k] <= leaf_size
            lo, hi = nd["lo"][:, k], nd["hi"][:, k]
            b = bounds[np.array(its)]
            assert np.all(b[:, :3] >= lo) and np.all(b[:, 3:] <= hi)


@pytest.mark.parametrize("n,leaf", [(1, 2), (2, 2), (3, 2), (37, 2), (500, 2), (500, 1)])
def test_widebvh_invariants_small(n, leaf):
    rng = np.random.default_rng(n)
    lo = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
    b = np.concatenate([lo, lo + rng.uniform(0, 1, (n, 3)).astype(np.float32)], 1)
    nodes, order, depth = oracle.widebvh_build(b, leaf, 20)
    _check_bvh(nodes, order, b, leaf, 20, depth)


def test_widebvh_invariants_parallel_path():
    # n >= 65536 defers subtrees below n/16 (Acceleration.swift:313-347).
    m = vt.icosphere_field(spheres=64, subdivisions=3, seed=9, half_extent=20)
    b = oracle.mesh_bounds(m.vertices, m.faces)
    assert len(b) >= 65536
    nodes, order, depth = oracle.widebvh_build(b, 2, 20)
    _check_bvh(nodes, order, b, 2, 20, depth)


def test_widebvh_coincident_centroids_median_fallback():
    # All centroids equal: no SAH split exists; the median path splits by position (:232-243).
    b = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (40, 1))
    nodes, order, depth = oracle.widebvh_build(b, 2, 20)
    _check_bvh(nodes, order, b, 2, 20, depth)


def test_widebvh_first_split_is_sah_optimal():
    # With leafSize = n - 1 the root is exactly the first binary split.  It must be the cheapest
    # of all 3 x 15 bin-boundary partitions (16 equal centroid bins per axis), costed in float64.
    for seed in range(20):
        rng = np.random.default_rng(100 + seed)
        n = int(rng.integers(8, 64))
        lo = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
        b = np.concatenate([lo, lo + rng.uniform(0.01, 2, (n, 3)).astype(np.float32)], 1)
        nodes, order, _ = oracle.widebvh_build(b, n - 1, 20)
        root = nodes[0]
        leaves = [(int(-root["child"][k] - 1), int(root["count"][k])) for k in range(4) if root["child"][k] < 0]
        assert len(leaves) == 2
        left = set(order[leaves[0][0]:leaves[0][0] + leaves[0][1]].tolist())
        c = (b[:, :3] + b[:, 3:]) * np.float32(0.5)
        clo, chi = c.min(0), c.max(0)

        def area(bb):
            e = np.maximum(bb[:, 3:].max(0) - bb[:, :3].min(0), 0)
            return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]

        best, best_set, costs = None, None, []
        for a in range(3):
            ext = chi[a] - clo[a]
            if ext <= 0:
                continue
            scale = np.float32(16) / np.float32(ext)
            bins = np.clip(((c[:, a] - clo[a]) * scale).astype(np.int64), 0, 15)
            for k in range(15):
                m = bins <= k
                if m.sum() in (0, n):
                    continue
                cost = area(b[m].astype(np.float64)) * m.sum() + area(b[~m].astype(np.float64)) * (~m).sum()
                costs.append(cost)
                if best is None or cost < best:
                    best, best_set = cost, set(np.nonzero(m)[0].tolist())
        costs = sorted(costs)
        if len(costs) > 1 and costs[1] - costs[0] < 1e-4 * costs[0]:
            continue  # near-tie: float32 costing may legitimately pick either
        assert left == best_set or set(range(n)) - left == best_set


# ---------------------------------------------------------------- traversal vs brute force

@pytest.fixture(scope="module")
def patch_scene():
    v, f, grid = scenes.patch(16)
    return oracle.Scene(v, f), grid


def test_watertight_edge_rays(patch_scene):
    # tests/Fix_renderer-followups.swift:141-238: rays aimed at every shared edge and vertex
    # never pass between triangles, and the BVH returns the brute-force hit.
    sc, grid = patch_scene
    targets = scenes.edge_targets(grid)
    rays = scenes.aimed_rays(scenes.EYES, targets)
    bvh = sc.closest(rays, 0)
    brute = sc.closest(rays, 1)
    assert np.all(bvh["face"] >= 0)                   # 0 closest-hit misses
    assert np.array_equal(bvh["t"], brute["t"])       # BVH == brute force, t bit-exact
    tie = bvh["face"] != brute["face"]
    assert np.all(bvh["t"][tie] == brute["t"][tie])   # differing faces only at exact ties
    short = rays.copy()
    dist = np.linalg.norm(np.tile(targets, (len(scenes.EYES), 1)) - np.repeat(scenes.EYES, len(targets), 0), axis=1)
    short[7] = (dist * 1.001 + 1e-3).astype(np.float32)
    assert np.all(sc.any(short, 0) == 1)              # 0 any-hit misses


def test_scattered_triangles_any_and_closest():
    # tests/Fix_geometry.swift:141-161: any-hit agrees with closest hit and brute force.
    v, f = scenes.scattered_triangles(3000)
    sc = oracle.Scene(v, f)
    rays = scenes.random_rays(4000, 3)
    bvh, brute = sc.closest(rays, 0), sc.closest(rays, 1)
    assert np.array_equal(bvh["t"], brute["t"])
    assert np.array_equal(bvh["face"] >= 0, brute["face"] >= 0)
    rays[7] = np.random.default_rng(4).uniform(0.05, 3.05, rays.shape[1]).astype(np.float32)
    occ = sc.any(rays, 0)
    expect = (brute["face"] >= 0) & (brute["t"] < rays[7])
    assert np.array_equal(occ.astype(bool), expect)
    assert np.array_equal(sc.any(rays, 1), occ)


def test_axis_parallel_rays():
    v, f = scenes.scattered_triangles(1500, seed=8)
    sc = oracle.Scene(v, f)
    rays = scenes.axis_rays(3000, 5)
    bvh, brute = sc.closest(rays, 0), sc.closest(rays, 1)
    assert np.array_equal(bvh["t"], brute["t"])


def test_icosphere_closed_mesh_watertight():
    # Rays from outside aimed at the vertices and edge midpoints of a closed icosphere, on the
    # side facing the eye (no silhouette tangency, as the reference's patch test arranges): the
    # target lies on the mesh, so any miss is a crack.
    m = vt.icosphere_field(spheres=1, subdivisions=3, mu=0.0, sigma=0.0, half_extent=0.0, seed=1)
    sc = oracle.Scene(m.vertices, m.faces)
    f = m.faces[:, :3]
    p = m.vertices[:, :3]
    targets = np.concatenate([p, (p[f[:, 0]] + p[f[:, 1]]) * 0.5, (p[f[:, 1]] + p[f[:, 2]]) * 0.5]).astype(np.float32)
    for eye in np.array([[3.1, 0.7, -2.2], [-0.3, 4.0, 0.2], [0.01, -0.02, 5.0]], np.float32):
        facing = targets[(targets * (eye - targets)).sum(1) > 0.3 * np.linalg.norm(eye - targets, axis=1)]
        rays = scenes.aimed_rays(eye[None], facing)
        hits = sc.closest(rays, 0)
        assert np.all(hits["face"] >= 0)
        assert np.array_equal(hits["t"], sc.closest(rays, 1)["t"])
        assert np.all(sc.any(rays, 0) == 1)
