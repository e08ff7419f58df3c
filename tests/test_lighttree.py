"""CPU tests of the light BVH: the library's builder against the oracle's restatement of
Sources/LightTree.swift, the reference's structural invariants (tests/Fix_light-tree.swift:6-50),
and the walk's sampling/pmf consistency and unbiasedness (:52-179)."""
import numpy as np
import pytest

import oracle
import vt


def _random_emitters(seed=1):
    rng = np.random.default_rng(seed)
    tris, power = [], []
    for i in range(500):
        o = np.array([rng.uniform(-10, 10), rng.uniform(-2, 2), rng.uniform(-10, 10)], np.float32)
        tris.append(np.concatenate([o, o + rng.uniform(-0.5, 0.5, 3), o + rng.uniform(-0.5, 0.5, 3)]))
        power.append(0.0 if i % 7 == 0 else rng.uniform(0.1, 10))
    tris += [tris[1]] * 40                                  # coincident emitters
    power += [power[1]] * 40
    tris.append(np.zeros(9))                                # degenerate
    power.append(3.0)
    return np.array(tris, np.float32), np.array(power, np.float64)


def _ceiling():
    """A 24 x 24 grid of small triangles facing down at y = 3 and a row of 24 facing up
    (tests/Fix_light-tree.swift:60-66 layout)."""
    tris, power = [], []
    s = 0.3
    for i in range(24):
        for j in range(24):
            x, z = -3.6 + 0.3 * i, -3.6 + 0.3 * j
            a, b, c = (x, 3, z), (x + s, 3, z), (x, 3, z + s)
            tris.append(a + b + c)                          # cross(b - a, c - a) points down
            power.append(0.5 * s * s * (5.0 if i < 12 else 1.0))
    for i in range(24):
        x, z = -3.6 + 0.3 * i, 4.0
        a, b, c = (x, 3, z), (x, 3, z + s), (x + s, 3, z)
        tris.append(a + b + c)                              # facing up
        power.append(0.5 * s * s * 20.0)
    return np.array(tris, np.float32), np.array(power, np.float64)


def test_library_builder_matches_oracle():
    for tris, power in (_random_emitters(1), _random_emitters(2), _ceiling()):
        lib = vt.light_tree_build_host(tris, power)
        ref = oracle.LightTree(tris, power)
        assert np.array_equal(lib["words"], ref.words)
        assert np.array_equal(lib["order"], ref.order)
        assert np.array_equal(lib["leaf_node"], ref.leaf_node)
        assert lib["depth"] == ref.depth
        assert np.array_equal(lib["grid_lo"], ref.grid_lo) and np.array_equal(lib["grid_hi"], ref.grid_hi)


def test_tree_is_well_formed():
    tris, power = _random_emitters(3)
    t = oracle.LightTree(tris, power)
    w = t.words
    n = len(w)
    phi = w[:, 1].view(np.float32).astype(np.float64)
    covered = np.zeros(len(tris), np.int64)
    leaves = 0
    for i in range(n):
        count = int(w[i, 7] >> 27)
        if count:
            leaves += 1
            assert count <= 8
            for pos in range(int(w[i, 6]), int(w[i, 6]) + count):
                covered[t.order[pos]] += 1
                assert t.leaf_node[pos] == i
            continue
        kids = [i + 1, int(w[i, 6])]
        for c in kids:
            assert c < n and int(w[c, 7] & 0x7ffffff) == i
        assert abs(phi[i] - phi[kids[0]] - phi[kids[1]]) <= 1e-5 * max(1, phi[i])
        outer = [w[i, 3] & 0xffff, w[i, 3] >> 16, w[i, 4] & 0xffff, w[i, 4] >> 16, w[i, 5] & 0xffff, w[i, 5] >> 16]
        for c in kids:
            if phi[c] > 0:
                inner = [w[c, 3] & 0xffff, w[c, 3] >> 16, w[c, 4] & 0xffff, w[c, 4] >> 16, w[c, 5] & 0xffff, w[c, 5] >> 16]
                assert all(inner[k] >= outer[k] for k in range(3)) and all(inner[k] <= outer[k] for k in range(3, 6))
    assert np.all(covered == 1)
    assert int(w[0, 7] & 0x7ffffff) == 0x7ffffff
    valid = np.linalg.norm(np.cross(tris[:, 3:6] - tris[:, :3], tris[:, 6:] - tris[:, :3]), axis=1) > 0
    root_power = power[(power > 0) & valid].sum()
    assert abs(phi[0] - root_power) <= 1e-5 * root_power
    assert t.depth <= 64 and n == 2 * leaves - 1


def test_one_lamp_gets_no_tree():
    # A lamp's triangles share place and facing: they stay one leaf, so no tree is built
    # (tests/Fix_light-tree.swift:180-187).
    quad = np.array([[-0.1, 1, -0.1, 0.1, 1, 0.1, -0.1, 1, 0.1], [-0.1, 1, -0.1, 0.1, 1, -0.1, 0.1, 1, 0.1]], np.float32)
    t = oracle.LightTree(quad, np.array([0.01, 0.01]))
    assert len(t.words) == 1


RECEIVERS = [(0.0, 0.0, 0.0), (-3.3, 2.5, -3.3), (3.0, 1.0, 0.0), (0.0, 2.7, 4.2), (0.0, 3.5, 4.2)]


def test_sampled_pmf_equals_hit_pmf_and_normalises():
    # Every sample's pmf equals the pmf the BSDF-hit walk recomputes from the leaf (within 1e-3,
    # tests/Fix_light-tree.swift:171-172); the pmfs plus the probability that the walk finds no light
    # reaching the receiver (light_tree_sample's break) sum to one; empirical choice frequencies
    # follow the pmf.
    tris, power = _ceiling()
    t = oracle.LightTree(tris, power)
    us = (np.arange(20000) + 0.5) / 20000
    for p in RECEIVERS:
        pm = np.array([t.pmf(p, int(t.positions[e])) for e in range(len(tris))])
        counts = np.zeros(len(tris))
        none = 0
        for k, u in enumerate(us):
            pos, pmf = t.sample(p, float(u))
            if pos is None:
                none += 1
                continue
            counts[t.order[pos]] += 1
            if k % 50 == 0:
                assert abs(pmf / t.pmf(p, pos) - 1.0) < 1e-3
        assert abs(pm.sum() + none / len(us) - 1.0) < 2e-3
        assert np.abs(counts / len(us) - pm).max() < 5e-3


def _contributions(tris, power, p):
    """Each emitter's light at p, L cos_l / d^2 dA over its area, by a 6 x 6 sub-triangle quadrature."""
    a, b, c = tris[:, :3].astype(np.float64), tris[:, 3:6].astype(np.float64), tris[:, 6:].astype(np.float64)
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    n = cr / np.linalg.norm(cr, axis=1, keepdims=True)
    radiance = power / area
    k = 6
    total = np.zeros(len(tris))
    for i in range(k):
        for j in range(k - i):
            for du, dv in ((1 / 3, 1 / 3), (2 / 3, 2 / 3)):
                u, v = i + du, j + dv
                if u + v > k:
                    continue
                q = a + (b - a) * (u / k) + (c - a) * (v / k)
                d = np.asarray(p) - q
                d2 = (d * d).sum(1)
                total += radiance * np.maximum(0, (n * d).sum(1) / np.sqrt(d2)) / d2 * area / (k * k)
    return total


def test_tree_selection_lowers_variance_near_lights():
    # The emitter-choice variance sum_e f_e^2 / p_e - (sum_e f_e)^2 with p = tree pmf is below half
    # the power-proportional choice's where nearby lights dominate (tests/Fix_light-tree.swift:173-176).
    tris, power = _ceiling()
    t = oracle.LightTree(tris, power)
    pw = power / power.sum()
    for p in RECEIVERS[1:4]:
        f = _contributions(tris, power, p)
        tree_p = np.array([t.pmf(p, int(t.positions[e])) for e in range(len(tris))])
        live = f > 0
        assert np.all(tree_p[live] > 0)                     # unbiased: every contributor can be chosen
        var_tree = (f[live] ** 2 / tree_p[live]).sum() - f.sum() ** 2
        var_power = (f[live] ** 2 / pw[live]).sum() - f.sum() ** 2
        assert var_tree < 0.5 * var_power, (p, var_tree, var_power)
