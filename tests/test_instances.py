"""CPU tests of the two-level oracle (oracle/instances.py), the checker of the vt_tlas path.

Pins: the instance inverse against exact rational arithmetic; instance_ray on exact cases; and the
two-level closest / any hit against the one-level oracle over the flattened world mesh for transforms
whose arithmetic is exact (signed axis permutations, power-of-two scales, translations on a coarse
grid, geometry and origins on a 1/8 grid): there every float operation of the object-space test is the
world test's up to a power of two and a signed permutation, so t, b1, b2 and the hit agree bit for bit
(the flattened face index orders hits as (instance, face) does).  A planted mistake in the oracle (a
transposed inverse) must break the comparison."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from oracle import instances as oinst

from scenes import mesh_from_triangles


def _exact_inverse(w):
    A = [[Fraction(float(w[r][c])) for c in range(3)] for r in range(3)]
    t = [Fraction(float(w[r][3])) for r in range(3)]
    det = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
           + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]))
    adj = [[A[1][1] * A[2][2] - A[1][2] * A[2][1], A[0][2] * A[2][1] - A[0][1] * A[2][2], A[0][1] * A[1][2] - A[0][2] * A[1][1]],
           [A[1][2] * A[2][0] - A[1][0] * A[2][2], A[0][0] * A[2][2] - A[0][2] * A[2][0], A[0][2] * A[1][0] - A[0][0] * A[1][2]],
           [A[1][0] * A[2][1] - A[1][1] * A[2][0], A[0][1] * A[2][0] - A[0][0] * A[2][1], A[0][0] * A[1][1] - A[0][1] * A[1][0]]]
    inv = [[adj[r][c] / det for c in range(3)] for r in range(3)]
    tr = [-(inv[r][0] * t[0] + inv[r][1] * t[1] + inv[r][2] * t[2]) for r in range(3)]
    return [[inv[r][0], inv[r][1], inv[r][2], tr[r]] for r in range(3)]


def test_inverse_rows_are_the_exact_inverse_rounded():
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(200):
        while True:
            w = np.zeros((3, 4), np.float32)
            w[:, :3] = rng.integers(-6, 7, (3, 3)) / 4.0
            w[:, 3] = rng.integers(-80, 81, 3) / 8.0
            if abs(np.linalg.det(w[:, :3].astype(np.float64))) > 0.05:
                break
        local, det, _ = oinst.affine_inverse(w)
        exact = _exact_inverse(w)
        for r in range(3):
            for c in range(4):
                e = exact[r][c]
                ulp = float(np.spacing(np.float32(abs(float(e)))))
                err = abs(Fraction(float(local[r, c])) - e)
                worst = max(worst, float(err) / ulp)
    # One rounding of the double result: within one float step of the exact value.
    assert worst <= 1.0, worst
    # General rotations and scales: the inverse times the matrix is the identity to float precision.
    for _ in range(50):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, c, d = q
        R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                      [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                      [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
        w = np.zeros((3, 4), np.float32)
        w[:, :3] = R * rng.uniform(0.2, 5.0, 3)
        w[:, 3] = rng.uniform(-100, 100, 3)
        local, _, _ = oinst.affine_inverse(w)
        ref = np.linalg.inv(np.vstack([w.astype(np.float64), [0, 0, 0, 1]]))[:3]
        assert np.allclose(local, ref, rtol=2e-7, atol=2e-7 * np.abs(ref).max())


def test_transform_checks_follow_the_reference():
    ok = np.eye(3, 4, dtype=np.float32)
    oinst.check_transform(ok)
    tiny = np.eye(3, 4, dtype=np.float32)
    tiny[0, 0] = 1e-7
    with pytest.raises(ValueError, match="scale exceeds"):
        oinst.check_transform(tiny)
    with pytest.raises(ValueError, match="Singular"):
        w = np.zeros((3, 4), np.float32)
        w[0, 0] = 1.0
        w[1, 0] = 1.0
        oinst.check_transform(w)


def test_instance_ray_exact_cases():
    rng = np.random.default_rng(1)
    rays = np.empty((8, 500), np.float32)
    rays[0:3] = rng.integers(-100, 100, (3, 500)) / 8.0
    rays[3:6] = rng.normal(size=(3, 500))
    rays[6] = 0.0
    rays[7] = 1e20
    ident, _, _ = oinst.affine_inverse(np.eye(3, 4, dtype=np.float32))
    assert np.array_equal(oinst.instance_rays(ident, rays), rays)
    w = np.eye(3, 4, dtype=np.float32)
    w[:, 3] = [2.5, -7.0, 0.125]
    local, _, _ = oinst.affine_inverse(w)
    got = oinst.instance_rays(local, rays)
    assert np.array_equal(got[0:3], rays[0:3] - w[:, 3:4])
    assert np.array_equal(got[3:], rays[3:])


def _grid_mesh(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.integers(-24, 25, (n, 1, 3)) / 8.0
    e = rng.integers(-4, 5, (n, 3, 3)) / 8.0
    return mesh_from_triangles((c + e).astype(np.float32))


def _exact_transforms(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        P = np.zeros((3, 3))
        P[np.arange(3), rng.permutation(3)] = rng.choice([-1.0, 1.0], 3)
        w = np.zeros((3, 4), np.float32)
        w[:, :3] = P * 2.0 ** rng.integers(-1, 2)
        w[:, 3] = rng.integers(-64, 65, 3) / 4.0
        out.append(w)
    return out


def _grid_rays(n, seed, targets):
    # Origins on a 1/8 grid, aimed at points near the instances' translations.
    rng = np.random.default_rng(seed)
    r = np.empty((8, n), np.float32)
    r[0:3] = rng.integers(-160, 161, (3, n)) / 8.0
    aim = np.asarray(targets)[rng.integers(0, len(targets), n)].T + rng.normal(scale=2.0, size=(3, n))
    d = aim - r[0:3]
    r[3:6] = d / np.linalg.norm(d, axis=0)
    r[6] = 0.0
    r[7] = 1e20
    return r


def _two_level_case(seed):
    meshes = [_grid_mesh(60, seed), _grid_mesh(90, seed + 1)]
    worlds = _exact_transforms(24, seed + 2)
    inst = [(k % 2, w) for k, w in enumerate(worlds)]
    assets = [oracle.Scene(v, f, with_bvh=False) for v, f in meshes]
    return meshes, inst, assets, _grid_rays(6000, seed + 3, [w[:, 3] for w in worlds])


def test_two_level_matches_the_flattened_mesh_bit_for_bit():
    meshes, inst, assets, rays = _two_level_case(11)
    got = oinst.closest(assets, inst, rays)
    vf, ff = oinst.flatten(meshes, inst)
    ref = oracle.Scene(vf, ff, with_bvh=False).closest(rays, mode=1)
    counts = [len(meshes[a][1]) for a, _ in inst]
    base = np.concatenate([[0], np.cumsum(counts)])
    hit = got["face"] >= 0
    assert hit.mean() > 0.2
    assert np.array_equal(got["t"], ref["t"]) and np.array_equal(got["b1"], ref["b1"]) and np.array_equal(got["b2"], ref["b2"])
    flat = np.where(hit, base[np.minimum(got["instance"], len(inst) - 1)] + got["face"], -1)
    assert np.array_equal(flat, ref["face"])
    occ = oinst.occluded(assets, inst, rays)
    assert np.array_equal(occ, oracle.Scene(vf, ff, with_bvh=False).any(rays, mode=1))


def test_planted_transposed_inverse_is_caught():
    meshes, inst, assets, rays = _two_level_case(12)
    vf, ff = oinst.flatten(meshes, inst)
    ref = oracle.Scene(vf, ff, with_bvh=False).closest(rays, mode=1)
    real = oinst.affine_inverse

    def transposed(world):
        local, det, inv = real(world)
        bad = local.copy()
        bad[:, :3] = local[:, :3].T
        return bad, det, inv

    oinst.affine_inverse = transposed
    try:
        got = oinst.closest(assets, inst, rays)
    finally:
        oinst.affine_inverse = real
    assert not np.array_equal(got["t"], ref["t"])
