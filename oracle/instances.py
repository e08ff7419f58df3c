"""Oracle (test infrastructure only): two-level scenes restated on the CPU with numpy over the
oracle's one-level traversal (oracle.Scene).

Follows the reference's software two-level path:
  MeshSceneLayout.addInstance   Sources/Acceleration.swift:470-516  transform checks (scale range,
                                singular matrix), world rows and their inverse (the record's local rows)
  instance_ray                  main.swift:1634-1637  object-space ray, direction not renormalised
  mesh_closest / blas_closest   main.swift:1788-1812, 1843-1900  closest hit over every instance's
                                asset with its object-space ray; the hit keeps the asset's face and
                                the instance's index
  mesh_occluded / blas_any      main.swift:1815-1838, 1902-1983  any hit over the instances
The instance hierarchy only culls; this restatement visits every instance, so a culling error of the
product's hierarchy shows as a difference.  Arithmetic this build fixes where the reference's rounding
is the Metal compiler's: the inverse by cofactors in double rounded once to float, and each dot product
of instance_ray as ((r0 x + r1 y) + r2 z) (+ r3) with every float32 operation rounded.  Ties at exactly
equal t go to the smaller instance index, then (within the asset) to the smaller face index.

Pinned by tests/test_instances.py: the inverse against exact rational arithmetic, instance_ray against
exact cases, and the two-level result against the one-level oracle on the flattened world mesh for
transforms whose arithmetic is exact (signed axis permutations, power-of-two scales, translations on a
coarse grid), where the two must agree bit for bit.
"""
from __future__ import annotations

import numpy as np


def affine_inverse(world):
    """world: (3, 4) object-to-world rows -> (local (3, 4) float32 rows of the inverse, det, inverse of
    the linear part (3, 3) float64), by cofactors in double."""
    w = np.asarray(world, np.float32).astype(np.float64).reshape(3, 4)
    a00, a01, a02, a10, a11, a12, a20, a21, a22 = (w[0, 0], w[0, 1], w[0, 2], w[1, 0], w[1, 1], w[1, 2],
                                                   w[2, 0], w[2, 1], w[2, 2])
    c = np.array([[a11 * a22 - a12 * a21, a02 * a21 - a01 * a22, a01 * a12 - a02 * a11],
                  [a12 * a20 - a10 * a22, a00 * a22 - a02 * a20, a02 * a10 - a00 * a12],
                  [a10 * a21 - a11 * a20, a01 * a20 - a00 * a21, a00 * a11 - a01 * a10]])
    det = a00 * c[0, 0] + a01 * c[1, 0] + a02 * c[2, 0]
    inv = c / det
    t = w[:, 3]
    local = np.empty((3, 4), np.float32)
    for r in range(3):
        local[r, :3] = inv[r].astype(np.float32)
        local[r, 3] = np.float32(-((inv[r, 0] * t[0] + inv[r, 1] * t[1]) + inv[r, 2] * t[2]))
    return local, float(det), inv


def check_transform(world):
    """The reference's addInstance checks (Sources/Acceleration.swift:477-486); raises ValueError with
    its message."""
    w = np.asarray(world, np.float32).astype(np.float64).reshape(3, 4)
    if not np.all(np.isfinite(w)):
        raise ValueError("Singular USD transform.")
    scale = np.sqrt(w[0, 0] ** 2 + w[1, 0] ** 2 + w[2, 0] ** 2)
    if not np.isfinite(scale) or not (1e-6 <= scale <= 1e6):
        raise ValueError("Combined hierarchy scale exceeds the supported range.")
    det = np.linalg.det(w[:, :3])
    if not np.isfinite(det) or abs(det) <= 1e-18:
        raise ValueError("Singular USD transform.")


def instance_rays(local, rays):
    """Object-space rays (8, n) float32 of world rays (8, n) under local rows (3, 4)."""
    L = np.asarray(local, np.float32)
    r = np.asarray(rays, np.float32)
    out = np.empty_like(r)
    for k in range(3):
        out[k] = ((L[k, 0] * r[0] + L[k, 1] * r[1]) + L[k, 2] * r[2]) + L[k, 3]
        out[3 + k] = (L[k, 0] * r[3] + L[k, 1] * r[4]) + L[k, 2] * r[5]
    out[6] = r[6]
    out[7] = r[7]
    return out


def closest(assets, instances, rays, mode=1):
    """assets: oracle.Scene list; instances: [(asset index, (3, 4) world)]; rays (8, n) float32 ->
    dict t, b1, b2, face, instance (-1 / tmax on a miss).  mode as Scene.closest (1: brute force)."""
    r = np.asarray(rays, np.float32)
    n = r.shape[1]
    best = {"t": r[7].copy(), "b1": np.zeros(n, np.float32), "b2": np.zeros(n, np.float32),
            "face": np.full(n, -1, np.int32), "instance": np.full(n, -1, np.int64)}
    for k, (a, world) in enumerate(instances):
        scene = assets[a]
        if len(scene.faces) == 0:
            continue
        local, _, _ = affine_inverse(world)
        h = scene.closest(instance_rays(local, r), mode=mode if scene.nodes is not None else 1)
        # Instances in increasing index: an equal t never replaces an earlier instance's hit.
        take = (h["face"] >= 0) & (h["t"] < best["t"])
        for key in ("t", "b1", "b2", "face"):
            best[key][take] = h[key][take]
        best["instance"][take] = k
    best["instance"] = np.where(best["face"] >= 0, best["instance"], 0xFFFFFFFF).astype(np.uint32)
    return best


def occluded(assets, instances, rays, mode=1):
    r = np.asarray(rays, np.float32)
    occ = np.zeros(r.shape[1], np.uint8)
    for a, world in instances:
        scene = assets[a]
        if len(scene.faces) == 0:
            continue
        local, _, _ = affine_inverse(world)
        occ |= scene.any(instance_rays(local, r), mode=mode if scene.nodes is not None else 1)
    return occ


def flatten(meshes, instances):
    """World-space vertices (V, 8) and faces (F, 4) of the instances, instance after instance (face
    index = the instance's first face + the asset's face index), positions mapped in double and
    rounded to float32."""
    verts, faces, base = [], [], 0
    for a, world in instances:
        v, f = meshes[a]
        w = np.asarray(world, np.float32).astype(np.float64).reshape(3, 4)
        p = np.asarray(v, np.float32)[:, :3].astype(np.float64)
        vw = np.zeros((len(v), 8), np.float32)
        vw[:, :3] = (p @ w[:, :3].T + w[:, 3]).astype(np.float32)
        f2 = np.asarray(f, np.uint32).copy()
        f2[:, :3] += base
        verts.append(vw)
        faces.append(f2)
        base += len(v)
    return np.concatenate(verts), np.concatenate(faces)
