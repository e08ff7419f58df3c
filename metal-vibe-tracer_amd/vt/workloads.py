"""Seeded scene recipes of the large-scene workloads: an fBm heightfield whose faces carry a few
dozen surface materials, lit by thousands of small emitting triangles hovering above it.

  'restir'   the ReSTIR DI workload: a 2237^2 grid (~10M triangles, 8 octaves, gain 0.5, seed 3),
             4096 emitting triangles with powers uniform in [1, 100], half of the 64 surface slots
             Lambert (albedo uniform in [0.2, 0.8]^3) and half GGX dielectric (OpenPBR, metalness 0,
             IOR 1.5, roughness uniform in [0.05, 1])
  'openpbr'  the learned-cache workload: the same terrain and emitters with an OpenPBR mix in the
             surface slots -- glossy dielectric (IOR U[1.3, 1.7]), metal, and transmissive dielectric
             (IOR U[1.3, 1.7]) in turn, roughness uniform in [0.02, 0.5]

Materials are dicts (type, albedo, emission, roughness, ior, metalness, transmission), the form the
oracle takes; `as_materials` turns them into vt_material records.  Each emitter is an equilateral
triangle of circumradius 0.35 at a random yaw, centred over a uniform point of [-90, 90]^2 and 2 to 6
units above the terrain's nearest grid vertex, emitting downwards (its front side) with radiance
power / (pi x area), white under D65.
"""
import numpy as np

TERRAIN_EYE = (0.0, 40.0, -20.0)
TERRAIN_TARGET = (0.0, 0.0, 30.0)
TERRAIN_FOV = 40.0


def _terrain_draws(kind, nf, emitters, slots, seed, power):
    """The recipe's random draws, in their fixed order: face tags, emitter centres, heights above the
    terrain, yaws, powers, then the surface palette (slot materials)."""
    if kind not in ("restir", "openpbr"):
        raise ValueError("kind must be 'restir' or 'openpbr'")
    rng = np.random.default_rng(seed)
    tags = rng.integers(0, slots, nf, dtype=np.uint32)
    cx, cz = rng.uniform(-90.0, 90.0, emitters), rng.uniform(-90.0, 90.0, emitters)
    lift = rng.uniform(2.0, 6.0, emitters)
    yaw = rng.uniform(0.0, 2.0 * np.pi, emitters)
    watts = rng.uniform(power[0], power[1], emitters)
    mats = []
    for k in range(slots):
        r = rng.random(4)
        if kind == "restir":
            if k % 2 == 0:
                mats.append(dict(type=0, albedo=tuple(0.2 + 0.6 * r[:3])))
            else:
                mats.append(dict(type=4, albedo=tuple(0.2 + 0.6 * r[:3]), roughness=0.05 + 0.95 * r[3], ior=1.5))
        else:
            rough = 0.02 + 0.48 * r[3]
            ior = 1.3 + 0.4 * rng.random()
            if k % 3 == 0:
                mats.append(dict(type=4, albedo=tuple(0.2 + 0.6 * r[:3]), roughness=rough, ior=ior))
            elif k % 3 == 1:
                mats.append(dict(type=4, albedo=tuple(0.5 + 0.45 * r[:3]), roughness=rough, ior=1.5, metalness=1.0))
            else:
                mats.append(dict(type=4, albedo=tuple(0.8 + 0.2 * r[:3]), roughness=rough, ior=ior, transmission=1.0))
    return tags, cx, cz, lift, yaw, watts, mats


def surface_palette(kind="restir", grid=2237, emitters=4096, slots=64, seed=3, power=(1.0, 100.0)):
    """The surface slots' materials of emitter_terrain(kind) without building the terrain."""
    nf = 2 * (grid - 1) * (grid - 1)
    return _terrain_draws(kind, nf, emitters, slots, seed, power)[-1]


def emitter_terrain(kind="restir", grid=2237, emitters=4096, slots=64, seed=3, power=(1.0, 100.0)):
    """Returns (mesh, materials): face tags 0..slots-1 pick a surface slot, slots + k emitter k."""
    import vt
    if kind not in ("restir", "openpbr"):
        raise ValueError("kind must be 'restir' or 'openpbr'")
    mesh = vt.heightfield(grid=grid, octaves=8, gain=0.5, half_extent=100.0, height=12.0, seed=seed)
    nf, nv = len(mesh.faces), len(mesh.vertices)
    tags, cx, cz, lift, yaw, watts, mats = _terrain_draws(kind, nf, emitters, slots, seed, power)
    # Emitters: unwelded triangles appended after the terrain.
    step = 200.0 / (grid - 1)
    gi = np.clip(np.rint((cx + 100.0) / step), 0, grid - 1).astype(np.int64)
    gj = np.clip(np.rint((cz + 100.0) / step), 0, grid - 1).astype(np.int64)
    cy = mesh.vertices[gj * grid + gi, 1].astype(np.float64) + lift
    # Corners at yaw, yaw + 120, yaw + 240 degrees: (b - a) x (c - a) points down (-y).
    ang = yaw[:, None] + np.array([0.0, 2.0, 4.0]) * np.pi / 3.0
    corners = np.stack([cx[:, None] + 0.35 * np.cos(ang), np.repeat(cy[:, None], 3, 1), cz[:, None] + 0.35 * np.sin(ang)], -1)
    ev = np.zeros((3 * emitters, 8), np.float32)
    ev[:, :3] = corners.reshape(-1, 3)
    ev[:, 4] = -1.0
    ef = np.zeros((emitters, 4), np.uint32)
    ef[:, :3] = (nv + np.arange(3 * emitters)).reshape(emitters, 3)
    ef[:, 3] = slots + np.arange(emitters, dtype=np.uint32)
    mesh.faces[:, 3] = tags
    mesh.vertices = np.concatenate([mesh.vertices, ev])
    mesh.faces = np.concatenate([mesh.faces, ef])
    area = 0.75 * np.sqrt(3.0) * 0.35 ** 2
    radiance = watts / (np.pi * area)
    for L in radiance:
        mats.append(dict(type=3, albedo=(0.0, 0.0, 0.0), emission=(float(L),) * 3))
    return mesh, mats


def as_materials(mats):
    """Material dicts -> vt_material records."""
    import vt
    return [vt.material(m["type"], m["albedo"], m.get("emission", (0.0, 0.0, 0.0)), m.get("roughness", 0.5),
                        m.get("ior", 1.5), m.get("metalness", 0.0), 1.0, m.get("transmission", 0.0)) for m in mats]


def stratified_pixels(width, height, nx, ny):
    """Flat indices of an nx x ny grid of pixel centres (the oracle's pixel subset)."""
    xs = (np.arange(nx) * width) // nx + width // (2 * nx)
    ys = (np.arange(ny) * height) // ny + height // (2 * ny)
    return (ys[:, None] * width + xs[None, :]).reshape(-1)


BATCH_SIZES = ((256, 256), (512, 512), (1024, 576), (1920, 1080))


def icosphere_scene(triangles, seed, emitter_fraction=0.02):
    """An icosphere field of about `triangles` triangles (subdivision 3, 1280 a sphere), radii
    lognormal(0, 0.5), centres uniform in a cube whose side keeps the density of 1024 spheres in
    [-50, 50]^3; spheres take one of 8 Lambert albedos, and a seeded 2% of them (at least 2) emit
    white light of radiance 4 to 16.  Returns (mesh, materials, half_extent)."""
    import vt
    spheres = max(2, int(round(triangles / 1280)))
    half = 50.0 * (spheres / 1024.0) ** (1.0 / 3.0)
    mesh = vt.icosphere_field(spheres=spheres, subdivisions=3, mu=0.0, sigma=0.5, half_extent=half, seed=seed)
    rng = np.random.default_rng(seed)
    per = len(mesh.faces) // spheres
    tags = rng.integers(0, 8, spheres).astype(np.uint32)
    lit = rng.choice(spheres, max(2, int(round(emitter_fraction * spheres))), replace=False)
    tags[lit] = 8 + np.arange(len(lit), dtype=np.uint32)
    mesh.faces[:, 3] = np.repeat(tags, per)
    mats = [dict(type=0, albedo=tuple(0.2 + 0.6 * rng.random(3))) for _ in range(8)]
    for _ in lit:
        L = float(rng.uniform(4.0, 16.0))
        mats.append(dict(type=3, albedo=(0.0, 0.0, 0.0), emission=(L, L, L)))
    return mesh, mats, half


def batch_workload(count=64, seed=11, tri_range=(10_000, 2_000_000)):
    """The batched workload: `count` icosphere scenes with triangle counts log-uniform over
    tri_range, each viewed at a size drawn from BATCH_SIZES.  Returns a list of
    (mesh, materials, width, height, eye, target, fov)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        tris = float(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
        w, h = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
        mesh, mats, half = icosphere_scene(tris, seed * 1000 + k)
        eye = (0.0, 0.6 * half, -3.4 * half)
        out.append((mesh, mats, w, h, eye, (0.0, 0.0, 0.0), 45.0))
    return out
