"""Small deterministic test scenes and ray sets shared by the CPU and GPU tests."""
import numpy as np


def mesh_from_triangles(tris: np.ndarray):
    """(F, 3, 3) float32 corner positions -> (vertices (3F, 8), faces (F, 4)), unwelded."""
    f = len(tris)
    v = np.zeros((3 * f, 8), np.float32)
    v[:, :3] = tris.reshape(-1, 3)
    v[:, 4] = 1.0
    faces = np.zeros((f, 4), np.uint32)
    faces[:, :3] = np.arange(3 * f, dtype=np.uint32).reshape(f, 3)
    return v, faces


def patch(cells=16):
    """The curved height-field patch of tests/Fix_renderer-followups.swift:14-35: a welded grid
    with alternating diagonals.  Returns (vertices (V, 8), faces (F, 4), grid positions)."""
    idx = np.arange(cells + 1, dtype=np.float32)
    x = idx[:, None] / np.float32(cells) * np.float32(2) - np.float32(1)
    z = idx[None, :] / np.float32(cells) * np.float32(2) - np.float32(1)
    x, z = np.broadcast_arrays(x, z)
    y = (np.float32(0.15) * np.sin(np.float32(1.7) * x + np.float32(0.3)) * np.cos(np.float32(1.3) * z)
         + np.float32(0.05) * x).astype(np.float32)
    grid = np.stack([x, y, z], -1).astype(np.float32)
    verts = np.zeros(((cells + 1) ** 2, 8), np.float32)
    verts[:, :3] = grid.reshape(-1, 3)
    verts[:, 4] = 1.0

    def vid(i, j):
        return i * (cells + 1) + j

    faces = []
    for i in range(cells):
        for j in range(cells):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            if (i + j) % 2 == 0:
                faces += [(a, b, c, 0), (a, c, d, 0)]
            else:
                faces += [(a, b, d, 0), (b, c, d, 0)]
    return verts, np.array(faces, np.uint32), grid


def edge_targets(grid, cells=16):
    """Interior vertices and the 1/4, 1/2, 3/4 points of every shared edge
    (tests/Fix_renderer-followups.swift:208-218)."""
    t = [grid[i, j] for i in range(1, cells) for j in range(1, cells)]

    def edge(a, b):
        for f in (0.25, 0.5, 0.75):
            t.append(a + (b - a) * np.float32(f))

    for i in range(cells):
        for j in range(cells):
            if j > 0:
                edge(grid[i, j], grid[i + 1, j])
            if i > 0:
                edge(grid[i, j], grid[i, j + 1])
            if (i + j) % 2 == 0:
                edge(grid[i, j], grid[i + 1, j + 1])
            else:
                edge(grid[i + 1, j], grid[i, j + 1])
    return np.array(t, np.float32)


EYES = np.array([[0.3, 3, 0.2], [4, 3.5, -3], [-7, 6, 5], [0.05, 20, 0.01], [2.5, 1.8, 1.7], [-0.4, 9, -12]],
                np.float32)


def aimed_rays(eyes, targets, tmin=0.0, tmax=1e20):
    """(8, len(eyes) * len(targets)) SoA rays from every eye through every target."""
    o = np.repeat(eyes, len(targets), axis=0)
    tg = np.tile(targets, (len(eyes), 1))
    d = tg - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r = np.empty((8, len(o)), np.float32)
    r[0:3] = o.T
    r[3:6] = d.T
    r[6] = tmin
    r[7] = tmax
    return r


def scattered_triangles(n=3000, seed=5):
    """Random small triangles in [-1.2, 1.2]^3 (tests/Fix_geometry.swift:141-161 style)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.2, 1.2, (n, 1, 3))
    e = rng.uniform(-0.08, 0.08, (n, 3, 3))
    return mesh_from_triangles((c + e).astype(np.float32))


def random_rays(n, seed, box=1.2, tmin=0.0, tmax=1e20):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-box, box, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r = np.empty((8, n), np.float32)
    r[0:3] = o.T
    r[3:6] = d.T
    r[6] = tmin
    r[7] = tmax
    return r


def axis_rays(n, seed, box=1.2):
    """Rays with exactly axis-aligned directions (the parallel-slab branches)."""
    r = random_rays(n, seed, box)
    rng = np.random.default_rng(seed + 1)
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n).astype(np.float32)
    r[3:6] = 0
    r[3 + axis, np.arange(n)] = sign
    return r


CORNELL_MATS = [dict(type=0, albedo=(0.73, 0.73, 0.73)), dict(type=0, albedo=(0.65, 0.05, 0.05)),
                dict(type=0, albedo=(0.12, 0.45, 0.15)), dict(type=3, albedo=(0, 0, 0), emission=(17, 17, 17))]


def _ceiling_quad(x0, z0, s, y, tag, flip=False):
    """Two triangles of an s x s square at height y; normal -y (facing down) unless flip."""
    quad = np.array([[x0, y, z0], [x0 + s, y, z0 + s], [x0, y, z0 + s],
                     [x0, y, z0], [x0 + s, y, z0], [x0 + s, y, z0 + s]], np.float32)
    if flip:
        quad = quad[[0, 2, 1, 3, 5, 4]]
    v = np.zeros((6, 8), np.float32)
    v[:, :3] = quad
    v[:, 4] = 1.0 if flip else -1.0
    return v, np.array([[0, 1, 2, tag], [3, 4, 5, tag]], np.uint32)


def _append(base, parts, mats, extra_mats):
    import vt
    verts, faces = [base.vertices], [base.faces]
    nv = len(base.vertices)
    for v, f in parts:
        f = f.copy()
        f[:, :3] += nv
        verts.append(v)
        faces.append(f)
        nv += len(v)
    return vt.Mesh(np.concatenate(verts), np.concatenate(faces)), list(mats) + list(extra_mats)


def multi_light_cornell():
    """The Cornell box with a 0.25 ceiling light plus a 4 x 4 grid of small ceiling lights of
    different intensities (tags 4-19)."""
    import vt
    parts, mats = [], []
    k = 0
    for i in range(4):
        for j in range(4):
            parts.append(_ceiling_quad(-0.8 + 0.45 * i, -0.8 + 0.45 * j, 0.12, 0.998, 4 + k))
            e = 2.0 + 3.0 * ((i * 4 + j) % 5)
            mats.append(dict(type=3, albedo=(0, 0, 0), emission=(e, e, e)))
            k += 1
    return _append(vt.cornell(0.25), parts, CORNELL_MATS, mats)


def backfacing_light_cornell():
    """The Cornell box plus a 0.25 emitter lying on the floor facing down (tag 4): it lights nothing
    inside the box, so half of the power-sampled light candidates have pdf 0."""
    import vt
    return _append(vt.cornell(0.25), [_ceiling_quad(0.3, 0.3, 0.25, -0.998, 4)], CORNELL_MATS,
                   [dict(type=3, albedo=(0, 0, 0), emission=(17, 17, 17))])


def materials_cornell(variant="layered", light_size=0.25):
    """The Cornell box with its tall box (tag 4), short box (tag 5) and floor (tag 6) given other
    materials.  'layered': rough gold-like metal (GLOSSY, roughness 0.3), rough glass (OPENPBR,
    transmission 1, roughness 0.2, IOR 1.5) and a glossy plastic floor (OPENPBR, roughness 0.4);
    'delta': a mirror (GLOSSY, roughness 0) and smooth glass (DIELECTRIC, IOR 1.5), Lambert floor;
    'opaque': the metal, a half-metallic blue OPENPBR box (roughness 0.25, IOR 1.6) and the plastic floor;
    'coated': coats (roughness 0.12, IOR 1.6) over them -- a fully coated metal, a half-coated rough glass,
    and a red plastic floor under a 0.7 coat; 'anisotropic': a brushed metal, an anisotropic rough glass
    and a half-metallic floor with a half coat, their specular lobes stretched along each face's tangent;
    'fuzzy': sheen (roughness 0.5) over a metal, a half-coated blue plastic and a rough glass."""
    import vt
    base = vt.cornell(light_size)
    faces = base.faces.copy()
    faces[10:22, 3] = 4
    faces[22:34, 3] = 5
    ys = base.vertices[faces[:, :3], 1]
    floor = np.all(ys == ys.min(), axis=1) & (faces[:, 3] == 0)
    faces[floor, 3] = 6
    mats = list(CORNELL_MATS)
    if variant == "opaque":
        mats += [dict(type=1, albedo=(0.95, 0.64, 0.54), roughness=0.3),
                 dict(type=4, albedo=(0.2, 0.4, 0.8), roughness=0.25, ior=1.6, metalness=0.5),
                 dict(type=4, albedo=(0.73, 0.73, 0.73), roughness=0.4, ior=1.5)]
    elif variant == "coated":
        mats += [dict(type=1, albedo=(0.95, 0.64, 0.54), roughness=0.3, coat=1.0),
                 dict(type=4, albedo=(1.0, 1.0, 1.0), roughness=0.3, ior=1.5, transmission=1.0, coat=0.5),
                 dict(type=4, albedo=(0.7, 0.15, 0.1), roughness=0.5, ior=1.5, coat=0.7)]
    elif variant == "anisotropic":
        mats += [dict(type=1, albedo=(0.95, 0.64, 0.54), roughness=0.35, anisotropy=0.8),
                 dict(type=4, albedo=(1.0, 1.0, 1.0), roughness=0.3, ior=1.5, transmission=1.0, anisotropy=0.5),
                 dict(type=4, albedo=(0.2, 0.4, 0.8), roughness=0.4, ior=1.5, metalness=0.6, anisotropy=0.7, coat=0.5)]
    elif variant == "fuzzy":
        mats += [dict(type=1, albedo=(0.95, 0.64, 0.54), roughness=0.4, fuzz=0.8),
                 dict(type=4, albedo=(0.3, 0.3, 0.8), roughness=0.5, ior=1.5, fuzz=1.0, coat=0.5),
                 dict(type=4, albedo=(1.0, 1.0, 1.0), roughness=0.3, ior=1.5, transmission=1.0, fuzz=0.6)]
    elif variant == "layered":
        mats += [dict(type=1, albedo=(0.95, 0.64, 0.54), roughness=0.3),
                 dict(type=4, albedo=(1.0, 1.0, 1.0), roughness=0.2, ior=1.5, transmission=1.0),
                 dict(type=4, albedo=(0.73, 0.73, 0.73), roughness=0.4, ior=1.5)]
    else:
        mats += [dict(type=1, albedo=(0.9, 0.9, 0.9), roughness=0.0),
                 dict(type=2, albedo=(1.0, 1.0, 1.0), ior=1.5),
                 dict(type=0, albedo=(0.73, 0.73, 0.73))]
    verts = base.vertices
    if variant == "anisotropic":
        # uv = (x + 0.3 y, z - 0.2 y): every face gets a tangent dP/du from its uv (none degenerate)
        verts = verts.copy()
        verts[:, 6] = verts[:, 0] + 0.3 * verts[:, 1]
        verts[:, 7] = verts[:, 2] - 0.2 * verts[:, 1]
    return vt.Mesh(verts, faces), mats


def vt_materials(mats):
    import vt
    return [vt.material(m["type"], m["albedo"], m.get("emission", (0, 0, 0)), m.get("roughness", 0.5),
                        m.get("ior", 1.5), m.get("metalness", 0.0), 1.0, m.get("transmission", 0.0), m.get("coat", 0.0),
                        m.get("anisotropy", 0.0), m.get("fuzz", 0.0)) for m in mats]


def sky_map(width=64, height=32, sun=(0.3, 0.28), sun_radiance=300.0, seed=2):
    """A lat-long rgba32f sky (row 0 at the zenith): blue toward the zenith and whiter toward the
    horizon, a dark brown ground below it, a bright 2 x 2-texel sun at (u, v) = `sun`, and mild
    per-texel variation except in the two pole rows, which are constant as in any map of a real sky
    (a pole is a single direction; a varying pole row would make lookups there depend on an azimuth
    that rounding decides)."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height, dtype=np.float64) + 0.5) / height
    up = np.cos(v * np.pi)[:, None] * np.ones((1, width))
    sky = np.stack([0.35 + 0.5 * (1 - up), 0.55 + 0.4 * (1 - up), np.full_like(up, 1.0)], -1)
    ground = np.broadcast_to(np.array([0.18, 0.12, 0.08]), sky.shape)
    noise = rng.uniform(0.85, 1.15, (height, width, 1))
    noise[[0, -1]] = 1.0
    img = np.where((up > 0)[..., None], sky, ground) * noise
    sx, sy = int(sun[0] * width), int(sun[1] * height)
    img[sy:sy + 2, sx:sx + 2] = (sun_radiance, 0.9 * sun_radiance, 0.75 * sun_radiance)
    out = np.zeros((height, width, 4), np.float32)
    out[..., :3] = img
    out[..., 3] = 1.0
    return out


def environment_scene(emitter=True, seed=4):
    """An open scene for environment lighting: a grey ground quad at y = -1 (tag 0), six diffuse
    icospheres (tag 1) and, with `emitter`, a small downward ceiling light at y = 1.6 (tag 2)."""
    import vt
    ground = np.array([[[-6, -1, -6], [6, -1, 6], [6, -1, -6]], [[-6, -1, -6], [-6, -1, 6], [6, -1, 6]]], np.float32)
    gv, gf = mesh_from_triangles(ground)
    spheres = vt.icosphere_field(spheres=6, subdivisions=2, mu=-0.7, sigma=0.2, half_extent=1.4, seed=seed)
    sf = spheres.faces.copy()
    sf[:, 3] = 1
    parts = [(spheres.vertices, sf)]
    mats = [dict(type=0, albedo=(0.6, 0.6, 0.6)), dict(type=0, albedo=(0.7, 0.45, 0.3))]
    if emitter:
        parts.append(_ceiling_quad(-0.3, -0.3, 0.6, 1.6, 2))
        mats.append(dict(type=3, albedo=(0, 0, 0), emission=(6.0, 6.0, 6.0)))
    return _append(vt.Mesh(gv, gf), parts, mats, [])


def furnace_sphere():
    """One unit icosphere at the origin with flat shading (vertex normals zero, so the shading normal is
    the geometric one and the convex body never sees itself), tag 0."""
    import vt
    s = vt.icosphere_field(spheres=1, subdivisions=3, mu=0.0, sigma=0.0, half_extent=1e-3, seed=1)
    v = s.vertices.copy()
    v[:, 3:6] = 0.0
    f = s.faces.copy()
    f[:, 3] = 0
    return vt.Mesh(v, f)


ENV_EYE = (0.0, 1.2, -6.5)
ENV_TARGET = (0.0, -0.3, 0.0)
