"""CPU tests of the oracle's environment lighting (oracle/environment.c, the environment part of
oracle/integrator.c), the checker of the library's environment path.

Pins, each independent of the oracle's own formulas:
- the sampling tables against a double-precision numpy statement of the reference's construction
  (luminance, 3 x 3 dilation wrapping in longitude, the sin-weighted mean, MIS compensation), cell by
  cell, and a map without texels above its mean sampling its luminance in both tables;
- the light-sample density integrating to one over the sphere, agreeing with the density the sampler
  reports for its own samples, and the samples' cell histogram matching the tables;
- the lookup against a map that is linear in both texel coordinates (bilinear filtering reproduces it
  exactly), with rotation and strength;
- the integrator on a white-furnace closed form: a flat-shaded convex sphere under a uniform sky
  reflects albedo x the sky's spectral round trip, in MIS, NEE-only and BSDF-only modes;
- the three sampling modes converging to one image with an uneven sky plus a triangle emitter (the
  emitter / environment split and the compensated table's MIS weights), and ReSTIR DI against MIS;
- planted mistakes (the density without its 1 / sin theta; the compensated table used where no BSDF
  continuation covers its zeros) breaking the agreement.
"""
import math

import numpy as np
import pytest

import oracle
import scenes
import vt


def _frame(w, h, depth, mode=1, env=None, strength=1.0, rotation=0.0, eye=scenes.ENV_EYE, target=scenes.ENV_TARGET,
           fov=45.0, **kw):
    f = dict(width=w, height=h, eye=eye, target=target, fov=fov, max_depth=depth, sampling_mode=mode,
             environment=env, env_strength=strength, env_rotation=rotation)
    f.update(kw)
    return f


def _numpy_cell_probabilities(img):
    """The reference's table construction (Sources/StudioModel.swift:1300-1352) in double."""
    h, w = img.shape[:2]
    lum = np.maximum(1e-6, 0.2126 * img[..., 0].astype(np.float64) + 0.7152 * img[..., 1] + 0.0722 * img[..., 2])
    hor = np.maximum(lum, np.maximum(np.roll(lum, 1, axis=1), np.roll(lum, -1, axis=1)))
    above = np.vstack([hor[:1], hor[:-1]])
    below = np.vstack([hor[1:], hor[-1:]])
    dil = np.maximum(hor, np.maximum(above, below))
    st = np.maximum(1e-4, np.sin(np.pi * (np.arange(h) + 0.5) / h))[:, None]
    mean = (lum.sum(1) * st[:, 0]).sum() / (st[:, 0].sum() * w)
    comp = np.maximum(0.0, dil - mean) * st
    orig = dil * st
    if not comp.sum() > 0:
        comp = orig.copy()
    out = []
    for wts in (comp, orig):
        out.append((wts.sum(1) / wts.sum(), wts / np.maximum(wts.sum(1, keepdims=True), 1e-300)))
    return out, dil, mean


def _table_probabilities(rows, cols):
    out = []
    for t in range(2):
        pr = np.diff(np.concatenate([[0.0], rows[:, t].astype(np.float64)]))
        pc = np.diff(np.concatenate([np.zeros((rows.shape[0], 1)), cols[t].astype(np.float64)], axis=1), axis=1)
        out.append((pr, pc))
    return out


def test_tables_follow_the_reference_construction():
    rng = np.random.default_rng(5)
    img = np.zeros((12, 24, 4), np.float32)
    img[..., :3] = rng.uniform(0.0, 1.0, (12, 24, 3))
    img[3, 7, :3] = 40.0
    img[9, 0, :3] = (0.0, 25.0, 0.0)   # next to the longitude seam: the dilation wraps
    rows, cols = oracle.env_tables(img)
    for t in range(2):
        assert rows[-1, t] == 1.0 and np.all(np.diff(rows[:, t]) >= 0)
        assert np.all(cols[t, :, -1] == 1.0) and np.all(np.diff(cols[t], axis=1) >= 0)
    expect, dil, mean = _numpy_cell_probabilities(img)
    got = _table_probabilities(rows, cols)
    for t in range(2):
        assert np.allclose(got[t][0], expect[t][0], rtol=2e-5, atol=1e-6), t  # float running sums
        assert np.allclose(got[t][1], expect[t][1], rtol=2e-5, atol=1e-6), t
    # The compensated table leaves exactly the texels at or below the mean to BSDF sampling (a row
    # with none above it has probability 0 and is never searched).
    live = got[0][0] > 0
    assert np.array_equal(got[0][0] > 0, (dil > mean).any(1))
    above = (dil > mean)[live]
    assert np.all(got[0][1][live][above] > 0)
    # (the last entry of a CDF row is set to 1, so a zero-weight last texel keeps the row's rounding)
    assert np.all(got[0][1][live][~above] < 1e-6)
    assert (dil <= mean).sum() > 40
    # The seam texel's dilation reaches the last column of its row and the rows above and below.
    assert expect[1][1][9, -1] > 10 * expect[1][1][9, 5]


def test_uniform_map_samples_its_luminance_in_both_tables():
    img = np.full((8, 16, 4), 0.5, np.float32)
    rows, cols = oracle.env_tables(img)
    assert np.array_equal(rows[:, 0], rows[:, 1]) and np.array_equal(cols[0], cols[1])
    pr = np.diff(np.concatenate([[0.0], rows[:, 1].astype(np.float64)]))
    st = np.sin(np.pi * (np.arange(8) + 0.5) / 8)
    assert np.allclose(pr, st / st.sum(), rtol=1e-5)


def _uv_directions(w, h, rotation, sub=4):
    # Directions at sub x sub points per texel, in the map's (u, v) parameterisation.
    uu = (np.arange(w * sub) + 0.5) / (w * sub)
    vv = (np.arange(h * sub) + 0.5) / (h * sub)
    U, Vv = np.meshgrid(uu, vv)
    theta = Vv * np.pi
    phi = (U - 0.5 - rotation / (2 * np.pi)) * 2 * np.pi
    d = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], -1)
    return d.reshape(-1, 3), theta.reshape(-1), (np.pi / (h * sub)) * (2 * np.pi / (w * sub))


@pytest.mark.parametrize("table", [0, 1])
def test_density_integrates_to_one_and_matches_sampling(table):
    img = scenes.sky_map(32, 16)
    mesh, mats = scenes.environment_scene(emitter=False)
    sc = oracle.Scene(mesh.vertices, mesh.faces, with_bvh=False)
    fr = _frame(8, 8, 3, env=img, strength=1.3, rotation=0.7)
    d, theta, dA = _uv_directions(32, 16, 0.7)
    _, pdf = oracle.env_query(sc, mats, fr, d, table=table)
    total = float((pdf.astype(np.float64) * np.sin(theta) * dA).sum())
    assert abs(total - 1.0) < 2e-3, total
    n = 200000
    s = oracle.light_samples(sc, mats, fr, (0.0, 0.0, 0.0), n, seed=9, table=table)
    assert np.all(s["light"] == 1)
    _, back = oracle.env_query(sc, mats, fr, s["wi"], table=table)
    agree = np.isclose(back, s["pdf"], rtol=1e-3)
    assert agree.mean() > 0.999, agree.mean()
    # Histogram of the sampled cells against the tables' cell probabilities.
    rows, cols = oracle.env_tables(img)
    pr, pc = _table_probabilities(rows, cols)[table]
    prob = pr[:, None] * pc
    wi = s["wi"].astype(np.float64)
    u = (np.arctan2(wi[:, 2], wi[:, 0]) / (2 * np.pi) + 0.5 + 0.7 / (2 * np.pi)) % 1.0
    v = np.arccos(np.clip(wi[:, 1], -1, 1)) / np.pi
    x = np.minimum((u * 32).astype(int), 31)
    y = np.minimum((v * 16).astype(int), 15)
    freq = np.bincount(y * 32 + x, minlength=512).reshape(16, 32) / n
    sigma = np.sqrt(np.maximum(prob, 1.0 / n) / n)
    assert np.all(np.abs(freq - prob) < 6 * sigma + 2e-5)


def test_lookup_is_bilinear_with_rotation_and_strength():
    w, h = 32, 16
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = np.arange(w, dtype=np.float32)[None, :] * 0.25 + 1.0   # linear in the column
    img[..., 1] = np.arange(h, dtype=np.float32)[:, None] * 0.5 + 2.0    # linear in the row
    img[..., 2] = 3.0
    mesh, mats = scenes.environment_scene(emitter=False)
    sc = oracle.Scene(mesh.vertices, mesh.faces, with_bvh=False)
    rng = np.random.default_rng(3)
    rot = 0.4
    # Directions whose texel coordinates stay inside the map (no seam, no pole clamp).
    fx = rng.uniform(0.5, w - 1.5, 4000)
    fy = rng.uniform(0.5, h - 1.5, 4000)
    u = (fx + 0.5) / w
    v = (fy + 0.5) / h
    theta = v * np.pi
    phi = (u - 0.5 - rot / (2 * np.pi)) * 2 * np.pi
    d = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], -1).astype(np.float32)
    rgb, _ = oracle.env_query(sc, mats, _frame(8, 8, 2, env=img, strength=2.5, rotation=rot), d, table=1)
    assert np.allclose(rgb[:, 0], 2.5 * (fx * 0.25 + 1.0), rtol=2e-4, atol=2e-4)
    assert np.allclose(rgb[:, 1], 2.5 * (fy * 0.5 + 2.0), rtol=2e-4, atol=2e-4)
    assert np.allclose(rgb[:, 2], 7.5, rtol=1e-6)


def _round_trip(c):
    """E over the wavelength number of spectrum_rgb(spectral_emission(grey c)): the sky's colour as
    the spectral integrator carries it (D65 basis, four hero wavelengths from the ICDF)."""
    _, (icdf, cmf, basis) = oracle.spectral_data()
    n = 1 << 16
    u = (np.arange(n) + 0.5) / n
    acc = np.zeros(3)
    for k in range(4):
        x = (u + k) * 0.25 * 1024
        i = np.minimum(x.astype(np.int64), 1023)
        low = icdf[i].astype(np.float64)
        width = icdf[i + 1].astype(np.float64) - low
        weight = 0.25 * 1024 * width
        lam = low + (x - i) * width
        b = np.minimum(np.clip(lam - 360.0 + 0.5, 0, 470).astype(np.int64), 470)
        s = c * basis[b, 0, 3].astype(np.float64)
        acc += (s * weight) @ cmf[b, :3].astype(np.float64) / n
    return acc


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_white_furnace_sphere(mode):
    img = np.full((16, 32, 4), 0.8, np.float32)
    strength = 1.5
    mesh = scenes.furnace_sphere()
    mats = [dict(type=0, albedo=(0.6, 0.6, 0.6))]
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    w = 48
    fr = _frame(w, w, 4, mode=mode, env=img, strength=strength, eye=(0.0, 0.0, -4.0), target=(0.0, 0.0, 0.0),
                fov=40.0)
    out = oracle.render(sc, mats, fr, 1, 64).reshape(w, w, 3)
    c = np.arange(w) + 0.5
    ndc = (c / w) * 2 - 1
    r = np.hypot(ndc[None, :], ndc[:, None])
    inner = r < 0.55
    outer = r > 0.85
    expect = 0.6 * _round_trip(0.8 * strength)
    got = out[inner].mean(0)
    tol = 0.03 if mode == 2 else 0.006
    assert np.all(np.abs(got / expect - 1) < tol), (got, expect)
    # Camera rays that miss see the sky's exact colour.
    assert np.allclose(out[outer], 0.8 * strength, rtol=1e-6)


def _mode_means(depth, samples=64, fault=0, sun=300.0, **kw):
    img = scenes.sky_map(sun_radiance=sun)
    mesh, mats = scenes.environment_scene(emitter=True)
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    oracle.set_fault(fault)
    try:
        return {m: oracle.render(sc, mats, _frame(48, 36, depth, mode=m, env=img, **kw), 1, samples).mean()
                for m in (1, 2, 3)}
    finally:
        oracle.set_fault(0)


@pytest.mark.parametrize("depth", [1, 3])
def test_sampling_modes_agree_with_sky_and_emitter(depth):
    # A dimmer sun keeps BSDF-only sampling's variance low enough for a 1% bar (measured within 0.4%).
    m = _mode_means(depth, samples=256, sun=30.0)
    assert abs(m[2] / m[1] - 1) < 0.01, m
    assert abs(m[3] / m[1] - 1) < 0.01, m


def test_planted_environment_mistakes_are_caught():
    # Without 1 / sin(theta) the NEE density is wrong: NEE-only and BSDF-only disagree.
    bad = _mode_means(1, samples=128, fault=11)
    assert abs(bad[2] / bad[3] - 1) > 0.05, bad
    # The compensated table at depth 1 leaves the texels below the mean to a BSDF sample that is never
    # taken: NEE-only comes out dark.
    dark = _mode_means(1, samples=128, fault=12)
    good = _mode_means(1, samples=128)
    assert dark[2] / good[2] < 0.95, (dark, good)


@pytest.mark.parametrize("depth", [1, 3])
def test_restir_di_with_sky_matches_mis(depth):
    img = scenes.sky_map()
    mesh, mats = scenes.environment_scene(emitter=True)
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    fr = _frame(48, 36, depth, env=img, indirect_reuse=1)
    a = oracle.render(sc, mats, fr, 1, 128).mean()
    b = oracle.render_restir(sc, mats, fr, 1, 128).mean()
    assert abs(b / a - 1) < 0.025, (a, b)


def test_environment_off_leaves_the_image_unchanged():
    # Strength 0 (or no map) is the scene without an environment: camera misses are black and the
    # emitters take every light sample.
    mesh, mats = scenes.environment_scene(emitter=True)
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    a = oracle.render(sc, mats, _frame(24, 16, 3, env=None), 1, 8)
    b = oracle.render(sc, mats, _frame(24, 16, 3, env=scenes.sky_map(), strength=0.0), 1, 8)
    assert np.array_equal(a, b)


def test_library_tables_equal_the_oracle_bit_for_bit():
    # The library's host construction (csrc/environment.cpp) and the oracle's restatement follow the
    # same float operations, so the tables agree exactly; sampling on the GPU then picks the oracle's cells.
    rng = np.random.default_rng(11)
    for h, w in ((16, 32), (7, 5), (1, 9), (33, 64)):
        img = np.zeros((h, w, 4), np.float32)
        img[..., :3] = rng.lognormal(0.0, 1.5, (h, w, 3)).astype(np.float32)
        a = vt.environment_tables(img)
        b = oracle.env_tables(img)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (h, w)
    flat = np.full((4, 8, 4), 2.0, np.float32)
    a, b = vt.environment_tables(flat), oracle.env_tables(flat)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    with pytest.raises(vt.VTError):
        vt.environment_tables(np.zeros((8193, 1, 4), np.float32))
