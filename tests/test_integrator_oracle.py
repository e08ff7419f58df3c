"""CPU tests that pin the integrator oracle (oracle/integrator.c): closed forms and precision."""
import math

import numpy as np
import pytest

import oracle
import scenes
import vt

CORNELL_MATS = [dict(type=0, albedo=(0.73, 0.73, 0.73)), dict(type=0, albedo=(0.65, 0.05, 0.05)),
                dict(type=0, albedo=(0.12, 0.45, 0.15)), dict(type=3, albedo=(0, 0, 0), emission=(17, 17, 17))]


def _quad(c, u, v):
    c, u, v = (np.asarray(x, np.float32) for x in (c, u, v))
    return [[c, c + u, c + u + v], [c, c + u + v, c + v]]


def _rect_form_factor(X, Y):
    # Point below one corner of an X by Y rectangle at unit height, parallel (Howell's catalogue).
    a, b = X / math.sqrt(1 + X * X), Y / math.sqrt(1 + Y * Y)
    return (a * math.atan(Y / math.sqrt(1 + X * X)) + b * math.atan(X / math.sqrt(1 + Y * Y))) / (2 * math.pi)


def test_direct_light_matches_form_factor_closed_form():
    # A white floor under a square D65 light of radiance 17: with depth 1 (direct light only) the
    # radiance leaving the floor point below the light's centre is rho * L * F, F the point-to-square
    # form factor; a 1-pixel, 0.05-degree camera sees exactly that point.
    h, a = 1.0, 0.5
    tris = _quad((-2, 0, -2), (0, 0, 4), (4, 0, 0))               # floor y = 0 (two-sided)
    tris += [[np.float32([-a / 2, h, -a / 2]), np.float32([a / 2, h, a / 2]), np.float32([-a / 2, h, a / 2])],
             [np.float32([-a / 2, h, -a / 2]), np.float32([a / 2, h, -a / 2]), np.float32([a / 2, h, a / 2])]]
    v, f = scenes.mesh_from_triangles(np.array(tris, np.float32))
    f[2:, 3] = 1
    sc = oracle.Scene(v, f)
    mats = [dict(type=0, albedo=(0.5, 0.5, 0.5)), dict(type=3, albedo=(0, 0, 0), emission=(17, 17, 17))]
    frame = dict(width=1, height=1, eye=(0.0, 0.6, 0.0), target=(0.0, -1.0, 1e-4), fov=0.05, max_depth=1)
    got = oracle.render(sc, mats, frame, 1, 200000, precision="f64")[0]
    F = 4 * _rect_form_factor(a / 2 / h, a / 2 / h)
    expect = 0.5 * 17.0 * F
    # 200k samples of four wavelengths: colour noise of the spectral estimate stays below 1%.
    assert np.allclose(got, expect, rtol=0.01), (got, expect)


def test_f32_reference_integrator_agrees_with_fp64():
    m = vt.cornell()
    sc = oracle.Scene(m.vertices, m.faces)
    frame = dict(width=48, height=48, eye=vt.orbit_eye(), target=(0, -0.1, 0), fov=42.0, max_depth=8)
    a = oracle.render(sc, CORNELL_MATS, frame, 1, 16, precision="f32")
    b = oracle.render(sc, CORNELL_MATS, frame, 1, 16, precision="f64")
    assert abs(a.mean() / b.mean() - 1) < 5e-4


def test_camera_sees_the_light_exactly():
    # Camera rays that hit the emitter return its RGB exactly (emitter_rgb with D65, main.swift:955-957).
    m = vt.cornell()
    sc = oracle.Scene(m.vertices, m.faces)
    frame = dict(width=1, height=1, eye=(0.0, 0.0, 0.0), target=(0.0, 1.0, 1e-4), fov=0.5, max_depth=8)
    got = oracle.render(sc, CORNELL_MATS, frame, 1, 4)[0]
    assert np.array_equal(got, [17.0, 17.0, 17.0])


def test_depth_limit_and_energy_monotone():
    # More path vertices only add light: mean radiance is non-decreasing in the depth.
    m = vt.cornell()
    sc = oracle.Scene(m.vertices, m.faces)
    means = []
    for depth in (1, 2, 4, 8):
        frame = dict(width=32, height=32, eye=vt.orbit_eye(), target=(0, -0.1, 0), fov=42.0, max_depth=depth)
        means.append(oracle.render(sc, CORNELL_MATS, frame, 1, 16).mean())
    assert all(b > a for a, b in zip(means, means[1:]))


def test_z_sampler_mode_is_an_unbiased_estimator_of_the_same_image():
    """Z mode (Sampler in Z mode, main.swift:213-293) changes only where the numbers come from: the
    image it converges to is the PCG mode's.  At 64 spp on a 24x24 Cornell box the two means agree to
    sampling noise, and the Z image is not a copy of the PCG one."""
    mesh = vt.cornell()
    sc = oracle.Scene(mesh.vertices, mesh.faces)
    base = dict(width=24, height=24, eye=vt.orbit_eye(), target=(0.0, -0.1, 0.0), fov=42.0, max_depth=4)
    pcg = oracle.render(sc, CORNELL_MATS, base, 1, 64)
    z = oracle.render(sc, CORNELL_MATS, dict(base, sampler=1, z_temporal=1), 1, 64)
    assert np.isfinite(z).all()
    assert abs(z.mean() / pcg.mean() - 1) < 0.03, (z.mean(), pcg.mean())
    assert not np.allclose(z, pcg)
    z3 = oracle.render(sc, CORNELL_MATS, dict(base, sampler=1, z_temporal=3), 1, 64)
    assert abs(z3.mean() / pcg.mean() - 1) < 0.03
    assert not np.allclose(z3, z)
