"""CPU tests of the spectral setup: the library's host tables against the oracle's restatement of the
reference generator, both against the reference's own uplift (tests/golden/spectral_uplift.npz, made by
running scripts/generate_spectral_tables.py), and both against what the mathematics fixes."""
import math
import os

import numpy as np
import pytest

import oracle
import vt
from oracle import spectral


@pytest.fixture(scope="module")
def data():
    return oracle.spectral_data()


def test_library_tables_match_oracle(data):
    d, (icdf, cmf, basis) = data
    i2, c2, b2 = vt.spectral_tables()
    assert np.array_equal(i2, icdf)
    assert np.array_equal(c2, cmf)
    assert np.allclose(b2, basis, atol=1e-6)


def test_icdf_inverts_the_density_exactly(data):
    # F(F^-1(n / 1024)) = n / 1024 for the piecewise-linear 1 nm density (main.swift:8171-8184).
    d, _ = data
    icdf, density = d.inverse_cdf()
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (density[:-1] + density[1:]))])
    total = cdf[-1]
    assert abs(total - 1.0) < 1e-12   # the density integrates to one
    for n in range(1, 1024, 37):
        lam = icdf[n]
        j = min(int(lam - 360), 469)
        t = lam - 360 - j
        p0, p1 = density[j], density[j + 1]
        f = cdf[j] + p0 * t + 0.5 * (p1 - p0) * t * t
        assert abs(f - n / 1024) < 1e-9
    assert np.all(np.diff(icdf) > 0)


def _lanes(icdf, u):
    """spectral_lambda (main.swift:795-801) for the four lanes of u."""
    out = []
    for k in range(4):
        x = (u + k) * 0.25 * 1024
        i = min(int(x), 1023)
        low, width = icdf[i], icdf[i + 1] - icdf[i]
        lam = low + (x - i) * width
        out.append((min(int(min(max(lam - 360 + 0.5, 0), 470)), 470), 0.25 * 1024 * width))
    return out


def test_wavelength_estimator_is_unbiased(data):
    # E_u[sum_k w_k c(bin(lambda_k))] = sum over 1 nm nodes of c: spectrum_rgb of a flat unit
    # spectrum integrates the CMFs; of normalised D65 it returns the white point (1, 1, 1).
    d, (icdf, cmf, basis) = data
    us = (np.arange(20000) + 0.5) / 20000
    flat = np.zeros(3)
    white = np.zeros(3)
    for u in us:
        for b, w in _lanes(icdf.astype(np.float64), u):
            flat += w * cmf[b, :3]
            white += w * cmf[b, :3] * basis[b, 0, 3]
    flat /= len(us)
    white /= len(us)
    assert np.allclose(flat, cmf[:, :3].sum(0), rtol=2e-3)
    assert np.allclose(white, 1.0, rtol=2e-3)


def test_white_point_and_500nm(data):
    d, (icdf, cmf, basis) = data
    # D65 maps to white (1, 1, 1) by construction of the matrix (generate_spectral_tables.py:161-165).
    assert np.allclose((d.rgb_cmf * d.d65[:, None]).sum(0), 1.0, atol=1e-12)
    # A 500 nm monochromatic stimulus: its CMF colour is the table entry (tests/PERFORMANCE.md:473).
    xyz500 = d.cmf[140]
    assert np.allclose(cmf[140, :3], d.xyz_to_rgb @ xyz500, atol=2e-5)


UPLIFT = os.path.join(os.path.dirname(__file__), "golden", "spectral_uplift.npz")


def _rho(lag, phases):
    x = np.cos(phases)
    return np.arctan(lag[0] + 2 * lag[1] * x + 2 * lag[2] * (2 * x * x - 1)) / math.pi + 0.5


def test_uplift_matches_reference_generator(data):
    # Lagrange multipliers of the Cornell walls, the emission corners, both terrain palettes, grid
    # nodes, arbitrary codes and colours in refined cells, as the reference's generator and grid
    # interpolation give them: the oracle to double rounding, the library to float rounding.
    d, _ = data
    z = np.load(UPLIFT)
    assert z["refined"].sum() >= 3 and len(z["rgb"]) > 150
    for c, want in zip(z["rgb"], z["lagrange"]):
        scale = max(1.0, np.abs(want).max())
        got_o = d.reflectance_lagrange(c)
        assert got_o[3] == 1.0 and np.abs(got_o[:3] - want).max() < 1e-9 * scale, c
        got_l = vt.spectral_reflectance(tuple(float(v) for v in c)).astype(np.float64)
        assert got_l[3] == 1.0 and np.abs(got_l[:3] - want).max() < 1e-6 * scale, c
        assert np.abs(_rho(got_l, d.phases) - _rho(want, d.phases)).max() < 1e-6, c


def test_uplift_fixture_catches_plausible_mistakes(data):
    # Planted mistakes the fixture must catch at the library's bar (1e-6 of the largest multiplier):
    # interpolating the coarse grid in code space (the 8-bit axis) instead of linear light, and
    # solving each colour exactly instead of interpolating the grid of exact solves.
    d, _ = data
    z = np.load(UPLIFT)
    k = int(np.argmin(np.abs(z["rgb"] - np.float32([0.65, 0.05, 0.05])).sum(1)))
    c = z["rgb"][k].astype(np.float64)
    code85 = np.array([spectral.srgb_oetf(v) * 85.0 for v in c])
    cell = [min(int(v), 84) for v in code85]
    bar = 1e-6 * np.abs(z["lagrange"][k]).max()
    wrong = spectral.fourier_lagrange(d._interpolate(c, [3 * i for i in cell], 3, code85 - np.array(cell)))
    assert np.abs(wrong - z["lagrange"][k]).max() > 50 * bar
    exact, _ = d.solve_lagrange(c)
    assert np.abs(exact - z["lagrange"][k]).max() > 5 * bar


@pytest.mark.parametrize("rgb", [(0.65, 0.05, 0.05), (0.12, 0.45, 0.15), (0.2, 0.3, 0.9), (0.95, 0.9, 0.1),
                                 (0.02, 0.0, 0.85), (0.7, 0.0, 0.6)])
def test_reflectance_round_trip(data, rgb):
    # The reconstructed rho(lambda) integrated against D65 x CMF returns the colour within half an
    # 8-bit step (the reference measures at most 0.459 steps over every code, tests/PERFORMANCE.md:
    # 471-473), and stays within [0, 1].
    d, _ = data
    lag = vt.spectral_reflectance(rgb)
    assert lag[3] == 1.0
    rho = d.reflectance_spectrum(lag.astype(np.float64))
    assert rho.min() >= 0.0 and rho.max() <= 1.0
    back = d.weights @ rho
    steps = [abs(spectral.srgb_oetf(back[k]) - spectral.srgb_oetf(rgb[k])) * 255 for k in range(3)]
    assert max(steps) < 0.5, steps


def test_greys_are_flat():
    for g in (0.0, 0.18, 0.73, 1.0):
        lag = vt.spectral_reflectance((g, g, g))
        assert lag[3] == 2.0 and abs(lag[0] - g) < 1e-7
