"""CPU tests that pin the oracle's fuzz layer (oracle/openpbr_impl.h: OpenPBR's fuzz lobe, the Disney
sheen of Zeltner, Burley and Chiang 2022 as linearly transformed cosines, weight from the material,
SurfaceSettings.detail.x, main.swift:2497, roughness 0.5 and white by OpenPBR's defaults):

- its LTC table is the vendored data (the committed fixture against the reference's header);
- the sheen it adds is the textbook LTC density (Heitz et al. 2016: D_o(M^-1 w / |M^-1 w|) |det M^-1| /
  |M^-1 w|^3 with M^-1 = [[a, 0, b], [0, a, 0], [0, 0, 1]] in the frame of the view's azimuth) times the
  table's reflectance R, over the layer below scaled by 1 - R (presence 1), with the coefficients
  interpolated here from the fixture;
- the sampler agrees with the evaluator, the pdf integrates to the valid fraction, a fuzzy white
  surface keeps its energy, inside hits ignore the fuzz; a planted mistake (the frame turned the wrong
  way) is caught."""
import math
import os
import re

import numpy as np
import pytest

import oracle

N = np.array([0.0, 0.0, 1.0])
REFERENCE = "/root/reference"


@pytest.fixture(autouse=True)
def _faithful():
    oracle.set_fault(0)
    yield
    oracle.set_fault(0)


def _dir(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def _ltc_coeffs(cos_t):
    t = oracle.reference_energy_tables()["ltc_sheen"].astype(np.float64)
    row, col = 0.5 * 31, min(max(cos_t, 0.0), 1 - 2 ** -24) * 31
    r, c = int(row), int(col)
    rf, cf = row - r, col - c
    lo = t[r, c] * (1 - cf) + t[r, c + 1] * cf
    hi = t[r + 1, c] * (1 - cf) + t[r + 1, c + 1] * cf
    return lo * (1 - rf) + hi * rf


def _ltc_density(wo, wi):
    """Textbook LTC density in the frame whose x axis is the view's azimuth."""
    a, b, _ = _ltc_coeffs(wo[2])
    phi = math.atan2(wo[1], wo[0])
    c, s = math.cos(phi), math.sin(phi)
    w = np.array([c * wi[0] + s * wi[1], -s * wi[0] + c * wi[1], wi[2]])
    minv = np.array([[a, 0, b], [0, a, 0], [0, 0, 1.0]])
    o = minv @ w
    length = np.linalg.norm(o)
    d_o = max(o[2] / length, 0.0) / np.pi
    return d_o * abs(np.linalg.det(minv)) / length ** 3


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "Vendor", "OpenPBR")), reason="reference checkout absent")
def test_ltc_fixture_is_the_vendored_table():
    text = open(os.path.join(REFERENCE, "Vendor", "OpenPBR", "impl", "data", "openpbr_ltc_data.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    vals = [float(v) for t in re.findall(r"vec3\(([^)]*)\)", re.sub(r"//[^\n]*", "", text)) for v in t.split(",")]
    assert np.array_equal(np.array(vals, np.float32).reshape(32, 32, 3), oracle.reference_energy_tables()["ltc_sheen"])


@pytest.mark.parametrize("cos_o", [0.9, 0.6, 0.3])
def test_sheen_is_the_textbook_ltc_over_the_scaled_base(cos_o):
    rho = (0.6, 0.3, 0.2, 0.4)
    fuzzy = oracle.BSDF(4, 0.4, 1.5, 0.0, 0.0, rho=rho, fuzz=1.0)
    bare = oracle.BSDF(4, 0.4, 1.5, 0.0, 0.0, rho=rho)
    rng = np.random.default_rng(int(cos_o * 10))
    wo = _dir(math.acos(cos_o), rng.uniform(0, 2 * np.pi))
    r_view = _ltc_coeffs(cos_o)[2]
    for _ in range(60):
        wi = _dir(rng.uniform(0, 1.5), rng.uniform(0, 2 * np.pi))
        f1, _ = fuzzy.eval(N, wo, wi)
        f0, _ = bare.eval(N, wo, wi)
        sheen = f1 - (1 - r_view) * f0
        want = _ltc_density(wo, wi) * r_view
        assert np.allclose(sheen, want, rtol=2e-4, atol=2e-6), (sheen, want)


def test_planted_frame_turn_is_caught():
    oracle.set_fault(15)
    with pytest.raises(AssertionError):
        test_sheen_is_the_textbook_ltc_over_the_scaled_base(0.6)


FUZZY = [dict(type=4, roughness=0.4, ior=1.5, metalness=0.0, transmission=0.0, rho=(0.6, 0.3, 0.2, 0.4), fuzz=1.0),
         dict(type=4, roughness=0.3, ior=1.5, metalness=0.3, transmission=0.0, rho=(0.2, 0.5, 0.8, 0.4), fuzz=0.5,
              coat=0.7),
         dict(type=1, roughness=0.5, ior=1.5, metalness=1.0, transmission=0.0, rho=(0.9, 0.8, 0.6, 0.7), fuzz=0.8)]


def _bsdf(m, fuzz=None):
    return oracle.BSDF(m["type"], m["roughness"], m["ior"], m["metalness"], m["transmission"], m["rho"],
                       coat=m.get("coat", 0.0), fuzz=m["fuzz"] if fuzz is None else fuzz)


def _sphere(n, rng):
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)


@pytest.mark.parametrize("mi", range(len(FUZZY)))
def test_fuzzy_sampler_matches_evaluator(mi):
    b = _bsdf(FUZZY[mi])
    rng = np.random.default_rng(90 + mi)
    n = 0
    for cos_t in (0.85, 0.5, 0.15):
        wo = _dir(math.acos(cos_t), 1.1)
        for _ in range(200):
            r = b.sample(N, wo, rng.random(3))
            if r is None:
                continue
            d, w, pdf = r
            f, p = b.eval(N, wo, d)
            assert abs(p / pdf - 1) < 1e-4
            assert np.allclose(w, f / pdf, rtol=1e-4, atol=1e-7)
            n += 1
    assert n > 450


@pytest.mark.parametrize("mi", [0, 2])
def test_fuzzy_pdf_integrates_to_the_valid_fraction(mi):
    b = _bsdf(FUZZY[mi])
    rng = np.random.default_rng(110 + mi)
    wo = _dir(0.9, 0.4)
    pdf_int = 4 * np.pi * np.mean([b.eval(N, wo, d)[1] for d in _sphere(60000, rng)])
    valid = np.mean([b.sample(N, wo, rng.random(3)) is not None for _ in range(20000)])
    assert abs(pdf_int - valid) < 0.03, (pdf_int, valid)


def test_fuzzy_white_surface_keeps_its_energy():
    # Sheen reflectance R plus (1 - R) times a white base's albedo (about 1 with the compensated lobes):
    # about 1, never above.
    b = oracle.BSDF(4, 0.5, 1.5, 0.0, 0.0, rho=(1, 1, 1, 1), fuzz=1.0)
    rng = np.random.default_rng(13)
    for cos_t in (0.9, 0.4):
        w = [r[1][0] if (r := b.sample(N, _dir(math.acos(cos_t), 0.2), rng.random(3))) else 0.0 for _ in range(12000)]
        assert 0.9 < np.mean(w) < 1.03, (cos_t, np.mean(w))


def test_inside_hits_ignore_the_fuzz():
    m = dict(type=4, roughness=0.4, ior=1.5, metalness=0.0, transmission=1.0, rho=(1, 1, 1, 1), fuzz=1.0)
    fuzzy, bare = _bsdf(m), _bsdf(m, fuzz=0.0)
    rng = np.random.default_rng(4)
    for _ in range(50):
        wo, wi = _sphere(2, rng)
        wo[2] = abs(wo[2]) + 0.1
        wo /= np.linalg.norm(wo)
        f1, p1 = fuzzy.eval(N, wo, wi, inside=True)
        f0, p0 = bare.eval(N, wo, wi, inside=True)
        assert f1.tolist() == f0.tolist() and p1 == p0
