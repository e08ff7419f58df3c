"""CPU tests that pin the oracle's layered-BSDF subset (oracle/openpbr_impl.h) and its energy tables
(oracle/energy.c): the tables against the reference's own data (OpenPBR's vendored dielectric arrays,
the reference's MetalEnergy.metal), then by what the mathematics fixes: the GGX single-scattering
albedo by independent quadrature, Fresnel at zero roughness, the black dielectric and the smooth metal
as textbook GGX reflections (value and pdf in closed form), sampler-evaluator consistency, pdf
normalisation, reciprocity, white-furnace energy conservation of metal, glass and plastic with
multiple-scattering compensation, and a planted mistake those checks must catch."""
import math
import os
import sys

import numpy as np
import pytest

import oracle

N = np.array([0.0, 0.0, 1.0])


def _wo(cos_t, phi=0.3):
    s = math.sqrt(1 - cos_t * cos_t)
    return np.array([s * math.cos(phi), s * math.sin(phi), cos_t])


@pytest.fixture(autouse=True)
def _faithful():
    oracle.set_fault(0)
    yield
    oracle.set_fault(0)


def _ggx(hz, alpha):
    return 1.0 / (math.pi * alpha * alpha * ((1 - hz * hz) / (alpha * alpha) + hz * hz) ** 2)


def _g1(z, alpha):
    z2 = z * z
    return 2.0 / (1.0 + np.sqrt(1.0 + alpha * alpha * (1 - z2) / z2))


def _metal_albedo_quadrature(alpha, mu, nt=4000, nphi=1000):
    # E(mu) = int D(h) G1(v) G1(l) / (4 mu) dl over the upper hemisphere, midpoint rule in (cos, phi).
    c = (np.arange(nt) + 0.5) / nt
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    C, P = np.meshgrid(c, phi, indexing="ij")
    S = np.sqrt(1 - C * C)
    l = np.stack([S * np.cos(P), S * np.sin(P), C], -1)
    v = np.array([math.sqrt(1 - mu * mu), 0.0, mu])
    h = l + v
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    f = _ggx(h[..., 2], alpha) * _g1(mu, alpha) * _g1(C, alpha) / (4 * mu)
    return float(f.mean() * 2 * np.pi)


REFERENCE = "/root/reference"


def test_energy_tables_are_the_reference_data():
    # The oracle's dielectric tables are OpenPBR's unorm16 arrays value for value, and its metal bake
    # (the recipe of scripts/generate_metal_energy.py, float sums) lands within float rounding of the
    # table the reference committed (Shaders/MetalEnergy.metal).
    table, avg, luts, _ = oracle.energy_tables()
    ref = oracle.reference_energy_tables()
    pairs = zip(("dielectric", "dielectric_avg", "dielectric_ratio", "opaque", "opaque_avg"), oracle.DIELECTRIC_KEYS)
    for nm, key in pairs:
        assert np.array_equal(np.rint(luts[nm] * 65535).astype(np.uint16), ref[key]), nm
    assert np.abs(table - ref["metal_complement"]).max() < 5e-6
    assert np.abs(avg - ref["metal_avg_complement"]).max() < 5e-6
    # Spot values read off the vendored files (reflection ratio row 0, opaque average node (31, 31)).
    assert ref["reflection_ratio"][0, 0] == 10108 and ref["reflection_ratio"][0, 1] == 39576
    assert ref["metal_complement"].shape == (129, 129) and np.all(ref["metal_complement"][0] == 0)


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "Vendor", "OpenPBR")), reason="reference checkout absent")
def test_energy_fixture_equals_vendored_arrays():
    # The committed fixture is the reference's data: re-read the vendored headers and compare.
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_energy_golden as g
    ref = oracle.reference_energy_tables()
    data = os.path.join(REFERENCE, "Vendor", "OpenPBR", "impl", "data")
    import re
    for key, fname, shape in g.DIELECTRIC:
        vals = [int(v) for v in re.findall(r"\d+", g.strip_comments(open(os.path.join(data, fname)).read()))]
        assert np.array_equal(np.array(vals, np.uint16).reshape(shape), ref[key]), key
    m = g.metal_arrays(os.path.join(REFERENCE, "Shaders", "MetalEnergy.metal"))
    assert np.array_equal(m["vibeMetalEnergy"].reshape(129, 129), ref["metal_complement"])
    assert np.array_equal(m["vibeMetalAverageEnergy"], ref["metal_avg_complement"])


def test_metal_table_matches_independent_quadrature():
    table, avg, _, _ = oracle.energy_tables()
    for i in (40, 64, 96, 128):
        r = i / 128
        for j in (40, 80, 128):
            mu = max(1e-5, (j / 128) ** 2)
            want = 1 - _metal_albedo_quadrature(r * r, mu)
            assert abs(table[i, j] - want) < 1e-3, (i, j, table[i, j], want)
    # The average integrates the sqrt-cosine-linear table exactly: compare with a fine quadrature of it.
    s = np.linspace(0, 1, 20001)
    for i in (32, 96):
        e = np.interp(s, np.arange(129) / 128, table[i])
        assert abs(avg[i] - np.trapezoid(4 * e * s ** 3, s)) < 1e-6
    assert np.all(table[0] == 0) and np.all(np.diff(avg[8:]) > 0)


def _fresnel(eta, c):
    s2t = (1 - c * c) / (eta * eta)
    if s2t >= 1:
        return 1.0
    ct = math.sqrt(1 - s2t)
    rs = (c - eta * ct) / (c + eta * ct)
    rp = (ct - eta * c) / (ct + eta * c)
    return 0.5 * (rs * rs + rp * rp)


def test_dielectric_tables_reduce_to_fresnel_when_smooth():
    # At roughness 0 the reflection is Fresnel's and nothing is lost (a property of OpenPBR's data).
    _, _, luts, iors = oracle.energy_tables()
    for i in (0, 8, 20, 31):
        for k in (3, 16, 31):
            mu = max(1e-5, k / 31)
            assert abs(luts["opaque"][i, 0, k] - (1 - _fresnel(iors[i], mu))) < 3e-5, (i, k)
            assert luts["dielectric"][i, 0, k] < 2e-5
    assert abs(_fresnel(1.5, 1.0) - 0.04) < 1e-12             # the closed form itself
    assert _fresnel(1 / 1.5, 0.5) == 1.0                        # total internal reflection


@pytest.mark.parametrize("rough", [0.1, 0.3, 0.6, 1.0])
@pytest.mark.parametrize("cos_t", [0.9, 0.5, 0.2])
def test_white_metal_furnace(rough, cos_t):
    # F = 1 (white F82 metal): single scattering plus the metal MMS lobe returns all the energy.
    b = oracle.BSDF(1, rough, rho=(1, 1, 1, 1))
    rng = np.random.default_rng(7)
    wo = _wo(cos_t)
    w = [r[1][0] if (r := b.sample(N, wo, rng.random(3))) else 0.0 for _ in range(6000)]
    assert abs(np.mean(w) - 1) < 0.03, np.mean(w)


def _sphere(n, rng):
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)


MATERIALS = [dict(type=4, roughness=0.3, ior=1.5, metalness=0.0, transmission=0.0, rho=(0.8, 0.5, 0.3, 0.2)),
             dict(type=4, roughness=0.7, ior=1.4, metalness=0.0, transmission=0.0, rho=(0.9, 0.9, 0.9, 0.9)),
             dict(type=1, roughness=0.25, ior=1.5, metalness=1.0, transmission=0.0, rho=(0.95, 0.64, 0.54, 0.5)),
             dict(type=4, roughness=0.45, ior=1.5, metalness=0.0, transmission=1.0, rho=(1, 1, 1, 1)),
             dict(type=4, roughness=0.5, ior=1.33, metalness=0.3, transmission=0.5, rho=(0.7, 0.6, 0.5, 0.4))]


def _bsdf(m):
    return oracle.BSDF(m["type"], m["roughness"], m["ior"], m["metalness"], m["transmission"], m["rho"])


@pytest.mark.parametrize("mi", range(len(MATERIALS)))
def test_sampler_matches_evaluator(mi):
    b = _bsdf(MATERIALS[mi])
    rng = np.random.default_rng(3 + mi)
    for cos_t, inside in ((0.8, False), (0.35, False), (0.6, True)):
        wo = _wo(cos_t)
        for _ in range(200):
            r = b.sample(N, wo, rng.random(3), inside)
            if r is None:
                continue
            d, w, pdf = r
            f, p = b.eval(N, wo, d, inside)
            assert abs(p / pdf - 1) < 1e-4
            assert np.allclose(w, f / pdf, rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize("mi", range(len(MATERIALS)))
def test_pdf_and_albedo_integrals_agree_with_sampling(mi):
    # int pdf dw = P(sample valid) and int f|cos| dw = E[weight]: uniform-sphere quadrature of the
    # evaluator against the sampler.
    b = _bsdf(MATERIALS[mi])
    rng = np.random.default_rng(11 + mi)
    wo = _wo(0.7)
    dirs = _sphere(60000, rng)
    ev = [b.eval(N, wo, d) for d in dirs]
    pdf_int = 4 * np.pi * np.mean([e[1] for e in ev])
    alb_int = 4 * np.pi * np.mean([e[0] for e in ev], axis=0)
    samples = [b.sample(N, wo, rng.random(3)) for _ in range(30000)]
    valid = np.mean([s is not None for s in samples])
    alb = np.mean([s[1] if s else np.zeros(4) for s in samples], axis=0)
    assert abs(pdf_int - valid) < 0.03, (pdf_int, valid)
    assert np.allclose(alb_int, alb, rtol=0.04, atol=0.01), (alb_int, alb)
    assert np.all(alb < 1.03)


@pytest.mark.parametrize("mi", [0, 1, 2])
def test_reflection_is_reciprocal(mi):
    b = _bsdf(MATERIALS[mi])
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, c = _sphere(2, rng)
        a[2], c[2] = abs(a[2]) + 0.05, abs(c[2]) + 0.05
        a /= np.linalg.norm(a)
        c /= np.linalg.norm(c)
        f1, _ = b.eval(N, a, c)
        f2, _ = b.eval(N, c, a)
        assert np.allclose(f1 / c[2], f2 / a[2], rtol=2e-4, atol=1e-7)


def test_glass_and_plastic_furnace():
    # A white rough glass interface with MMS compensation keeps about all the energy (reflection
    # plus transmission); a white rough plastic (GGX coat over the compensated diffuse) too.
    rng = np.random.default_rng(9)
    for m in (dict(type=4, roughness=0.5, ior=1.5, metalness=0.0, transmission=1.0, rho=(1, 1, 1, 1)),
              dict(type=4, roughness=0.5, ior=1.5, metalness=0.0, transmission=0.0, rho=(1, 1, 1, 1))):
        b = _bsdf(m)
        for cos_t in (0.9, 0.4):
            w = [r[1][0] if (r := b.sample(N, _wo(cos_t), rng.random(3))) else 0.0 for _ in range(8000)]
            assert abs(np.mean(w) - 1) < 0.05, (m, cos_t, np.mean(w))


def test_planted_shadowing_fault_is_caught():
    b = _bsdf(MATERIALS[0])
    rng = np.random.default_rng(2)
    wo = _wo(0.5)
    oracle.set_fault(5)
    bad = 0
    for _ in range(100):
        r = b.sample(N, wo, rng.random(3))
        if r is None:
            continue
        f, p = b.eval(N, wo, r[0])
        bad += not np.allclose(r[1], f / p, rtol=1e-4, atol=1e-7)
    assert bad > 10


def _ggx_reflection(wo, wi, alpha, fresnel):
    """Textbook GGX reflection (Walter et al. 2007): f |cos wi| = F D G1(wo) G1(wi) / (4 cos wo), and the
    visible-normal density D G1(wo) / (4 cos wo) (Heitz 2018)."""
    h = (wo + wi) / np.linalg.norm(wo + wi)
    d = _ggx(h[2], alpha)
    g = _g1(wo[2], alpha) * _g1(wi[2], alpha)
    return fresnel(float(wo @ h)) * d * g / (4 * wo[2]), d * _g1(wo[2], alpha) / (4 * wo[2])


@pytest.mark.parametrize("rough,ior", [(0.3, 1.5), (0.6, 1.33), (0.2, 2.0)])
def test_black_dielectric_is_textbook_ggx_reflection(rough, ior):
    # A black opaque dielectric has no diffuse, transmission or multiple-scattering term: its BSDF is
    # the GGX reflection with dielectric Fresnel, and the aggregate pdf that lobe's visible-normal pdf.
    b = oracle.BSDF(4, rough, ior, 0.0, 0.0, rho=(0, 0, 0, 0))
    rng = np.random.default_rng(int(rough * 100 + ior * 10))
    for _ in range(60):
        wo = _wo(rng.uniform(0.15, 0.95), rng.uniform(0, 6.2))
        wi = _wo(rng.uniform(0.1, 0.99), rng.uniform(0, 6.2))
        f, pdf = b.eval(N, wo, wi)
        want_f, want_pdf = _ggx_reflection(wo, wi, rough * rough, lambda c: _fresnel(ior, c))
        assert np.allclose(f, want_f, rtol=1e-4, atol=1e-8), (f, want_f)
        assert abs(pdf - want_pdf) <= 1e-4 * want_pdf + 1e-8, (pdf, want_pdf)


def test_smooth_metal_is_textbook_ggx_with_schlick():
    # Below roughness 0.04 the metal's multiple-scattering lobe is off: the BSDF is GGX with the F82
    # Fresnel, which for a white tint is Schlick's f0 + (1 - f0) (1 - c)^5.
    rho = np.array([0.95, 0.64, 0.54, 0.5])
    rough = 0.038
    b = oracle.BSDF(1, rough, rho=tuple(rho))
    rng = np.random.default_rng(17)
    n = 0
    for _ in range(80):
        wo = _wo(rng.uniform(0.2, 0.95), rng.uniform(0, 6.2))
        wi = np.array([-wo[0], -wo[1], wo[2]]) + rng.normal(0, 0.01, 3)
        wi /= np.linalg.norm(wi)
        f, pdf = b.eval(N, wo, wi)
        for k in range(4):
            want, want_pdf = _ggx_reflection(wo, wi, rough * rough,
                                             lambda c: rho[k] + (1 - rho[k]) * (1 - min(max(c, 0), 1)) ** 5)
            assert abs(f[k] - want) <= 2e-4 * want + 1e-8, (k, f[k], want)
        assert abs(pdf - want_pdf) <= 2e-4 * want_pdf + 1e-8
        n += 1
    assert n == 80
