"""CPU tests that pin the oracle's anisotropic specular lobe (oracle/openpbr_impl.h: OpenPBR's
specular_roughness_anisotropy, SurfaceSettings.surface.w, main.swift:2496; the alphas of
openpbr_compute_anisotropic_alpha along the hit's tangent frame, main.swift:2508-2509):

- on a white metal (Fresnel exactly 1, the multiple-scattering lobe isotropic), turning the tangent by
  90 degrees changes the BSDF by exactly the difference of two textbook anisotropic GGX reflections
  (Heitz 2014: D and the Smith G1 with alphas (a_x, a_y) and (a_y, a_x)), with the OpenPBR spec's
  mapping a_x = alpha sqrt(2 / (1 + (1 - anisotropy)^2)), a_y = (1 - anisotropy) a_x;
- anisotropy 0 ignores the tangent; the sampler agrees with the evaluator and the pdf integrates to the
  valid fraction; reflection stays reciprocal; a planted mistake (swapped alphas) is caught."""
import math

import numpy as np
import pytest

import oracle

N = np.array([0.0, 0.0, 1.0])


@pytest.fixture(autouse=True)
def _faithful():
    oracle.set_fault(0)
    yield
    oracle.set_fault(0)


def _dir(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def _aniso_ggx(wo, wi, ax, ay):
    """f |cos wi| of the textbook anisotropic GGX reflection with F = 1, in the (x, y, z) frame."""
    h = (wo + wi) / np.linalg.norm(wo + wi)
    d = 1 / (np.pi * ax * ay * (h[0] ** 2 / ax ** 2 + h[1] ** 2 / ay ** 2 + h[2] ** 2) ** 2)

    def g1(v):
        return 2 / (1 + math.sqrt(1 + (ax * ax * v[0] ** 2 + ay * ay * v[1] ** 2) / v[2] ** 2))

    return d * g1(wo) * g1(wi) / (4 * wo[2])


def _alphas(rough, aniso):
    a = rough * rough
    ax = a * math.sqrt(2 / (1 + (1 - aniso) ** 2))
    return max(ax, 1e-6), max((1 - aniso) * ax, 1e-6)


@pytest.mark.parametrize("rough,aniso", [(0.3, 0.5), (0.5, 0.8), (0.2, 0.3)])
def test_quarter_turn_of_the_tangent_is_the_textbook_difference(rough, aniso):
    along = oracle.BSDF(1, rough, rho=(1, 1, 1, 1), anisotropy=aniso, tangent=(1, 0, 0))
    across = oracle.BSDF(1, rough, rho=(1, 1, 1, 1), anisotropy=aniso, tangent=(0, 1, 0))
    ax, ay = _alphas(rough, aniso)
    rng = np.random.default_rng(int(rough * 10 + aniso * 100))
    n = 0
    for _ in range(80):
        wo = _dir(rng.uniform(0.1, 1.2), rng.uniform(0, 2 * np.pi))
        mirror = np.array([-wo[0], -wo[1], wo[2]])
        wi = mirror + rng.normal(0, 0.25, 3)
        wi /= np.linalg.norm(wi)
        if wi[2] < 0.05:
            continue
        d_lib = along.eval(N, wo, wi)[0][0] - across.eval(N, wo, wi)[0][0]
        # the second frame has tangent +y and bitangent n x y = -x: local (y, -x, z)
        rot = lambda v: np.array([v[1], -v[0], v[2]])
        want = _aniso_ggx(wo, wi, ax, ay) - _aniso_ggx(rot(wo), rot(wi), ax, ay)
        scale = _aniso_ggx(wo, wi, ax, ay) + _aniso_ggx(rot(wo), rot(wi), ax, ay)
        assert abs(d_lib - want) <= 2e-5 * scale + 1e-7, (d_lib, want)
        n += 1
    assert n > 50


def test_planted_alpha_swap_is_caught():
    oracle.set_fault(14)
    with pytest.raises(AssertionError):
        test_quarter_turn_of_the_tangent_is_the_textbook_difference(0.3, 0.5)


def test_zero_anisotropy_ignores_the_tangent():
    rng = np.random.default_rng(3)
    a = oracle.BSDF(4, 0.35, 1.5, 0.2, 0.0, rho=(0.7, 0.4, 0.2, 0.3), tangent=(1, 0, 0))
    b = oracle.BSDF(4, 0.35, 1.5, 0.2, 0.0, rho=(0.7, 0.4, 0.2, 0.3), tangent=(0.6, 0.8, 0.0))
    for _ in range(50):
        wo, wi = _dir(rng.uniform(0, 1.4), rng.uniform(0, 6.3)), _dir(rng.uniform(0, 3.1), rng.uniform(0, 6.3))
        fa, pa = a.eval(N, wo, wi)
        fb, pb = b.eval(N, wo, wi)
        assert fa.tolist() == fb.tolist() and pa == pb


def _tangent(rng):
    t = rng.normal(size=3)
    t[2] *= 0.2
    return tuple(t / np.linalg.norm(t))


ANISO = [dict(type=1, roughness=0.35, ior=1.5, metalness=1.0, transmission=0.0, rho=(0.95, 0.64, 0.54, 0.5),
              anisotropy=0.7),
         dict(type=4, roughness=0.45, ior=1.5, metalness=0.0, transmission=0.0, rho=(0.2, 0.5, 0.8, 0.4),
              anisotropy=0.5, coat=0.6),
         dict(type=4, roughness=0.4, ior=1.5, metalness=0.0, transmission=1.0, rho=(1, 1, 1, 1), anisotropy=0.6)]


def _bsdf(m, tangent):
    return oracle.BSDF(m["type"], m["roughness"], m["ior"], m["metalness"], m["transmission"], m["rho"],
                       coat=m.get("coat", 0.0), anisotropy=m["anisotropy"], tangent=tangent)


def _sphere(n, rng):
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], 1)


@pytest.mark.parametrize("mi", range(len(ANISO)))
def test_anisotropic_sampler_matches_evaluator_and_pdf_normalises(mi):
    rng = np.random.default_rng(70 + mi)
    b = _bsdf(ANISO[mi], _tangent(rng))
    n = 0
    for cos_t, inside in ((0.85, False), (0.4, False), (0.6, True)):
        wo = _dir(math.acos(cos_t), 0.7)
        for _ in range(200):
            r = b.sample(N, wo, rng.random(3), inside)
            if r is None:
                continue
            d, w, pdf = r
            f, p = b.eval(N, wo, d, inside)
            assert abs(p / pdf - 1) < 1e-4
            assert np.allclose(w, f / pdf, rtol=1e-4, atol=1e-7)
            n += 1
    assert n > 450
    if mi == 0:
        # Uniform-sphere quadrature is only tight for a reflective lobe this wide (a refracted lobe or
        # the coat's narrow peak needs importance sampling; the check above covers them).
        wo = _dir(0.8, 0.3)
        dirs = _sphere(60000, rng)
        pdf_int = 4 * np.pi * np.mean([b.eval(N, wo, d)[1] for d in dirs])
        valid = np.mean([b.sample(N, wo, rng.random(3)) is not None for _ in range(20000)])
        assert abs(pdf_int - valid) < 0.03, (pdf_int, valid)


def test_anisotropic_reflection_is_reciprocal():
    rng = np.random.default_rng(5)
    b = _bsdf(ANISO[0], _tangent(rng))
    for _ in range(200):
        a, c = _sphere(2, rng)
        a[2], c[2] = abs(a[2]) + 0.05, abs(c[2]) + 0.05
        a /= np.linalg.norm(a)
        c /= np.linalg.norm(c)
        f1, _ = b.eval(N, a, c)
        f2, _ = b.eval(N, c, a)
        assert np.allclose(f1 / c[2], f2 / a[2], rtol=2e-4, atol=1e-7)
