"""Oracle (test infrastructure only): numpy restatement of the reference's spectral table setup.

Follows scripts/generate_spectral_tables.py (the reference's offline generator):
  Data.__init__            :145-170  D65 normalised to Y = 1, xyz->linear sRGB from the sRGB primaries
                                     and D65's own white, reflectance->rgb weights, warped phases
  Data.phase               :172-177  Peters' even warp, piecewise linear in 5 nm steps
  sampling_density         :490-498  0.9 S s / int(S s) + 0.1 s / int(s), s = |r| + |g| + |b|
  inverse_cdf              :507-530  exact inversion of the piecewise-linear density, 1025 nodes
  rgb_and_jacobian,
  solve_lagrange           :276-340  Levenberg-Marquardt in Lagrange-multiplier space
  moments_of_lagrange      :227-239  bounded moments by the 1024-node midpoint rule
  solve_code               :381-386  exact moments of an 8-bit code (the coarse grid's nodes)
  code_error               :357-366  8-bit round trip of a code through the interpolated moments
and main.swift's conversion of a colour to a reflectance: spectral_moments_coarse /
spectral_moments :855-884 (trilinear interpolation in linear light on the 86^3 grid of exact solves,
refined cells at one-code spacing), the refinement rule of spectral_refine_flags_kernel :8223-8233
(a cell is refined when one of its codes comes back more than 0.35 steps off; node codes counted in
the cell above them), SpectralColour.fourierLagrange :8640-8680 (bounded-MESE multipliers with
Levinson's recursion and biasing), spectral_lagrange / spectral_from_lagrange :886-903 (greys flat,
reflectance atan(L0 + 2 L1 cos p + 2 L2 cos 2p) / pi + 1/2).  Input data: the CIE tables and
Peters' warp as vendored by the reference (numbers in metal-vibe-tracer_amd/csrc/spectral_data.inc).
Pinned by tests/golden/spectral_uplift.npz, made with the reference's own generator.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "metal-vibe-tracer_amd", "csrc", "spectral_data.inc")
LAMBDA_MIN, LAMBDA_MAX, SAMPLES, SEGMENTS = 360, 830, 471, 1024
EPSILON = 1e-4


def _arrays():
    text = open(DATA).read()
    out = {}
    for name in ("kCieXyz", "kD65", "kPhaseWarp"):
        m = re.search(name + r"\[[^=]*=\s*\{(.*?)\};", text, re.S)
        body = re.sub(r"//[^\n]*", "", m.group(1))
        out[name] = np.array([float(v) for v in re.findall(r"[-+0-9.eE]+", body)])
    return out["kCieXyz"].reshape(SAMPLES, 3), out["kD65"], out["kPhaseWarp"]


def srgb_eotf(v):
    return v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4


def srgb_oetf(v):
    v = min(max(float(v), 0.0), 1.0)
    return 12.92 * v if v <= 0.0031308 else 1.055 * v ** (1.0 / 2.4) - 0.055


def moments_of_lagrange(lam, nodes=1024):
    """c_j = (1/pi) int_{-pi}^{0} g(phi) cos(j phi) dphi, midpoint rule with `nodes` nodes."""
    h = math.pi / nodes
    cp = np.cos(-math.pi + (np.arange(nodes) + 0.5) * h)
    c2p = 2.0 * cp * cp - 1.0
    v = np.arctan(lam[0] + 2.0 * lam[1] * cp + 2.0 * lam[2] * c2p) / math.pi + 0.5
    return np.array([v.sum(), (v * cp).sum(), (v * c2p).sum()]) * h / math.pi


def fourier_lagrange(c):
    """Bounded-MESE Lagrange multipliers of moments c (Peters et al. 2019, Eqs. 6, 7, 10), Levinson's
    recursion with the biasing that pulls invalid moments back (keep 0.9999, then 0)."""
    pi = math.pi
    c0 = min(max(float(c[0]), EPSILON), 1.0 - EPSILON)
    g0p = complex(math.sin(pi * c0), -math.cos(pi * c0)) / (4.0 * pi)
    g0 = 2.0 * g0p.real
    g1 = 2j * pi * c[1] * g0p
    g2 = 1j * pi * (2.0 * c[2] * g0p + c[1] * g1)
    q0 = 1.0 / g0
    u1 = q0 * g1
    n1, keep = abs(u1) ** 2, 0.9999
    if n1 >= 1.0:
        u1 *= keep / math.sqrt(n1)
        g1 = u1 / q0
        n1, keep = keep * keep, 0.0
    d1 = 1.0 / (1.0 - n1)
    a0, a1 = q0 * d1, -u1 * q0 * d1
    u2 = a0 * g2 + a1 * g1
    n2 = abs(u2) ** 2
    if n2 >= 1.0:
        u2 *= keep / math.sqrt(n2)
        g2 = (u2 - a1 * g1) / a0
        n2 = keep * keep
    d2 = 1.0 / (1.0 - n2)
    b = (a0 * d2, (a1 - u2 * a1.conjugate()) * d2, -u2 * a0 * d2)
    gamma = (g0p, g1, g2)
    out = []
    for l in range(3):
        acc = 0j
        for k in range(3 - l):
            acc += gamma[k] * sum(b[j + k + l].conjugate() * b[j] for j in range(3 - k - l))
        out.append(2.0 * acc.imag / b[0].real)
    return np.array(out)


class Data:
    def __init__(self):
        self.cmf, d65, self.warp = _arrays()
        y = (d65 * self.cmf[:, 1]).sum()
        self.d65 = d65 / y
        white = (self.d65[:, None] * self.cmf).sum(0)
        prim = np.array([[x / yy for x, yy in ((0.64, 0.33), (0.30, 0.60), (0.15, 0.06))], [1.0, 1.0, 1.0],
                         [(1 - x - yy) / yy for x, yy in ((0.64, 0.33), (0.30, 0.60), (0.15, 0.06))]])
        scale = np.linalg.solve(prim, white)
        self.rgb_to_xyz = prim * scale[None, :]
        self.xyz_to_rgb = np.linalg.inv(self.rgb_to_xyz)
        self.rgb_cmf = self.cmf @ self.xyz_to_rgb.T            # (471, 3)
        self.weights = (self.rgb_cmf * self.d65[:, None]).T   # (3, 471)
        lam = np.arange(LAMBDA_MIN, LAMBDA_MAX + 1, dtype=np.float64)
        self.phases = np.array([self.phase(l) for l in lam])
        self.cos1 = 2.0 * np.cos(self.phases)
        self.cos2 = 2.0 * np.cos(2.0 * self.phases)

    def phase(self, lam):
        t = (lam - 360.0) * 0.2
        i = min(max(int(math.floor(t)), 0), 93)
        f = t - i
        return self.warp[i] * (1.0 - f) + self.warp[i + 1] * f

    # -------------------------------------------------------------------- wavelength sampling
    def sampling_density(self):
        s = np.abs(self.rgb_cmf).sum(1)

        def normalized(v):
            return v / (0.5 * (v[:-1] + v[1:])).sum()
        return 0.9 * normalized(self.d65 * s) + 0.1 * normalized(s)

    def inverse_cdf(self):
        density = self.sampling_density()
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (density[:-1] + density[1:]))])
        total = cdf[-1]
        out = [float(LAMBDA_MIN)]
        j = 0
        for n in range(1, SEGMENTS):
            y = total * n / SEGMENTS
            while cdf[j + 1] < y:
                j += 1
            p0, p1, rest = density[j], density[j + 1], y - cdf[j]
            dp = p1 - p0
            t = rest / p0 if abs(dp) < 1e-12 * max(p0, 1e-30) else (-p0 + math.sqrt(max(p0 * p0 + 2 * dp * rest, 0.0))) / dp
            out.append(LAMBDA_MIN + j + min(max(t, 0.0), 1.0))
        out.append(float(LAMBDA_MAX))
        return np.array(out), density

    # -------------------------------------------------------------------- bounded MESE
    def rgb_and_jacobian(self, lam):
        s = lam[0] + lam[1] * self.cos1 + lam[2] * self.cos2
        v = np.arctan(s) / math.pi + 0.5
        dv = 1.0 / (math.pi * (1.0 + s * s))
        rgb = self.weights @ v
        basis = np.stack([np.ones_like(s), self.cos1, self.cos2])     # (3, 471)
        jac = (self.weights * dv[None, :]) @ basis.T                   # (3 rgb, 3 multipliers)
        return rgb, jac

    def solve_lagrange(self, target, tolerance=1e-10, iterations=80):
        target = np.clip(np.asarray(target, np.float64), EPSILON, 1.0 - EPSILON)
        lam = np.zeros(3)
        rgb, jac = self.rgb_and_jacobian(lam)
        f = target - rgb
        err = np.abs(f).max()
        mu = 0.0
        n = 0
        while err > tolerance and n < iterations:
            n += 1
            normal = jac.T @ jac
            gradient = jac.T @ f
            while True:
                damped = normal.copy()
                damped[np.diag_indices(3)] *= 1.0 + mu
                step = np.linalg.solve(damped, gradient)
                trial = lam + step
                rgb_t, jac_t = self.rgb_and_jacobian(trial)
                f_t = target - rgb_t
                err_t = np.abs(f_t).max()
                if err_t < err:
                    mu = mu * 0.1 if mu > 1e-12 else 0.0
                    break
                mu = max(mu * 10.0, 1e-6)
                if mu > 1e8:
                    return lam, err
            lam, jac, f, err = trial, jac_t, f_t, err_t
        return lam, err

    # -------------------------------------------------------------------- uplift grid
    def code_moments(self, code):
        """Exact moments of an 8-bit code, float32 as the grid file stores them."""
        cache = self.__dict__.setdefault("_exact", {})
        code = tuple(int(v) for v in code)
        if code not in cache:
            lam, _ = self.solve_lagrange([srgb_eotf(v / 255.0) for v in code])
            cache[code] = moments_of_lagrange(lam).astype(np.float32)
        return cache[code]

    def _interpolate(self, c, base, step, t):
        m = np.zeros(3)
        for a in (0, 1):
            for b in (0, 1):
                for d in (0, 1):
                    w = (t[0] if a else 1 - t[0]) * (t[1] if b else 1 - t[1]) * (t[2] if d else 1 - t[2])
                    if w == 0.0:
                        continue
                    code = [min(base[k] + (a, b, d)[k] * step, 255) for k in range(3)]
                    m += w * self.code_moments(code).astype(np.float64)
        return m

    def coarse_moments(self, c, cell):
        nodes = np.float32([srgb_eotf(3.0 * i / 255.0) for i in range(86)]).astype(np.float64)
        lo, hi = nodes[list(cell)], nodes[[i + 1 for i in cell]]
        t = np.clip((np.asarray(c, np.float64) - lo) / (hi - lo), 0.0, 1.0)
        return self._interpolate(c, [3 * i for i in cell], 3, t)

    def refined(self, cell):
        """A cell whose codes come back from the coarse conversion more than 0.35 8-bit steps off."""
        flags = self.__dict__.setdefault("_refined", {})
        cell = tuple(int(i) for i in cell)
        if cell not in flags:
            ranges = [range(3 * i, 3 * i + (4 if i == 84 else 3)) for i in cell]
            worst = 0.0
            for r in ranges[0]:
                for g in ranges[1]:
                    for b in ranges[2]:
                        if r == g == b:
                            continue
                        c = [srgb_eotf(v / 255.0) for v in (r, g, b)]
                        lam = fourier_lagrange(self.coarse_moments(c, cell))
                        rgb = self.weights @ (np.arctan(lam[0] + lam[1] * self.cos1 + lam[2] * self.cos2) / math.pi + 0.5)
                        worst = max(worst, max(abs(srgb_oetf(rgb[k]) * 255.0 - (r, g, b)[k]) for k in range(3)))
            flags[cell] = worst > 0.35
        return flags[cell]

    def uplift_moments(self, c):
        c = np.asarray(c, np.float64)
        code85 = np.array([srgb_oetf(v) * 85.0 for v in c])
        cell = [min(max(int(v), 0), 84) for v in code85]
        if not self.refined(cell):
            return self.coarse_moments(c, cell)
        s = [min(max(int(code85[k] * 3.0 - 3.0 * cell[k]), 0), 2) for k in range(3)]
        base = [3 * cell[k] + s[k] for k in range(3)]
        lo = np.array([srgb_eotf(b / 255.0) for b in base])
        hi = np.array([srgb_eotf((b + 1) / 255.0) for b in base])
        return self._interpolate(c, base, 1, np.clip((c - lo) / (hi - lo), 0.0, 1.0))

    def reflectance_lagrange(self, rgb):
        c = np.clip(np.asarray(rgb, np.float32), 0.0, 1.0)
        if c.max() - c.min() <= np.float32(1e-4):
            return np.array([(c[0] + c[1] + c[2]) / 3.0, 0.0, 0.0, 2.0])
        lam = fourier_lagrange(self.uplift_moments(c.astype(np.float64)))
        return np.array([lam[0], lam[1], lam[2], 1.0])

    def reflectance_spectrum(self, lagrange):
        """rho(lambda) at every nanometre for a (L0, L1, L2, w) record."""
        if lagrange[3] == 2.0:
            return np.full(SAMPLES, lagrange[0])
        x = np.cos(self.phases)
        return np.arctan(lagrange[0] + 2 * lagrange[1] * x + 2 * lagrange[2] * (2 * x * x - 1)) / math.pi + 0.5

    def tables(self):
        """float32 tables in the layout the integrators read: icdf (1025), cmf (471, 4), basis (471, 2, 4)."""
        icdf, _ = self.inverse_cdf()
        icdf32 = icdf.astype(np.float32)
        for i in range(1, len(icdf32)):
            if icdf32[i] <= icdf32[i - 1]:
                icdf32[i] = np.nextafter(icdf32[i - 1], np.float32(np.inf))
        cmf = np.zeros((SAMPLES, 4), np.float32)
        cmf[:, :3] = self.rgb_cmf
        cmf[:, 3] = np.cos(self.phases)
        corners = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]
        spectra = [self.reflectance_spectrum(self.reflectance_lagrange(c)) for c in corners]
        basis = np.zeros((SAMPLES, 2, 4), np.float32)
        for k in range(3):
            basis[:, 0, k] = spectra[k]
            basis[:, 1, k] = spectra[3 + k]
        basis[:, 0, 3] = self.d65
        return icdf32, cmf, basis
