// spectral.cpp -- host setup of the spectral tables (spectral.h).
#include "spectral.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace vt {
namespace {

#include "spectral_data.inc"

constexpr double kPi = 3.14159265358979323846;
constexpr double kEpsilon = 1e-4;  // moment / target clamp of the bounded MESE

void inverse3(const double m[3][3], double r[3][3]) {
    double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0], h = m[2][1],
           i = m[2][2];
    double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    r[0][0] = (e * i - f * h) / det; r[0][1] = (c * h - b * i) / det; r[0][2] = (b * f - c * e) / det;
    r[1][0] = (f * g - d * i) / det; r[1][1] = (a * i - c * g) / det; r[1][2] = (c * d - a * f) / det;
    r[2][0] = (d * h - e * g) / det; r[2][1] = (b * g - a * h) / det; r[2][2] = (a * e - b * d) / det;
}

// Peters' even warp, piecewise linear in 5 nm steps.
double phase_of(double lambda) {
    double t = (lambda - 360.0) * 0.2;
    int i = std::min(std::max((int)std::floor(t), 0), 93);
    double f = t - i;
    return kPhaseWarp[i] * (1.0 - f) + kPhaseWarp[i + 1] * f;
}

struct SolveData {
    double w[3][kSpectralSamples];  // reflectance -> linear sRGB weights under D65
    double c1[kSpectralSamples], c2[kSpectralSamples];
};

void solve_data(const SpectralTables &t, SolveData &s) {
    for (int i = 0; i < kSpectralSamples; ++i) {
        for (int k = 0; k < 3; ++k) {
            double rgb = t.xyz_to_rgb[k][0] * kCieXyz[i][0] + t.xyz_to_rgb[k][1] * kCieXyz[i][1] +
                         t.xyz_to_rgb[k][2] * kCieXyz[i][2];
            s.w[k][i] = rgb * t.d65[i];
        }
        double p = phase_of(kLambdaMin + i);
        s.c1[i] = 2.0 * std::cos(p);
        s.c2[i] = 2.0 * std::cos(2.0 * p);
    }
}

void rgb_and_jacobian(const SolveData &s, const double l[3], double rgb[3], double jac[9]) {
    for (int k = 0; k < 3; ++k) rgb[k] = 0;
    for (int k = 0; k < 9; ++k) jac[k] = 0;
    for (int i = 0; i < kSpectralSamples; ++i) {
        double a1 = s.c1[i], a2 = s.c2[i];
        double x = l[0] + l[1] * a1 + l[2] * a2;
        double v = std::atan(x) / kPi + 0.5, dv = 1.0 / (kPi * (1.0 + x * x));
        for (int k = 0; k < 3; ++k) {
            rgb[k] += s.w[k][i] * v;
            double d = s.w[k][i] * dv;
            jac[3 * k] += d;
            jac[3 * k + 1] += d * a1;
            jac[3 * k + 2] += d * a2;
        }
    }
}

bool solve3(const double m[9], const double f[3], double x[3]) {
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], g_ = m[5], g = m[6], h = m[7], i = m[8];
    double det = a * (e * i - g_ * h) - b * (d * i - g_ * g) + c * (d * h - e * g);
    if (det == 0.0 || !std::isfinite(det)) return false;
    x[0] = (f[0] * (e * i - g_ * h) - b * (f[1] * i - g_ * f[2]) + c * (f[1] * h - e * f[2])) / det;
    x[1] = (a * (f[1] * i - g_ * f[2]) - f[0] * (d * i - g_ * g) + c * (d * f[2] - f[1] * g)) / det;
    x[2] = (a * (e * f[2] - f[1] * h) - b * (d * f[2] - f[1] * g) + f[0] * (d * h - e * g)) / det;
    return true;
}

}  // namespace

double solve_lagrange(const SpectralTables &t, const double target_in[3], double lam[3]) {
    SolveData s;
    solve_data(t, s);
    double target[3];
    for (int k = 0; k < 3; ++k) target[k] = std::min(std::max(target_in[k], kEpsilon), 1.0 - kEpsilon);
    lam[0] = lam[1] = lam[2] = 0.0;
    double rgb[3], jac[9];
    rgb_and_jacobian(s, lam, rgb, jac);
    double f[3] = {target[0] - rgb[0], target[1] - rgb[1], target[2] - rgb[2]};
    double err = std::max(std::fabs(f[0]), std::max(std::fabs(f[1]), std::fabs(f[2])));
    double mu = 0.0;
    for (int n = 0; n < 80 && err > 1e-10; ++n) {
        // Normal equations J^T J step = J^T f, damped on the diagonal (Levenberg-Marquardt).
        double normal[9], grad[3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                double v = 0;
                for (int k = 0; k < 3; ++k) v += jac[3 * k + r] * jac[3 * k + c];
                normal[3 * r + c] = v;
            }
            double g = 0;
            for (int k = 0; k < 3; ++k) g += jac[3 * k + r] * f[k];
            grad[r] = g;
        }
        double trial[3], rgb_t[3], jac_t[9], f_t[3], err_t;
        for (;;) {
            double damped[9];
            std::memcpy(damped, normal, sizeof damped);
            for (int d : {0, 4, 8}) damped[d] *= 1.0 + mu;
            double step[3];
            if (!solve3(damped, grad, step)) return err;
            for (int k = 0; k < 3; ++k) trial[k] = lam[k] + step[k];
            rgb_and_jacobian(s, trial, rgb_t, jac_t);
            for (int k = 0; k < 3; ++k) f_t[k] = target[k] - rgb_t[k];
            err_t = std::max(std::fabs(f_t[0]), std::max(std::fabs(f_t[1]), std::fabs(f_t[2])));
            if (err_t < err) {
                mu = mu > 1e-12 ? mu * 0.1 : 0.0;
                break;
            }
            mu = std::max(mu * 10.0, 1e-6);
            if (mu > 1e8) return err;
        }
        std::memcpy(lam, trial, sizeof trial);
        std::memcpy(jac, jac_t, sizeof jac_t);
        std::memcpy(f, f_t, sizeof f_t);
        err = err_t;
    }
    return err;
}

void reflectance_lagrange(const SpectralTables &t, const float rgb[3], float out[4]) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = std::min(std::max(rgb[k], 0.0f), 1.0f);
    float lo = std::min(c[0], std::min(c[1], c[2])), hi = std::max(c[0], std::max(c[1], c[2]));
    if (hi - lo <= 1e-4f) {  // greys are exactly flat (main.swift:892-895)
        out[0] = (c[0] + c[1] + c[2]) / 3.0f;
        out[1] = out[2] = 0.0f;
        out[3] = 2.0f;
        return;
    }
    double target[3] = {c[0], c[1], c[2]}, l[3];
    solve_lagrange(t, target, l);
    out[0] = (float)l[0];
    out[1] = (float)l[1];
    out[2] = (float)l[2];
    out[3] = 1.0f;
}

void build_spectral_tables(SpectralTables &t) {
    // D65 normalised to luminance one with 1 nm sums; the xyz -> linear sRGB matrix from the sRGB
    // primaries with D65's own XYZ as white.
    double y = 0;
    for (int i = 0; i < kSpectralSamples; ++i) y += kD65[i] * kCieXyz[i][1];
    double white[3] = {0, 0, 0};
    for (int i = 0; i < kSpectralSamples; ++i) {
        t.d65[i] = kD65[i] / y;
        for (int k = 0; k < 3; ++k) white[k] += t.d65[i] * kCieXyz[i][k];
    }
    const double px[3] = {0.64, 0.30, 0.15}, py[3] = {0.33, 0.60, 0.06};
    double prim[3][3], inv[3][3], scale[3], rgb_to_xyz[3][3];
    for (int k = 0; k < 3; ++k) {
        prim[0][k] = px[k] / py[k];
        prim[1][k] = 1.0;
        prim[2][k] = (1.0 - px[k] - py[k]) / py[k];
    }
    inverse3(prim, inv);
    for (int r = 0; r < 3; ++r) scale[r] = inv[r][0] * white[0] + inv[r][1] * white[1] + inv[r][2] * white[2];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) rgb_to_xyz[r][k] = prim[r][k] * scale[k];
    inverse3(rgb_to_xyz, t.xyz_to_rgb);

    // Per-nanometre CMF table and sampling density s = |r| + |g| + |b|.
    std::vector<double> s(kSpectralSamples), is(kSpectralSamples);
    for (int i = 0; i < kSpectralSamples; ++i) {
        double rgb[3];
        for (int k = 0; k < 3; ++k)
            rgb[k] = t.xyz_to_rgb[k][0] * kCieXyz[i][0] + t.xyz_to_rgb[k][1] * kCieXyz[i][1] +
                     t.xyz_to_rgb[k][2] * kCieXyz[i][2];
        for (int k = 0; k < 3; ++k) t.cmf[i][k] = (float)rgb[k];
        t.cmf[i][3] = (float)std::cos(phase_of(kLambdaMin + i));
        s[i] = std::fabs(rgb[0]) + std::fabs(rgb[1]) + std::fabs(rgb[2]);
        is[i] = t.d65[i] * s[i];
    }
    // Density 0.9 S s / int(S s) + 0.1 s / int(s) (trapezoid integrals), inverted exactly.
    auto trapezoid = [](const std::vector<double> &v) {
        double a = 0;
        for (size_t i = 0; i + 1 < v.size(); ++i) a += 0.5 * (v[i] + v[i + 1]);
        return a;
    };
    double ns = trapezoid(s), nis = trapezoid(is);
    std::vector<double> p(kSpectralSamples), cdf(kSpectralSamples, 0.0);
    for (int i = 0; i < kSpectralSamples; ++i) p[i] = 0.9 * is[i] / nis + 0.1 * s[i] / ns;
    for (int i = 1; i < kSpectralSamples; ++i) cdf[i] = cdf[i - 1] + 0.5 * (p[i - 1] + p[i]);
    double total = cdf[kSpectralSamples - 1];
    std::vector<double> out(kIcdfSegments + 1);
    out[0] = kLambdaMin;
    int j = 0;
    for (int n = 1; n < kIcdfSegments; ++n) {
        double yv = total * n / kIcdfSegments;
        while (cdf[j + 1] < yv) ++j;
        double p0 = p[j], dp = p[j + 1] - p0, rest = yv - cdf[j];
        double tt = std::fabs(dp) < 1e-12 * std::max(p0, 1e-30) ? rest / p0
                                                                : (-p0 + std::sqrt(std::max(p0 * p0 + 2.0 * dp * rest, 0.0))) / dp;
        out[n] = kLambdaMin + j + std::min(std::max(tt, 0.0), 1.0);
    }
    out[kIcdfSegments] = kLambdaMin + kSpectralSamples - 1;
    for (int i = 0; i <= kIcdfSegments; ++i) {
        float v = (float)out[i];
        if (i > 0 && v <= t.icdf[i - 1]) v = std::nextafter(t.icdf[i - 1], 1e30f);
        t.icdf[i] = v;
    }

    // Emission basis: the bounded-MESE spectra of the six saturated corners (spectral_icdf_kernel,
    // main.swift:8117-8131), and the normalised D65.
    const double corners[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
    double lag[6][3];
    for (int c = 0; c < 6; ++c) solve_lagrange(t, corners[c], lag[c]);
    for (int i = 0; i < kSpectralSamples; ++i) {
        double ph = phase_of(kLambdaMin + i), c1 = std::cos(ph);
        float v[6];
        for (int c = 0; c < 6; ++c)
            v[c] = (float)(std::atan(lag[c][0] + 2 * lag[c][1] * c1 + 2 * lag[c][2] * (2 * c1 * c1 - 1)) / kPi + 0.5);
        t.basis[i][0][0] = v[0]; t.basis[i][0][1] = v[1]; t.basis[i][0][2] = v[2];
        t.basis[i][0][3] = (float)t.d65[i];
        t.basis[i][1][0] = v[3]; t.basis[i][1][1] = v[4]; t.basis[i][1][2] = v[5];
        t.basis[i][1][3] = 0.0f;
    }
}

}  // namespace vt
