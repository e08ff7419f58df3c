// spectral.cpp -- host setup of the spectral tables (spectral.h).
#include "spectral.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
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

namespace {

double solve_with(const SolveData &s, const double target_in[3], double lam[3]) {
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

double srgb_eotf(double v) { return v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4); }
double srgb_oetf(double v) {
    v = std::min(std::max(v, 0.0), 1.0);
    return v <= 0.0031308 ? 12.92 * v : 1.055 * std::pow(v, 1.0 / 2.4) - 0.055;
}

// Bounded trigonometric moments c_j = (1/pi) int_{-pi}^{0} g(phi) cos(j phi) dphi of the reflectance
// of Lagrange multipliers l, by the 1024-node midpoint rule (generate_spectral_tables.py:227-239).
void moments_of_lagrange(const double l[3], double m[3]) {
    const double h = kPi / 1024.0;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (int i = 0; i < 1024; ++i) {
        double cp = std::cos(-kPi + (i + 0.5) * h), c2p = 2.0 * cp * cp - 1.0;
        double v = std::atan(l[0] + 2.0 * l[1] * cp + 2.0 * l[2] * c2p) / kPi + 0.5;
        c0 += v;
        c1 += v * cp;
        c2 += v * c2p;
    }
    m[0] = c0 * h / kPi;
    m[1] = c1 * h / kPi;
    m[2] = c2 * h / kPi;
}

// Lagrange multipliers of the bounded MESE of three real moments (Peters et al. 2019, Eqs. 6, 7, 10;
// Levinson's recursion with the biasing of FOURIERSRGB2019 Alg. 2), as the reference's host copy
// SpectralColour.fourierLagrange (main.swift:8640-8680) and MSL vibe_fourier_lagrange: moments
// outside the valid set are pulled back, so every input yields a reflectance.
struct Cx {
    double x, y;
};
Cx cmul(Cx a, Cx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
Cx conj(Cx a) { return {a.x, -a.y}; }
Cx operator+(Cx a, Cx b) { return {a.x + b.x, a.y + b.y}; }
Cx operator-(Cx a, Cx b) { return {a.x - b.x, a.y - b.y}; }
Cx operator*(double s, Cx a) { return {s * a.x, s * a.y}; }
double cdot(Cx a, Cx b) { return a.x * b.x + a.y * b.y; }

void fourier_lagrange(const double c[3], double out[3]) {
    const double c0 = std::min(std::max(c[0], kEpsilon), 1.0 - kEpsilon);
    const Cx g0p = (1.0 / (4.0 * kPi)) * Cx{std::sin(kPi * c0), -std::cos(kPi * c0)};
    const double g0 = 2.0 * g0p.x;
    Cx g1 = (2.0 * kPi * c[1]) * Cx{-g0p.y, g0p.x};
    Cx g2 = kPi * Cx{-(2.0 * c[2] * g0p.y + c[1] * g1.y), 2.0 * c[2] * g0p.x + c[1] * g1.x};
    const double q0 = 1.0 / g0;
    Cx u1 = q0 * g1;
    double n1 = cdot(u1, u1), keep = 0.9999;
    if (n1 >= 1.0) {
        u1 = (keep / std::sqrt(n1)) * u1;
        g1 = (1.0 / q0) * u1;
        n1 = keep * keep;
        keep = 0.0;
    }
    const double d1 = 1.0 / (1.0 - n1), a0 = q0 * d1;
    const Cx a1 = (-(q0 * d1)) * u1;
    Cx u2 = a0 * g2 + cmul(a1, g1);
    double n2 = cdot(u2, u2);
    if (n2 >= 1.0) {
        u2 = (keep / std::sqrt(n2)) * u2;
        g2 = (1.0 / a0) * (u2 - cmul(a1, g1));
        n2 = keep * keep;
    }
    const double d2 = 1.0 / (1.0 - n2), b0 = a0 * d2;
    const Cx b1 = d2 * (a1 - cmul(u2, conj(a1))), b2 = (-(a0 * d2)) * u2;
    const Cx r0 = {b0 * b0 + cdot(b1, b1) + cdot(b2, b2), 0.0};
    const Cx r1 = cmul(conj(b1), Cx{b0, 0.0}) + cmul(conj(b2), b1), r2 = cmul(conj(b2), Cx{b0, 0.0});
    const Cx s0 = cmul(g0p, r0) + cmul(g1, r1) + cmul(g2, r2), s1 = cmul(g0p, r1) + cmul(g1, r2), s2 = cmul(g0p, r2);
    const double scale = 2.0 / b0;
    out[0] = s0.y * scale;
    out[1] = s1.y * scale;
    out[2] = s2.y * scale;
}

// Linear sRGB under D65 of the reflectance with multipliers l (the generator's rgb_of_lagrange).
void rgb_of_lagrange(const SolveData &s, const double l[3], double rgb[3]) {
    rgb[0] = rgb[1] = rgb[2] = 0.0;
    for (int i = 0; i < kSpectralSamples; ++i) {
        double v = std::atan(l[0] + l[1] * s.c1[i] + l[2] * s.c2[i]) / kPi + 0.5;
        for (int k = 0; k < 3; ++k) rgb[k] += s.w[k][i] * v;
    }
}

// The reference's uplift grid, evaluated lazily: exact moments of the coarse nodes (8-bit codes
// 0, 3, ..., 255; FourierSRGB86.bin) and of the codes of refined cells, and the refinement flag of
// each cell, cached for the process (every scene's reflectances are taken under the same D65 data).
struct UpliftGrid {
    SolveData s;
    std::mutex mu;
    std::unordered_map<uint32_t, std::array<float, 3>> exact;
    std::unordered_map<uint32_t, bool> refined;
    float nodes[86];

    explicit UpliftGrid(const SpectralTables &t) {
        solve_data(t, s);
        for (int i = 0; i < 86; ++i) nodes[i] = (float)srgb_eotf(3.0 * i / 255.0);
    }

    // Exact moments of code (r, g, b): Levenberg-Marquardt to the code's linear colour, then the
    // 1024-node moments (the generator's solve_code, :381-386), stored as float like the grid file.
    std::array<float, 3> code_moments(int r, int g, int b) {
        const uint32_t key = (uint32_t)r << 16 | (uint32_t)g << 8 | (uint32_t)b;
        {
            std::lock_guard<std::mutex> lock(mu);
            auto it = exact.find(key);
            if (it != exact.end()) return it->second;
        }
        const double target[3] = {srgb_eotf(r / 255.0), srgb_eotf(g / 255.0), srgb_eotf(b / 255.0)};
        double l[3], m[3];
        solve_with(s, target, l);
        moments_of_lagrange(l, m);
        const std::array<float, 3> v = {(float)m[0], (float)m[1], (float)m[2]};
        std::lock_guard<std::mutex> lock(mu);
        exact.emplace(key, v);
        return v;
    }

    // Trilinear interpolation in linear light of the coarse cell (i) at colour c (MSL
    // spectral_moments_coarse, main.swift:855-869).
    void coarse(const double c[3], const int i[3], double m[3]) {
        double t[3];
        for (int k = 0; k < 3; ++k) {
            const double lo = nodes[i[k]], hi = nodes[i[k] + 1];
            t[k] = std::min(std::max((c[k] - lo) / (hi - lo), 0.0), 1.0);
        }
        m[0] = m[1] = m[2] = 0.0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) {
                    const double w = (a ? t[0] : 1.0 - t[0]) * (b ? t[1] : 1.0 - t[1]) * (d ? t[2] : 1.0 - t[2]);
                    if (w == 0.0) continue;
                    const auto e = code_moments(3 * (i[0] + a), 3 * (i[1] + b), 3 * (i[2] + d));
                    for (int k = 0; k < 3; ++k) m[k] += w * (double)e[k];
                }
    }

    // Whether the reference refines cell i: some 8-bit code of the cell comes back from the coarse
    // conversion more than 0.35 steps off (spectral_code_error / spectral_refine_flags_kernel,
    // main.swift:8212-8233).  The codes of cell i are 3i, 3i + 1, 3i + 2 per axis (and 255 in the
    // last cell); greys are skipped.
    bool is_refined(const int i[3]) {
        const uint32_t key = (uint32_t)i[0] << 16 | (uint32_t)i[1] << 8 | (uint32_t)i[2];
        {
            std::lock_guard<std::mutex> lock(mu);
            auto it = refined.find(key);
            if (it != refined.end()) return it->second;
        }
        bool flag = false;
        int n[3];
        for (int k = 0; k < 3; ++k) n[k] = i[k] == 84 ? 4 : 3;
        for (int a = 0; a < n[0] && !flag; ++a)
            for (int b = 0; b < n[1] && !flag; ++b)
                for (int d = 0; d < n[2] && !flag; ++d) {
                    const int code[3] = {3 * i[0] + a, 3 * i[1] + b, 3 * i[2] + d};
                    if (code[0] == code[1] && code[1] == code[2]) continue;
                    double c[3], m[3], l[3], rgb[3];
                    for (int k = 0; k < 3; ++k) c[k] = srgb_eotf(code[k] / 255.0);
                    coarse(c, i, m);
                    fourier_lagrange(m, l);
                    rgb_of_lagrange(s, l, rgb);
                    for (int k = 0; k < 3; ++k)
                        if (std::fabs(srgb_oetf(rgb[k]) * 255.0 - code[k]) > 0.35) flag = true;
                }
        std::lock_guard<std::mutex> lock(mu);
        refined.emplace(key, flag);
        return flag;
    }

    // Moments of a bounded linear colour (MSL spectral_moments, main.swift:870-884; host copy
    // SpectralColour.lagrange, :8596-8635): coarse cell, or in a refined cell the exact moments of
    // its codes interpolated at one-code spacing.
    void moments(const double c[3], double m[3]) {
        int i[3];
        double code85[3];
        for (int k = 0; k < 3; ++k) {
            code85[k] = srgb_oetf(c[k]) * 85.0;
            i[k] = std::min(std::max((int)code85[k], 0), 84);
        }
        if (!is_refined(i)) {
            coarse(c, i, m);
            return;
        }
        int n0[3];
        double t[3];
        for (int k = 0; k < 3; ++k) {
            const int sk = std::min(std::max((int)(code85[k] * 3.0 - 3.0 * i[k]), 0), 2);
            n0[k] = 3 * i[k] + sk;
            const double lo = srgb_eotf(n0[k] / 255.0), hi = srgb_eotf((n0[k] + 1) / 255.0);
            t[k] = std::min(std::max((c[k] - lo) / (hi - lo), 0.0), 1.0);
        }
        m[0] = m[1] = m[2] = 0.0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) {
                    const double w = (a ? t[0] : 1.0 - t[0]) * (b ? t[1] : 1.0 - t[1]) * (d ? t[2] : 1.0 - t[2]);
                    if (w == 0.0) continue;
                    const auto e = code_moments(std::min(n0[0] + a, 255), std::min(n0[1] + b, 255), std::min(n0[2] + d, 255));
                    for (int k = 0; k < 3; ++k) m[k] += w * (double)e[k];
                }
    }
};

UpliftGrid &uplift_grid(const SpectralTables &t) {
    static UpliftGrid *g = new UpliftGrid(t);
    return *g;
}

}  // namespace

void uplift_solve_tables(const SpectralTables &t, double *w, double *c1, double *c2, float *nodes) {
    SolveData s;
    solve_data(t, s);
    for (int i = 0; i < kSpectralSamples; ++i) {
        for (int k = 0; k < 3; ++k) w[k * kSpectralSamples + i] = s.w[k][i];
        c1[i] = s.c1[i];
        c2[i] = s.c2[i];
    }
    for (int i = 0; i < 86; ++i) nodes[i] = (float)srgb_eotf(3.0 * i / 255.0);
}

double solve_lagrange(const SpectralTables &t, const double target[3], double lam[3]) {
    SolveData s;
    solve_data(t, s);
    return solve_with(s, target, lam);
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
    const double cd[3] = {c[0], c[1], c[2]};
    double m[3], l[3];
    uplift_grid(t).moments(cd, m);
    fourier_lagrange(m, l);
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
    const float corners[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
    double lag[6][3];
    for (int c = 0; c < 6; ++c) {
        float l[4];
        reflectance_lagrange(t, corners[c], l);
        for (int k = 0; k < 3; ++k) lag[c][k] = l[k];
    }
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
