// spectral.h -- host setup of the 4-wavelength spectral transport.
//
// The integrator keeps scene colours in linear sRGB and converts them where they are used
// (main.swift:717-1111): reflectances through the bounded MESE of three trigonometric moments
// (Peters et al. 2019), emitters as white + saturated-corner spectra times the illuminant (Smits
// 1999), contributions back to linear sRGB through per-nanometre colour-matching functions.
// Everything here runs once per scene on the host, in double precision; the GPU reads the float
// tables below.
#pragma once
#include <cstdint>

namespace vt {

constexpr int kLambdaMin = 360;
constexpr int kSpectralSamples = 471;  // 360-830 nm at 1 nm
constexpr int kIcdfSegments = 1024;

struct SpectralTables {
    // Wavelength inverse CDF at u = i / 1024 of the D65 sampling density (1025 nodes).
    float icdf[kIcdfSegments + 1];
    // Per nanometre: linear-sRGB colour-matching functions (xyz -> sRGB) and cos of the warped phase.
    float cmf[kSpectralSamples][4];
    // Per nanometre: emission basis R, G, B (x, y, z) and the D65 SPD normalised to luminance one
    // (w); then C, M, Y (x, y, z), 0.
    float basis[kSpectralSamples][2][4];
    double xyz_to_rgb[3][3];
    double d65[kSpectralSamples];  // normalised to luminance Y = 1 (1 nm sums)
};

// Builds the tables for a scene lit by D65 emitters only.
void build_spectral_tables(SpectralTables &t);

// Reflectance representation of a linear-sRGB colour (spectral_lagrange, main.swift:889-896):
// (L0, L1, L2, 1) bounded-MESE Lagrange multipliers, or (grey, 0, 0, 2) for an exactly flat grey.
void reflectance_lagrange(const SpectralTables &t, const float rgb[3], float out[4]);

// Lagrange multipliers reproducing `target` (linear sRGB, channels clamped to [1e-4, 1 - 1e-4]) under
// D65: Levenberg-Marquardt in multiplier space.  Returns the max channel error.
double solve_lagrange(const SpectralTables &t, const double target[3], double lagrange[3]);

// The uplift solve's per-nanometre data: w = 3 rows of kSpectralSamples reflectance -> linear sRGB
// weights under D65, c1 = 2 cos(phi), c2 = 2 cos(2 phi) of the warped phase; nodes = the 86 coarse
// nodes' linear values (codes 0, 3, ..., 255).
void uplift_solve_tables(const SpectralTables &t, double *w, double *c1, double *c2, float *nodes);

}  // namespace vt
