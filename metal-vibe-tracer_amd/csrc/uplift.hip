// uplift.hip -- the reflectance-uplift grid built on the device, as the reference refines it once when
// the grid is loaded (PathTracerRenderer.refineSpectralGrid: spectral_refine_flags_kernel and
// spectral_refine_solve_kernel, main.swift:8187-8278).
//
// coarse_node_kernel: the exact moments of the 86^3 coarse nodes (8-bit codes 0, 3, ..., 255), the
//   grid the reference loads from FourierSRGB86.bin: the generator's solve (scripts/
//   generate_spectral_tables.py:381-386, Levenberg-Marquardt from zero to 1e-10 in double) and its
//   1024-node midpoint moments, one node per lane in double.
// refine_flags_kernel: every 8-bit code (2^24 lanes) through the coarse conversion -- trilinear moments
//   in linear light, the bounded MESE's multipliers, the reflectance's linear sRGB under D65, back to
//   8 bits -- in float as the reference's kernel; a code more than 0.35 steps off flags its cell.
// refine_solve_kernel: the 4 x 4 x 4 codes of each flagged cell, Levenberg-Marquardt in float from the
//   interpolated start (40 iterations, 24 damping trials, 2e-7) and the midpoint moments, as the
//   reference's kernel.
// The reflectance -> linear sRGB weights and the warped-phase cosines come from the host tables
// (uplift_solve_tables), in double and rounded to float.
#include <hip/hip_runtime.h>

#include "uplift.h"

namespace vt {
namespace {

constexpr int kSamples = 471;
constexpr uint32_t kNodes = 86, kCells = 85;
constexpr float kRcpPiF = 0.318309886f;
constexpr double kPiD = 3.14159265358979323846;
constexpr float kEpsF = 1e-4f;
constexpr double kEpsD = 1e-4;

struct Tables {
    const double *w, *c1, *c2;   // w: 3 rows of kSamples
    const float *wf, *c1f, *c2f;
};

// ---------------------------------------------------------------- float (the reference's kernels)
__device__ __forceinline__ float eotf_f(float v) { return v <= 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f); }
__device__ __forceinline__ float oetf_f(float c) {
    return c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}
// Linear sRGB of the reflectance with multipliers L (spectral_rgb_of_lagrange), with the Jacobian
// columns d rgb / d L0, L1, L2 when jac is set.
__device__ void rgb_of_lagrange_f(const Tables &t, const float L[3], float rgb[3], float (*jac)[3]) {
    float r[3] = {0.0f, 0.0f, 0.0f}, j[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    for (int i = 0; i < kSamples; ++i) {
        const float c1 = t.c1f[i], c2 = t.c2f[i];
        const float s = L[0] + L[1] * c1 + L[2] * c2;
        const float v = atanf(s) * kRcpPiF + 0.5f, dv = kRcpPiF / (1.0f + s * s);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float w = t.wf[k * kSamples + i];
            r[k] += w * v;
            if (jac) {
                const float d = w * dv;
                j[0][k] += d;
                j[1][k] += d * c1;
                j[2][k] += d * c2;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rgb[k] = r[k];
        if (jac)
            for (int c = 0; c < 3; ++c) jac[c][k] = j[c][k];
    }
}

// The bounded MESE's multipliers of three trigonometric moments (vibe_fourier_lagrange: Peters et al.
// 2019 with Levinson's recursion and the biasing of FOURIERSRGB2019 Alg. 2), in T.
template <class T>
struct Cx {
    T x, y;
};
template <class T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) {
    return {a.x + b.x, a.y + b.y};
}
template <class T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) {
    return {a.x - b.x, a.y - b.y};
}
template <class T>
__device__ __forceinline__ Cx<T> cscale(T s, Cx<T> a) {
    return {s * a.x, s * a.y};
}
template <class T>
__device__ __forceinline__ T cdot(Cx<T> a, Cx<T> b) {
    return a.x * b.x + a.y * b.y;
}
template <class T>
__device__ __forceinline__ Cx<T> cconj(Cx<T> a) {
    return {a.x, -a.y};
}
template <class T>
__device__ void fourier_lagrange(const T c[3], T out[3]) {
    const T pi = (T)kPiD, eps = (T)kEpsD;
    const T c0 = fmin(fmax(c[0], eps), (T)1 - eps);
    const Cx<T> g0p = cscale((T)1 / ((T)4 * pi), Cx<T>{sin(pi * c0), -cos(pi * c0)});
    const T g0 = (T)2 * g0p.x;
    Cx<T> g1 = cscale((T)2 * pi * c[1], Cx<T>{-g0p.y, g0p.x});
    Cx<T> g2 = cscale(pi, Cx<T>{-((T)2 * c[2] * g0p.y + c[1] * g1.y), (T)2 * c[2] * g0p.x + c[1] * g1.x});
    const T q0 = (T)1 / g0;
    Cx<T> u1 = cscale(q0, g1);
    T n1 = cdot(u1, u1), keep = (T)0.9999;
    if (n1 >= (T)1) {
        u1 = cscale(keep / sqrt(n1), u1);
        g1 = cscale((T)1 / q0, u1);
        n1 = keep * keep;
        keep = 0;
    }
    const T d1 = (T)1 / ((T)1 - n1), a0 = q0 * d1;
    const Cx<T> a1 = cscale(-(q0 * d1), u1);
    Cx<T> u2 = cadd(cscale(a0, g2), cmul(a1, g1));
    T n2 = cdot(u2, u2);
    if (n2 >= (T)1) {
        u2 = cscale(keep / sqrt(n2), u2);
        g2 = cscale((T)1 / a0, csub(u2, cmul(a1, g1)));
        n2 = keep * keep;
    }
    const T d2 = (T)1 / ((T)1 - n2), b0 = a0 * d2;
    const Cx<T> b1 = cscale(d2, csub(a1, cmul(u2, cconj(a1)))), b2 = cscale(-(a0 * d2), u2);
    const Cx<T> r0 = {b0 * b0 + cdot(b1, b1) + cdot(b2, b2), (T)0};
    const Cx<T> r1 = cadd(cmul(cconj(b1), Cx<T>{b0, (T)0}), cmul(cconj(b2), b1)), r2 = cmul(cconj(b2), Cx<T>{b0, (T)0});
    const Cx<T> s0 = cadd(cadd(cmul(g0p, r0), cmul(g1, r1)), cmul(g2, r2)), s1 = cadd(cmul(g0p, r1), cmul(g1, r2)),
                s2 = cmul(g0p, r2);
    const T scale = (T)2 / b0;
    out[0] = s0.y * scale;
    out[1] = s1.y * scale;
    out[2] = s2.y * scale;
}

// Trilinear moments of coarse cell i at linear colour c (spectral_moments_coarse): nodes at the linear
// values of codes 3i.
__device__ void coarse_moments_f(const float4 *grid, const float *nodes, const float c[3], const uint32_t i[3],
                                 float m[3]) {
    float t[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo = nodes[i[k]], hi = nodes[i[k] + 1];
        t[k] = fminf(fmaxf((c[k] - lo) / (hi - lo), 0.0f), 1.0f);
    }
    m[0] = m[1] = m[2] = 0.0f;
    for (uint32_t a = 0; a < 2; ++a)
        for (uint32_t b = 0; b < 2; ++b) {
            const uint32_t row = ((i[0] + a) * kNodes + (i[1] + b)) * kNodes + i[2];
            const float w = (a == 0 ? 1.0f - t[0] : t[0]) * (b == 0 ? 1.0f - t[1] : t[1]);
            const float4 p = grid[row], q = grid[row + 1];
            m[0] += w * ((1.0f - t[2]) * p.x + t[2] * q.x);
            m[1] += w * ((1.0f - t[2]) * p.y + t[2] * q.y);
            m[2] += w * ((1.0f - t[2]) * p.z + t[2] * q.z);
        }
}
__device__ __forceinline__ void code85_cell(const float c[3], float code85[3], uint32_t cell[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        code85[k] = oetf_f(c[k]) * 85.0f;
        cell[k] = min((uint32_t)fmaxf(code85[k], 0.0f), kCells - 1u);
    }
}

__global__ __launch_bounds__(256) void refine_flags_kernel(Tables t, const float4 *__restrict__ grid,
                                                           const float *__restrict__ nodes, uint8_t *__restrict__ flags) {
    const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
    if (tid >= (1u << 24)) return;
    const uint32_t code[3] = {tid >> 16, (tid >> 8) & 255u, tid & 255u};
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = eotf_f((float)code[k] / 255.0f);
    if (c[0] == c[1] && c[1] == c[2]) return;
    float code85[3], m[3], L[3], rgb[3];
    uint32_t cell[3];
    code85_cell(c, code85, cell);
    coarse_moments_f(grid, nodes, c, cell, m);
    fourier_lagrange<float>(m, L);
    rgb_of_lagrange_f(t, L, rgb, nullptr);
    float err = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const float v = fminf(fmaxf(rgb[k], 0.0f), 1.0f);
        err = fmaxf(err, fabsf(oetf_f(v) * 255.0f - (float)code[k]));
    }
    if (err > 0.35f) flags[(cell[0] * kCells + cell[1]) * kCells + cell[2]] = 1u;
}

// spectral_inverse3: the rows of the inverse are the cross products of the columns, over the
// determinant; m[c] is column c.
__device__ void inverse3_f(const float m[3][3], float r[3][3]) {
    const float *a = m[0], *b = m[1], *c = m[2];
    const float r0[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    const float r1[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    const float r2[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const float det = a[0] * r0[0] + a[1] * r0[1] + a[2] * r0[2];
    for (int k = 0; k < 3; ++k) {
        r[0][k] = r0[k] / det;
        r[1][k] = r1[k] / det;
        r[2][k] = r2[k] / det;
    }
}

__global__ __launch_bounds__(256) void refine_solve_kernel(Tables t, const float4 *__restrict__ grid,
                                                           const float *__restrict__ nodes,
                                                           const uint32_t *__restrict__ cells, uint32_t count,
                                                           float4 *__restrict__ refined) {
    const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
    if (tid >= 64u * count) return;
    const uint32_t slot = tid / 64u, entry = tid % 64u, cell = cells[slot];
    const uint32_t i[3] = {cell / (kCells * kCells), (cell / kCells) % kCells, cell % kCells};
    const uint32_t code[3] = {min(3u * i[0] + entry / 16u, 255u), min(3u * i[1] + (entry / 4u) % 4u, 255u),
                              min(3u * i[2] + entry % 4u, 255u)};
    float c[3], target[3], m[3], L[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = eotf_f((float)code[k] / 255.0f);
        target[k] = fminf(fmaxf(c[k], kEpsF), 1.0f - kEpsF);
    }
    coarse_moments_f(grid, nodes, c, i, m);  // the start: the refined cell's own coarse interpolation
    fourier_lagrange<float>(m, L);
    float J[3][3], rgb[3], f[3];
    rgb_of_lagrange_f(t, L, rgb, J);
    for (int k = 0; k < 3; ++k) f[k] = target[k] - rgb[k];
    float error = fmaxf(fabsf(f[0]), fmaxf(fabsf(f[1]), fabsf(f[2]))), mu = 0.0f;
    for (int n = 0; n < 40 && error > 2e-7f; ++n) {
        // normal = J^T J and gradient = J^T f with J's columns J[c]
        float normal[3][3], grad[3];
        for (int a = 0; a < 3; ++a) {
            for (int b = 0; b < 3; ++b) normal[b][a] = J[a][0] * J[b][0] + J[a][1] * J[b][1] + J[a][2] * J[b][2];
            grad[a] = J[a][0] * f[0] + J[a][1] * f[1] + J[a][2] * f[2];
        }
        bool improved = false;
        for (int k = 0; k < 24 && !improved; ++k) {
            float damped[3][3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) damped[a][b] = normal[a][b] * (a == b ? 1.0f + mu : 1.0f);
            const float det = damped[0][0] * (damped[1][1] * damped[2][2] - damped[2][1] * damped[1][2]) -
                              damped[1][0] * (damped[0][1] * damped[2][2] - damped[2][1] * damped[0][2]) +
                              damped[2][0] * (damped[0][1] * damped[1][2] - damped[1][1] * damped[0][2]);
            if (!(fabsf(det) > 0.0f)) break;
            float inv[3][3];  // inv[row][col]
            inverse3_f(damped, inv);
            float trial[3];
            for (int a = 0; a < 3; ++a) trial[a] = L[a] + (inv[a][0] * grad[0] + inv[a][1] * grad[1] + inv[a][2] * grad[2]);
            float Jt[3][3], rt[3], ft[3];
            rgb_of_lagrange_f(t, trial, rt, Jt);
            for (int a = 0; a < 3; ++a) ft[a] = target[a] - rt[a];
            const float et = fmaxf(fabsf(ft[0]), fmaxf(fabsf(ft[1]), fabsf(ft[2])));
            if (et < error) {
                for (int a = 0; a < 3; ++a) {
                    L[a] = trial[a];
                    f[a] = ft[a];
                    for (int b = 0; b < 3; ++b) J[a][b] = Jt[a][b];
                }
                error = et;
                mu = mu > 1e-12f ? mu * 0.1f : 0.0f;
                improved = true;
            } else {
                mu = fmaxf(mu * 10.0f, 1e-6f);
            }
        }
        if (!improved) break;
    }
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int k = 0; k < 1024; ++k) {
        const float x = cosf(-3.14159265f + ((float)k + 0.5f) * (3.14159265f / 1024.0f)), x2 = 2.0f * x * x - 1.0f;
        const float g = atanf(L[0] + 2.0f * L[1] * x + 2.0f * L[2] * x2) * kRcpPiF + 0.5f;
        s0 += g;
        s1 += g * x;
        s2 += g * x2;
    }
    refined[64u * slot + entry] = make_float4(s0 / 1024.0f, s1 / 1024.0f, s2 / 1024.0f, 0.0f);
}

// ---------------------------------------------------------------- double (the generator's solve)
__device__ void rgb_jac_d(const Tables &t, const double l[3], double rgb[3], double jac[9]) {
    for (int k = 0; k < 3; ++k) rgb[k] = 0.0;
    for (int k = 0; k < 9; ++k) jac[k] = 0.0;
    for (int i = 0; i < kSamples; ++i) {
        const double a1 = t.c1[i], a2 = t.c2[i];
        const double x = l[0] + l[1] * a1 + l[2] * a2;
        const double v = atan(x) / kPiD + 0.5, dv = 1.0 / (kPiD * (1.0 + x * x));
        for (int k = 0; k < 3; ++k) {
            const double w = t.w[k * kSamples + i];
            rgb[k] += w * v;
            const double d = w * dv;
            jac[3 * k] += d;
            jac[3 * k + 1] += d * a1;
            jac[3 * k + 2] += d * a2;
        }
    }
}
__device__ bool solve3_d(const double m[9], const double f[3], double x[3]) {
    const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], g_ = m[5], g = m[6], h = m[7], i = m[8];
    const double det = a * (e * i - g_ * h) - b * (d * i - g_ * g) + c * (d * h - e * g);
    if (det == 0.0 || !isfinite(det)) return false;
    x[0] = (f[0] * (e * i - g_ * h) - b * (f[1] * i - g_ * f[2]) + c * (f[1] * h - e * f[2])) / det;
    x[1] = (a * (f[1] * i - g_ * f[2]) - f[0] * (d * i - g_ * g) + c * (d * f[2] - f[1] * g)) / det;
    x[2] = (a * (e * f[2] - f[1] * h) - b * (d * f[2] - f[1] * g) + f[0] * (d * h - e * g)) / det;
    return true;
}

__global__ __launch_bounds__(256) void coarse_node_kernel(Tables t, float4 *__restrict__ grid) {
    const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
    if (tid >= kNodes * kNodes * kNodes) return;
    const uint32_t node[3] = {tid / (kNodes * kNodes), (tid / kNodes) % kNodes, tid % kNodes};
    double target[3];
    for (int k = 0; k < 3; ++k) {
        const double v = (double)(3u * node[k]) / 255.0;
        const double lin = v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4);
        target[k] = fmin(fmax(lin, kEpsD), 1.0 - kEpsD);
    }
    double lam[3] = {0.0, 0.0, 0.0}, rgb[3], jac[9];
    rgb_jac_d(t, lam, rgb, jac);
    double f[3] = {target[0] - rgb[0], target[1] - rgb[1], target[2] - rgb[2]};
    double err = fmax(fabs(f[0]), fmax(fabs(f[1]), fabs(f[2]))), mu = 0.0;
    bool stuck = false;
    for (int n = 0; n < 80 && err > 1e-10 && !stuck; ++n) {
        double normal[9], grad[3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) normal[3 * r + c] = jac[r] * jac[c] + jac[3 + r] * jac[3 + c] + jac[6 + r] * jac[6 + c];
            grad[r] = jac[r] * f[0] + jac[3 + r] * f[1] + jac[6 + r] * f[2];
        }
        double trial[3], rgb_t[3], jac_t[9], f_t[3], err_t = err;
        bool accepted = false;
        while (!accepted) {
            double damped[9];
            for (int k = 0; k < 9; ++k) damped[k] = normal[k];
            damped[0] *= 1.0 + mu;
            damped[4] *= 1.0 + mu;
            damped[8] *= 1.0 + mu;
            double step[3];
            if (!solve3_d(damped, grad, step)) {
                stuck = true;
                break;
            }
            for (int k = 0; k < 3; ++k) trial[k] = lam[k] + step[k];
            rgb_jac_d(t, trial, rgb_t, jac_t);
            for (int k = 0; k < 3; ++k) f_t[k] = target[k] - rgb_t[k];
            err_t = fmax(fabs(f_t[0]), fmax(fabs(f_t[1]), fabs(f_t[2])));
            if (err_t < err) {
                mu = mu > 1e-12 ? mu * 0.1 : 0.0;
                accepted = true;
            } else {
                mu = fmax(mu * 10.0, 1e-6);
                if (mu > 1e8) {
                    stuck = true;
                    break;
                }
            }
        }
        if (!accepted) break;
        for (int k = 0; k < 3; ++k) {
            lam[k] = trial[k];
            f[k] = f_t[k];
        }
        for (int k = 0; k < 9; ++k) jac[k] = jac_t[k];
        err = err_t;
    }
    // 1024-node midpoint moments (generate_spectral_tables.py:227-239).
    const double h = kPiD / 1024.0;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (int i = 0; i < 1024; ++i) {
        const double cp = cos(-kPiD + (i + 0.5) * h), c2p = 2.0 * cp * cp - 1.0;
        const double v = atan(lam[0] + 2.0 * lam[1] * cp + 2.0 * lam[2] * c2p) / kPiD + 0.5;
        c0 += v;
        c1 += v * cp;
        c2 += v * c2p;
    }
    grid[tid] = make_float4((float)(c0 * h / kPiD), (float)(c1 * h / kPiD), (float)(c2 * h / kPiD), 0.0f);
}

}  // namespace

hipError_t launch_uplift_coarse(const UpliftTables &tb, float4 *grid, hipStream_t stream) {
    const Tables t{tb.w, tb.c1, tb.c2, tb.wf, tb.c1f, tb.c2f};
    const uint32_t n = kNodes * kNodes * kNodes;
    hipLaunchKernelGGL(coarse_node_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, t, grid);
    return hipGetLastError();
}
hipError_t launch_uplift_flags(const UpliftTables &tb, const float4 *grid, const float *nodes, uint8_t *flags,
                               hipStream_t stream) {
    const Tables t{tb.w, tb.c1, tb.c2, tb.wf, tb.c1f, tb.c2f};
    hipLaunchKernelGGL(refine_flags_kernel, dim3((1u << 24) / 256), dim3(256), 0, stream, t, grid, nodes, flags);
    return hipGetLastError();
}
hipError_t launch_uplift_refine(const UpliftTables &tb, const float4 *grid, const float *nodes, const uint32_t *cells,
                                uint32_t count, float4 *refined, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const Tables t{tb.w, tb.c1, tb.c2, tb.wf, tb.c1f, tb.c2f};
    hipLaunchKernelGGL(refine_solve_kernel, dim3((64u * count + 255) / 256), dim3(256), 0, stream, t, grid, nodes,
                       cells, count, refined);
    return hipGetLastError();
}

}  // namespace vt
