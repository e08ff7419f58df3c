// render.hip -- gfx950 kernels of the wavefront spectral path tracer (render_kernels.h).
//
// Estimator: the reference's pass-2 path loop in MIS mode (main.swift:7370-7914) for Lambertian
// surfaces and triangle emitters: camera rays see emitters directly; at every non-delta vertex a
// next-event sample of an emitter chosen by area x luminance (sample_direct_light, imported branch,
// main.swift:2888-2917) is weighted by the power heuristic against the BSDF pdf; BSDF continuations
// that reach an emitter are weighted by emission_weight (main.swift:3462-3466); Russian roulette after
// the third bounce (main.swift:7819-7827); four stratified wavelengths per path (main.swift:795-833).
// Random numbers follow the reference's per-pixel PCG streams draw for draw.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "render_kernels.h"
#include "zsampler.h"
#include "traverse8.h"

constexpr int kNrcInputCount = 11;

namespace vt {
namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr int kIcdfSegments = 1024;
constexpr int kSamples = 471;

// ---------------------------------------------------------------- RNG (main.swift:91-117)
__device__ __forceinline__ uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float rand_f(uint32_t &seed) {
    seed = pcg_hash(seed);
    return (float)(seed >> 8u) * (1.0f / 16777216.0f);
}
// A path's random numbers (Sampler, main.swift:213-293): its PCG stream, or in Z mode the Z++ sampler
// under the pixel's key, each sampling event (a path vertex and a stream) owning 32 dimensions.
// rand_f3 is one 3D constituent in Z mode and three consecutive draws of the stream otherwise, so the
// PCG mode draws exactly what it did with single draws.
struct Rng {
    uint32_t pcg;
    ZSampler zs;
    bool z;
};
enum : uint32_t { kZBsdf = 0u, kZNee = 1u, kZRoulette = 2u, kZDi = 4u, kZGiBsdf = 5u, kZGiNee = 6u, kZWavelength = 12u };
__device__ __forceinline__ float rand_f(Rng &r) { return r.z ? r.zs.rand1() : rand_f(r.pcg); }
__device__ __forceinline__ void rand_f3(uint32_t &seed, float &a, float &b, float &c) {
    a = rand_f(seed);
    b = rand_f(seed);
    c = rand_f(seed);
}
__device__ __forceinline__ void rand_f2(uint32_t &seed, float &a, float &b) {
    a = rand_f(seed);
    b = rand_f(seed);
}
__device__ __forceinline__ void rand_f2(Rng &r, float &a, float &b) {
    if (r.z) r.zs.rand2(a, b);
    else rand_f2(r.pcg, a, b);
}
__device__ __forceinline__ void rand_f3(Rng &r, float &a, float &b, float &c) {
    if (r.z) r.zs.rand3(a, b, c);
    else rand_f3(r.pcg, a, b, c);
}
// sampler_event (main.swift:293-295); the PCG stream ignores events.
__device__ __forceinline__ void rng_event(uint32_t &, uint32_t, uint32_t) {}
__device__ __forceinline__ void rng_event(Rng &r, uint32_t vertex, uint32_t stream) {
    if (r.z) r.zs.event(vertex, stream);
}

// ---------------------------------------------------------------- small vector helpers
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ float max_abs(V3 p) { return fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))); }
__device__ __forceinline__ V3 load_v3(const float *p) { return v3(p[0], p[1], p[2]); }

struct S4 {
    float v[4];
};
__device__ __forceinline__ S4 s4(float a) { return S4{{a, a, a, a}}; }
__device__ __forceinline__ S4 mul(S4 a, S4 b) {
    return S4{{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}};
}
__device__ __forceinline__ S4 scale(S4 a, float s) { return S4{{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}}; }
__device__ __forceinline__ float smax(S4 a) { return fmaxf(fmaxf(a.v[0], a.v[1]), fmaxf(a.v[2], a.v[3])); }

// ---------------------------------------------------------------- tolerances (main.swift:1340-1369)
__device__ __forceinline__ float ray_epsilon(V3 p) { return fmaxf(1e-7f, 9.536743e-7f * max_abs(p)); }
__device__ __forceinline__ float ray_t_min(V3 o) { return ray_epsilon(o) * 0.25f; }
__device__ __forceinline__ V3 ray_origin(V3 p, V3 ng, V3 dir, float error) {
    float offset = fmaxf(error, ray_epsilon(p));
    return p + ng * (dot(dir, ng) >= 0.0f ? offset : -offset);
}
__device__ __forceinline__ float endpoint_tolerance(V3 origin, V3 endpoint, float d) {
    float legacy = 2.0f * ray_epsilon(endpoint);
    return fmaxf(legacy, 1.907349e-6f * (max_abs(origin) + max_abs(endpoint) + d));
}

// ---------------------------------------------------------------- spectral (main.swift:795-944)
struct Spectral {
    const float *icdf;
    const float4 *cmf;
    const float4 *basis;
};
__device__ __forceinline__ float spectral_lambda(const Spectral &sp, float u, int k, float &weight) {
    float x = (u + (float)k) * 0.25f * (float)kIcdfSegments;
    uint32_t i = min((uint32_t)x, (uint32_t)kIcdfSegments - 1u);
    float low = sp.icdf[i], width = sp.icdf[i + 1] - low;
    weight = 0.25f * (float)kIcdfSegments * width;
    return low + (x - (float)i) * width;
}
__device__ __forceinline__ uint32_t spectral_bin(float lambda) {
    return min((uint32_t)fminf(fmaxf(lambda - 360.0f + 0.5f, 0.0f), (float)(kSamples - 1)), (uint32_t)kSamples - 1u);
}
struct Lanes {
    uint32_t bin[4];
    float weight[4];
};
__device__ __forceinline__ Lanes lanes_of(const Spectral &sp, float u) {
    Lanes l;
    for (int k = 0; k < 4; ++k) l.bin[k] = spectral_bin(spectral_lambda(sp, u, k, l.weight[k]));
    return l;
}
// spectrum_rgb: sum_k s_k / (4 p(lambda_k)) * rgb CMF(lambda_k).
__device__ __forceinline__ V3 spectrum_rgb(const Spectral &sp, const Lanes &l, S4 s) {
    V3 rgb = v3(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < 4; ++k) {
        float4 c = sp.cmf[l.bin[k]];
        float w = s.v[k] * l.weight[k];
        rgb = rgb + v3(c.x, c.y, c.z) * w;
    }
    return rgb;
}
// spectral_from_lagrange (main.swift:898-903): bounded reflectance at each lane.
__device__ __forceinline__ S4 reflectance(const Spectral &sp, const Lanes &l, float4 L) {
    if (L.w == 2.0f) return s4(L.x);
    S4 r;
    for (int k = 0; k < 4; ++k) {
        float x = sp.cmf[l.bin[k]].w;
        float series = L.x + 2.0f * L.y * x + 2.0f * L.z * (2.0f * x * x - 1.0f);
        r.v[k] = atanf(series) * 0.318309886f + 0.5f;
    }
    return r;
}
// spectral_emission with the D65 illuminant (main.swift:918-944): white + secondary + primary.
__device__ __forceinline__ S4 emission_spectrum(const Spectral &sp, const Lanes &l, V3 e) {
    e = v3(fmaxf(e.x, 0.0f), fmaxf(e.y, 0.0f), fmaxf(e.z, 0.0f));
    float hi = fmaxf(e.x, fmaxf(e.y, e.z));
    if (!(hi > 0.0f)) return s4(0.0f);
    float lo = fminf(e.x, fminf(e.y, e.z)), mid = e.x + e.y + e.z - hi - lo;
    uint32_t pi = e.x == hi ? 0u : (e.y == hi ? 1u : 2u);
    uint32_t si = pi != 2u && e.z == lo ? 2u : (pi != 1u && e.y == lo ? 1u : 0u);
    float ws = fmaxf(mid - lo, 0.0f), wp = fmaxf(hi - mid, 0.0f);
    S4 r;
    for (int k = 0; k < 4; ++k) {
        float4 a = sp.basis[2 * l.bin[k]], b = sp.basis[2 * l.bin[k] + 1];
        float prim = pi == 0u ? a.x : (pi == 1u ? a.y : a.z);
        float sec = si == 0u ? b.x : (si == 1u ? b.y : b.z);
        r.v[k] = (lo + ws * sec + wp * prim) * a.w;
    }
    return r;
}

// ---------------------------------------------------------------- hit attributes (main.swift:1587-1602, 2000-2014)
struct Surface {
    V3 p, n, ng;
    float error;
    bool front;
    uint32_t slot, face;
};
__device__ __forceinline__ void face_corners(const RenderScene &sc, uint32_t face, V3 &a, V3 &b, V3 &c, uint32_t &tag,
                                             const float **va, const float **vb, const float **vc) {
    const uint32_t *f = sc.faces + 4 * (size_t)face;
    *va = sc.vertices + 8 * (size_t)f[0];
    *vb = sc.vertices + 8 * (size_t)f[1];
    *vc = sc.vertices + 8 * (size_t)f[2];
    a = load_v3(*va);
    b = load_v3(*vb);
    c = load_v3(*vc);
    tag = f[3];
}
__device__ __forceinline__ Surface surface_at(const RenderScene &sc, uint32_t face, float b1, float b2, V3 dir) {
    V3 a, b, c;
    uint32_t tag;
    const float *va, *vb, *vc;
    face_corners(sc, face, a, b, c, tag, &va, &vb, &vc);
    Surface s;
    float b0 = 1.0f - b1 - b2;
    s.p = a * b0 + b * b1 + c * b2;
    V3 ng = normalize(cross(b - a, c - a));
    s.front = dot(ng, dir) < 0.0f;
    s.ng = s.front ? ng : -ng;
    V3 smooth = load_v3(va + 3) * b0 + load_v3(vb + 3) * b1 + load_v3(vc + 3) * b2;
    s.n = dot(smooth, smooth) > 1e-12f ? normalize(smooth) : s.ng;
    if (dot(s.n, s.ng) < 0.0f) s.n = -s.n;
    s.error = fmaxf(1e-7f, 1.907349e-6f * fmaxf(max_abs(a), fmaxf(max_abs(b), max_abs(c))));
    s.slot = tag;
    s.face = face;
    return s;
}
// The hit's tangent dP/du made perpendicular to the shading normal (mesh_hit_attributes,
// main.swift:1594-1599); read only by anisotropic lobes.
__device__ __forceinline__ V3 surface_tangent(const RenderScene &sc, const Surface &s) {
    V3 a, b, c;
    uint32_t tag;
    const float *va, *vb, *vc;
    face_corners(sc, s.face, a, b, c, tag, &va, &vb, &vc);
    const V3 e1 = b - a, e2 = c - a;
    const float d1x = vb[6] - va[6], d1y = vb[7] - va[7], d2x = vc[6] - va[6], d2y = vc[7] - va[7];
    const float det = d1x * d2y - d1y * d2x;
    V3 t = fabsf(det) > 1e-8f ? (e1 * d2y - e2 * d1y) * (1.0f / det) : e1;
    t = t - s.n * dot(t, s.n);
    if (dot(t, t) < 1e-12f) t = cross(s.n, fabsf(s.n.y) < 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f));
    return normalize(t);
}
struct Mat {
    uint32_t type;
    V3 albedo;
    float4 lagrange;
    V3 emission;
    float weight;
    float roughness, ior, metalness, transmission, coat, anisotropy, fuzz;
};
__device__ __forceinline__ Mat material(const RenderScene &sc, uint32_t slot) {
    Mat m;
    if (slot >= sc.material_count) {
        m.type = kDiffuse;
        m.albedo = v3(0.7f, 0.7f, 0.7f);
        m.lagrange = make_float4(0.7f, 0.0f, 0.0f, 2.0f);
        m.emission = v3(0.0f, 0.0f, 0.0f);
        m.weight = 0.0f;
        m.roughness = 0.5f;
        m.ior = 1.5f;
        m.metalness = 0.0f;
        m.transmission = 0.0f;
        m.coat = 0.0f;
        m.anisotropy = 0.0f;
        m.fuzz = 0.0f;
        return m;
    }
    const float4 *r = reinterpret_cast<const float4 *>(sc.materials) + (kMaterialFloats / 4) * (size_t)slot;
    float4 r0 = r[0], r2 = r[2], r3 = r[3];
    m.roughness = r3.x;
    m.ior = r3.y;
    m.metalness = r3.z;
    m.transmission = r3.w;
    const float4 r4 = r[4];
    m.coat = r4.x;
    m.anisotropy = r4.y;
    m.fuzz = r4.z;
    m.type = __float_as_uint(r0.x);
    m.albedo = v3(r0.y, r0.z, r0.w);
    m.lagrange = r[1];
    m.emission = v3(r2.x, r2.y, r2.z);
    m.weight = r2.w;
    if (m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f) m.type = kEmissive;
    return m;
}

// ---------------------------------------------------------------- MIS helpers (main.swift:2454-2476, 3462-3473)
__device__ __forceinline__ float power_heuristic(float pf, float pg) {
    float s = fmaxf(pf, pg);
    if (s <= 0.0f) return 0.0f;
    float f = pf / s, g = pg / s;
    return f * f / (f * f + g * g);
}
__device__ __forceinline__ float emission_weight(bool prev_delta, bool prev_nee, bool prev_mis, float bsdf_pdf,
                                                 float light_pdf) {
    if (prev_delta || !prev_nee || light_pdf <= 0.0f) return 1.0f;
    return prev_mis ? power_heuristic(bsdf_pdf, light_pdf) : 0.0f;
}
__device__ __forceinline__ V3 sample_cosine_hemisphere(V3 n, float r0, float r1) {
    float phi = kTwoPi * r0;
    float cos_th = sqrtf(r1 + (0.5f / 16777216.0f));
    float sin_th = sqrtf(fmaxf(0.0f, 1.0f - cos_th * cos_th));
    V3 up = fabsf(n.z) < 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
    V3 u = normalize(cross(up, n));
    V3 v = cross(n, u);
    return normalize(u * (cosf(phi) * sin_th) + v * (sinf(phi) * sin_th) + n * cos_th);
}

#include "openpbr_device.h"

// is_delta (main.swift:3285-3287).
__device__ __forceinline__ bool is_delta(const Mat &m) { return m.type == kDielectric || (m.type == kGlossy && m.roughness < 0.02f); }

// eval_bsdf_with_pdf (main.swift:2659-2691): f (not cosine-weighted) at the path's lanes, and the pdf.
// kLayered = false compiles the Lambert branch alone (scenes without GLOSSY / OPENPBR slots).
template <bool kLayered>
__device__ __forceinline__ S4 eval_bsdf_pdf(const RenderScene &sc, const Spectral &sp, const Lanes &lanes, const Surface &s,
                                            const Mat &m, V3 wo, V3 wi, float &pdf) {
    if (m.type == kDiffuse) {
        pdf = dot(s.n, wo) <= 0.0f ? 0.0f : fmaxf(0.0f, dot(s.n, wi)) / kPi;
        float cosine = fabsf(dot(s.n, wi));
        bool ok = dot(wi, s.ng) > 0.0f && cosine >= 1e-7f && dot(s.n, wo) > 0.0f && dot(s.n, wi) > 0.0f;
        return ok ? scale(reflectance(sp, lanes, m.lagrange), 1.0f / kPi) : s4(0.0f);
    }
    pdf = 0.0f;
    if constexpr (!kLayered) return s4(0.0f);
    float cosine = fabsf(dot(s.n, wi));
    if (cosine < 1e-7f || dot(s.n, wo) <= 0.0f) return s4(0.0f);
    const EnergyTables E{sc.energy_metal, sc.energy_luts};
    Pbr p;
    pbr_prepare(p, E, m.type, m.roughness, m.ior, m.metalness, m.transmission, m.coat, m.anisotropy,
                m.anisotropy > 0.0f ? surface_tangent(sc, s) : v3(0.0f, 0.0f, 0.0f), m.fuzz, s.n, wo, !s.front,
                reflectance(sp, lanes, m.lagrange));
    pdf = pbr_pdf(p, E, wi);
    if (!(m.transmission > 0.0f) && dot(wi, s.ng) <= 0.0f) return s4(0.0f);
    return scale(pbr_eval(p, E, wi), 1.0f / cosine);
}

// sample_bsdf (main.swift:3304-3368): direction, weight = f |cos| / pdf per lane, pdf (0 for the delta
// branches); false ends the path.
template <bool kLayered, class R>
__device__ __forceinline__ bool sample_bsdf(const RenderScene &sc, const Spectral &sp, const Lanes &lanes, const Surface &s,
                                           const Mat &m, V3 in, R &seed, V3 &dir, S4 &weight, float &pdf) {
    pdf = 0.0f;
    if (m.type == kDielectric) {
        const float ior = m.ior;
        const S4 tint = reflectance(sp, lanes, m.lagrange);
        const float eta = s.front ? 1.0f / ior : ior;
        const float r0 = (1.0f - ior) / (1.0f + ior);
        for (int attempt = 0; attempt < 2; ++attempt) {
            const V3 mm = attempt == 0 ? s.n : s.ng;
            const float ci = fminf(fmaxf(dot(-in, mm), 0.0f), 1.0f);
            const float s2t = eta * eta * (1.0f - ci * ci);
            const float c = eta > 1.0f && s2t < 1.0f ? sqrtf(1.0f - s2t) : ci;
            const float omc = 1.0f - c;
            const float fr = r0 * r0 + (1.0f - r0 * r0) * (omc * omc * omc * omc * omc);
            const bool reflected = s2t >= 1.0f || rand_f(seed) < fr;
            if (reflected) {
                dir = normalize(in - mm * (2.0f * dot(mm, in)));
                weight = tint;
            } else {
                dir = normalize(in * eta + mm * (eta * ci - sqrtf(1.0f - s2t)));
                weight = scale(tint, eta * eta);
            }
            if ((dot(dir, s.ng) > 0.0f) == reflected) return true;
        }
        return false;
    }
    if (is_delta(m)) {
        const V3 rn = in - s.n * (2.0f * dot(s.n, in));
        const V3 mm = dot(rn, s.ng) > 0.0f ? s.n : s.ng;
        dir = normalize(in - mm * (2.0f * dot(mm, in)));
        const float ci = fminf(fmaxf(dot(-in, mm), 0.0f), 1.0f);
        const S4 f0 = reflectance(sp, lanes, m.lagrange);
        const float omc = 1.0f - ci, p5 = omc * omc * omc * omc * omc;
        for (int k = 0; k < 4; ++k) weight.v[k] = f0.v[k] + (1.0f - f0.v[k]) * p5;
        return dot(dir, s.ng) > 0.0f;
    }
    if (m.type == kDiffuse) {
        float r0, r1, r2;  // r0: the lobe-choice draw the generic sampler keeps (main.swift:3354-3356)
        rand_f3(seed, r0, r1, r2);
        (void)r0;
        dir = sample_cosine_hemisphere(s.n, r1, r2);
        pdf = fmaxf(0.0f, dot(s.n, dir)) / kPi;
        weight = reflectance(sp, lanes, m.lagrange);
        return pdf > 0.0f && dot(dir, s.ng) > 0.0f;
    }
    if constexpr (!kLayered) return false;
    const EnergyTables E{sc.energy_metal, sc.energy_luts};
    Pbr p;
    pbr_prepare(p, E, m.type, m.roughness, m.ior, m.metalness, m.transmission, m.coat, m.anisotropy,
                m.anisotropy > 0.0f ? surface_tangent(sc, s) : v3(0.0f, 0.0f, 0.0f), m.fuzz, s.n, -in, !s.front,
                reflectance(sp, lanes, m.lagrange));
    float r0, r1, r2;
    rand_f3(seed, r0, r1, r2);
    if (!pbr_sample(p, E, r0, r1, r2, dir, weight, pdf)) {
        pdf = 0.0f;
        return false;
    }
    if (!(pdf > 0.0f)) return false;
    if (!(m.transmission > 0.0f) && dot(dir, s.ng) <= 0.0f) return false;
    for (int k = 0; k < 4; ++k)
        if (!isfinite(weight.v[k]) || weight.v[k] < 0.0f) return false;
    return true;
}

// ---------------------------------------------------------------- emitters (main.swift:2720-2724, 2853-2917, 3070-3075)
__device__ __forceinline__ float imported_geometry(V3 p, V3 position, V3 a, V3 b, V3 c) {
    V3 d = position - p;
    float d2 = dot(d, d);
    V3 n = cross(b - a, c - a);
    return d2 > 1e-12f && dot(n, n) > 1e-20f ? fmaxf(0.0f, dot(normalize(n), -d * (1.0f / sqrtf(d2)))) / d2 : 0.0f;
}
// ---------------------------------------------------------------- light BVH walk (main.swift:2743-2844)
struct LtNode {
    V3 lo, hi, w;
    float phi, cos_o, cos_e;
    uint32_t child, parent, count;
};
constexpr uint32_t kNoParent = 0x7ffffffu;
__device__ __forceinline__ LtNode lt_node(const RenderScene &sc, uint32_t i) {
    const uint4 a = reinterpret_cast<const uint4 *>(sc.tree_nodes)[2 * i];
    const uint4 b = reinterpret_cast<const uint4 *>(sc.tree_nodes)[2 * i + 1];
    LtNode n;
    float fx = (float)(a.x & 0xffffu) / 65535.0f * 2.0f - 1.0f, fy = (float)(a.x >> 16) / 65535.0f * 2.0f - 1.0f;
    V3 v = v3(fx, fy, 1.0f - fabsf(fx) - fabsf(fy));
    if (v.z < 0.0f) {
        float nx = (1.0f - fabsf(fy)) * (fx >= 0.0f ? 1.0f : -1.0f), ny = (1.0f - fabsf(fx)) * (fy >= 0.0f ? 1.0f : -1.0f);
        v.x = nx;
        v.y = ny;
    }
    n.w = v * (1.0f / sqrtf(dot(v, v)));
    n.phi = __uint_as_float(a.y);
    n.cos_o = (float)(a.z & 0xffffu) / 65535.0f * 2.0f - 1.0f;
    n.cos_e = (float)(a.z >> 16) / 65535.0f * 2.0f - 1.0f;
    const V3 lo = v3(sc.tree_lo[0], sc.tree_lo[1], sc.tree_lo[2]);
    n.lo = v3(lo.x + (float)(a.w & 0xffffu) * sc.tree_scale[0], lo.y + (float)(a.w >> 16) * sc.tree_scale[1],
              lo.z + (float)(b.x & 0xffffu) * sc.tree_scale[2]);
    n.hi = v3(lo.x + (float)(b.x >> 16) * sc.tree_scale[0], lo.y + (float)(b.y & 0xffffu) * sc.tree_scale[1],
              lo.z + (float)(b.y >> 16) * sc.tree_scale[2]);
    n.child = b.z;
    n.parent = b.w & kNoParent;
    n.count = b.w >> 27;
    return n;
}
__device__ __forceinline__ float lt_power(const RenderScene &sc, uint32_t position) {
    return __uint_as_float(sc.tree_leaves[3 * (size_t)position + 1]);
}
__device__ __forceinline__ float lt_leaf_sum(const RenderScene &sc, const LtNode &n) {
    float s = 0.0f;
    for (uint32_t i = 0; i < n.count; ++i) s += lt_power(sc, n.child + i);
    return s;
}
__device__ __forceinline__ float cos_sub(float sa, float ca, float sb, float cb) { return ca > cb ? 1.0f : ca * cb + sa * sb; }
__device__ __forceinline__ float sin_sub(float sa, float ca, float sb, float cb) { return ca > cb ? 0.0f : sa * cb - ca * sb; }
__device__ __forceinline__ float lt_importance(const LtNode &n, V3 p) {
    if (!(n.phi > 0.0f)) return 0.0f;
    V3 pc = (n.lo + n.hi) * 0.5f;
    V3 dpc = p - pc;
    float d2 = fmaxf(dot(dpc, dpc), fmaxf(0.5f * length(n.hi - n.lo), 1e-20f));
    if (p.x >= n.lo.x && p.y >= n.lo.y && p.z >= n.lo.z && p.x <= n.hi.x && p.y <= n.hi.y && p.z <= n.hi.z)
        return n.phi / d2;
    V3 wi = dpc * (1.0f / sqrtf(dot(dpc, dpc)));
    float cw = dot(n.w, wi), sw = sqrtf(fmaxf(0.0f, 1.0f - cw * cw));
    V3 dh = pc - n.hi;
    float s2 = dot(dh, dh) / dot(dpc, dpc);
    float cb = s2 >= 1.0f ? -1.0f : sqrtf(fmaxf(0.0f, 1.0f - s2)), sb = sqrtf(fmaxf(0.0f, 1.0f - cb * cb));
    float so = sqrtf(fmaxf(0.0f, 1.0f - n.cos_o * n.cos_o));
    float cx = cos_sub(sw, cw, so, n.cos_o), sx = sin_sub(sw, cw, so, n.cos_o);
    float cp = cos_sub(sx, cx, sb, cb);
    return cp <= n.cos_e ? 0.0f : n.phi * cp / d2;
}
// A light-tree node decoded once per scene (decode_light_tree_kernel, lt_node's arithmetic) into the
// quantities lt_importance derives from it: the walk then spends no instructions on dequantisation or
// the octahedral normal.  Stored as 64 bytes, one cache line each, in four 16-byte words:
// (lo, phi), (hi, m), (w, cos_o), (cos_e, child, links, dh2); the centre and the cone's sine are
// re-derived with lt_node's own formulas, which costs fewer cycles than a fifth load per node.
struct LtDec {
    V3 pc;
    float m;  // max(0.5 |hi - lo|, 1e-20)
    V3 lo;
    float phi;
    V3 hi;
    float dh2;  // |pc - hi|^2
    V3 w;
    float cos_o, so, cos_e;
    uint32_t child, links;  // links: parent | leaf count << 27, as the node's word
};
constexpr int kLtDecWords = 4;
__device__ __forceinline__ LtDec lt_dec(const RenderScene &sc, uint32_t i) {
    const float4 *q = sc.tree_dec + kLtDecWords * (size_t)i;
    const float4 a = q[0], b = q[1], c = q[2], d = q[3];
    LtDec n;
    n.lo = v3(a.x, a.y, a.z);
    n.phi = a.w;
    n.hi = v3(b.x, b.y, b.z);
    n.m = b.w;
    n.w = v3(c.x, c.y, c.z);
    n.cos_o = c.w;
    n.cos_e = d.x;
    n.child = __float_as_uint(d.y);
    n.links = __float_as_uint(d.z);
    n.dh2 = d.w;
    n.pc = (n.lo + n.hi) * 0.5f;
    n.so = sqrtf(fmaxf(0.0f, 1.0f - n.cos_o * n.cos_o));
    return n;
}
// The word holding a node's child and links (x: cos_e, y: child, z: links).
__device__ __forceinline__ float4 lt_links(const RenderScene &sc, uint32_t i) {
    return sc.tree_dec[kLtDecWords * (size_t)i + 3];
}
// lt_importance on a decoded node (the same operations in the same order).
__device__ __forceinline__ float lt_importance(const LtDec &n, V3 p) {
    if (!(n.phi > 0.0f)) return 0.0f;
    const V3 dpc = p - n.pc;
    const float dd = dot(dpc, dpc);
    const float d2 = fmaxf(dd, n.m);
    if (p.x >= n.lo.x && p.y >= n.lo.y && p.z >= n.lo.z && p.x <= n.hi.x && p.y <= n.hi.y && p.z <= n.hi.z)
        return n.phi / d2;
    const V3 wi = dpc * (1.0f / sqrtf(dd));
    const float cw = dot(n.w, wi), sw = sqrtf(fmaxf(0.0f, 1.0f - cw * cw));
    const float s2 = n.dh2 / dd;
    const float cb = s2 >= 1.0f ? -1.0f : sqrtf(fmaxf(0.0f, 1.0f - s2)), sb = sqrtf(fmaxf(0.0f, 1.0f - cb * cb));
    const float cx = cos_sub(sw, cw, n.so, n.cos_o), sx = sin_sub(sw, cw, n.so, n.cos_o);
    const float cp = cos_sub(sx, cx, sb, cb);
    return cp <= n.cos_e ? 0.0f : n.phi * cp / d2;
}
__device__ __forceinline__ float lt_leaf_sum(const RenderScene &sc, uint32_t child, uint32_t count) {
    float s = 0.0f;
    for (uint32_t i = 0; i < count; ++i) s += lt_power(sc, child + i);
    return s;
}

// The light-tree walk (LightTree sampling, main.swift's imported light tree) from node `index` at
// `depth` with first child `child` and leaf count `count`, u the remaining uniform and pmf the
// probability of the path so far.  The chosen child's links (first child, leaf count) are carried
// down, so a level costs one round of two independent node loads (the children) instead of two
// dependent rounds.
__device__ __forceinline__ uint32_t lt_walk(const RenderScene &sc, V3 p, float u, float &pmf, uint32_t index,
                                            uint32_t child, uint32_t count, uint32_t depth) {
    for (; depth <= 64u; ++depth) {
        if (count != 0u) {
            float sum = lt_leaf_sum(sc, child, count), target = u * sum, cum = 0.0f, chosen = 0.0f;
            uint32_t position = 0xffffffffu;
            for (uint32_t i = 0; i < count && sum > 0.0f; ++i) {
                float power = lt_power(sc, child + i);
                if (!(power > 0.0f)) continue;
                position = child + i;
                chosen = power;
                cum += power;
                if (target < cum) break;
            }
            if (position == 0xffffffffu) break;
            pmf *= chosen / sum;
            return position;
        }
        const LtDec left = lt_dec(sc, index + 1u), right = lt_dec(sc, child);
        float l = lt_importance(left, p), r = lt_importance(right, p);
        if (!(l + r > 0.0f)) break;
        float pl = l / (l + r);
        if (u < pl) {
            u = fminf(u / pl, 0x1.fffffep-1f);
            index = index + 1u;
            pmf *= pl;
            child = left.child;
            count = left.links >> 27;
        } else {
            u = fminf((u - pl) / (1.0f - pl), 0x1.fffffep-1f);
            index = child;
            pmf *= r / (l + r);
            child = right.child;
            count = right.links >> 27;
        }
    }
    pmf = 0.0f;
    return 0xffffffffu;
}
__device__ __forceinline__ uint32_t lt_sample(const RenderScene &sc, V3 p, float u, float &pmf) {
    pmf = 1.0f;
    const float4 root = lt_links(sc, 0);
    return lt_walk(sc, p, u, pmf, 0u, __float_as_uint(root.y), __float_as_uint(root.z) >> 27, 0u);
}

// The light tree's top kLtTopLevels levels for one shading point.  Every candidate of a pixel walks the
// tree from the root at the same point, so the importances of the children of the top nodes are the
// same for all of them: they are computed once per pixel (2 x 15 evaluations instead of 2 per level per
// candidate, 256 for 32 candidates) and kept in LDS, [node][thread].  The block shares the top nodes'
// indices and links (heap order: node h has children 2h + 1 and 2h + 2; kNoNode below a leaf).  The
// walk through the cached levels performs the same operations on the same values as lt_walk, so the
// samples are identical.
constexpr int kLtTopLevels = 4;
constexpr int kLtTopInner = (1 << kLtTopLevels) - 1;       // heap nodes whose children are cached
constexpr int kLtTopAll = (1 << (kLtTopLevels + 1)) - 1;   // heap nodes with their links in the table
constexpr uint32_t kNoNode = 0xffffffffu;
struct LtTop {
    uint32_t index[kLtTopAll], child[kLtTopAll], count[kLtTopAll];
};
// Heap node h of the tree's top (one thread per h < kLtTopAll): follows h's path from the root.
__device__ __forceinline__ void lt_top_fill(const RenderScene &sc, LtTop &top, uint32_t h) {
    uint32_t path = 0, depth = 0;
    for (uint32_t g = h; g != 0u; g = (g - 1u) >> 1) {
        path = (path << 1) | ((g & 1u) ? 0u : 1u);  // 0: left child (odd heap index), 1: right
        ++depth;
    }
    uint32_t index = 0;
    float4 e = lt_links(sc, 0);
    for (uint32_t d = 0; d < depth; ++d, path >>= 1) {
        if ((__float_as_uint(e.z) >> 27) != 0u) {  // a leaf has no children
            index = kNoNode;
            break;
        }
        index = (path & 1u) ? __float_as_uint(e.y) : index + 1u;
        e = lt_links(sc, index);
    }
    top.index[h] = index;
    top.child[h] = index == kNoNode ? 0u : __float_as_uint(e.y);
    top.count[h] = index == kNoNode ? 0u : __float_as_uint(e.z) >> 27;
}
template <int kBlk>
__device__ __forceinline__ void lt_top_importances(const RenderScene &sc, const LtTop &top, V3 p,
                                                   float (*tl)[kBlk], float (*tr)[kBlk]) {
    for (int h = 0; h < kLtTopInner; ++h) {
        if (top.index[h] == kNoNode || top.count[h] != 0u) continue;
        tl[h][threadIdx.x] = lt_importance(lt_dec(sc, top.index[h] + 1u), p);
        tr[h][threadIdx.x] = lt_importance(lt_dec(sc, top.child[h]), p);
    }
}
template <int kBlk>
__device__ __forceinline__ uint32_t lt_sample_top(const RenderScene &sc, V3 p, float u, float &pmf, const LtTop &top,
                                                  const float (*tl)[kBlk], const float (*tr)[kBlk]) {
    pmf = 1.0f;
    uint32_t h = 0, child = top.child[0], count = top.count[0], depth = 0;
    for (; depth < (uint32_t)kLtTopLevels; ++depth) {
        if (count != 0u) break;
        const float l = tl[h][threadIdx.x], r = tr[h][threadIdx.x];
        if (!(l + r > 0.0f)) {
            pmf = 0.0f;
            return 0xffffffffu;
        }
        const float pl = l / (l + r);
        if (u < pl) {
            u = fminf(u / pl, 0x1.fffffep-1f);
            pmf *= pl;
            h = 2u * h + 1u;
        } else {
            u = fminf((u - pl) / (1.0f - pl), 0x1.fffffep-1f);
            pmf *= r / (l + r);
            h = 2u * h + 2u;
        }
        child = top.child[h];
        count = top.count[h];
    }
    return lt_walk(sc, p, u, pmf, top.index[h], child, count, depth);
}
__device__ __forceinline__ float lt_pmf(const RenderScene &sc, V3 p, uint32_t position) {
    uint32_t index = sc.tree_leaves[3 * (size_t)position + 2];
    LtDec node = lt_dec(sc, index);
    float sum = lt_leaf_sum(sc, node.child, node.links >> 27);
    float pmf = sum > 0.0f ? lt_power(sc, position) / sum : 0.0f;
    for (uint32_t depth = 0; depth <= 64u && pmf > 0.0f; ++depth) {
        uint32_t parent = node.links & kNoParent;
        if (parent == kNoParent) return pmf;
        const LtDec up = lt_dec(sc, parent);
        float l = lt_importance(lt_dec(sc, parent + 1u), p), r = lt_importance(lt_dec(sc, up.child), p);
        float mine = index == parent + 1u ? l : r;
        if (!(l + r > 0.0f) || !(mine > 0.0f)) return 0.0f;
        pmf *= mine / (l + r);
        index = parent;
        node = up;
    }
    return 0.0f;
}

// light_visible's shadow ray (main.swift:3254-3264) from surface s toward a light sample: to the point,
// stopped short by the endpoint tolerance, or toward the environment (light 1) along wi for 1e6 less
// twice the ray epsilon.  False when the sample is too close to need one (it is then not visible).
__device__ __forceinline__ bool shadow_ray(const Surface &s, uint32_t light, V3 wi, V3 position, V3 &o, V3 &d,
                                           float &tmin, float &tmax) {
    o = ray_origin(s.p, s.ng, wi, s.error);
    V3 delta;
    float dist, tol;
    if (light == 1u) {
        delta = wi;
        dist = 1e6f;
        tol = 2.0f * ray_epsilon(s.p);
    } else {
        delta = position - o;
        dist = length(delta);
        tol = endpoint_tolerance(o, position, dist);
    }
    if (!(dist > tol)) return false;
    d = normalize(delta);
    tmin = ray_t_min(o);
    tmax = dist - tol;
    return true;
}

struct LightSample {
    V3 position, wi, emission;
    float dist, pdf, geometry;
    uint32_t slot, light;  // light: face + 2 (LightSample.isDirectional), 1 = the environment, 0 = none
};

// ---------------------------------------------------------------- environment map (main.swift:2077-2135, 2700-2719, 2880-2949)
// An image environment without a sun, in an imported scene: lat-long lookup with bilinear filtering,
// image sampling from the two tables (kEnvCompensated where a BSDF continuation covers the texels the
// MIS-compensated table leaves out, kEnvOriginal elsewhere), and the split of light samples between
// the triangle emitters and the environment.
constexpr uint32_t kEnvCompensated = 0u, kEnvOriginal = 1u;
// a / b correctly rounded.  This file is compiled with the hardware's approximate fp32 division; the
// environment's lookups divide in double (exact once rounded back: 53 >= 2 x 24 + 2 bits) so that a
// coordinate lands in the texel the host's computation of the same lookup picks.
__device__ __forceinline__ float div_rn(float a, float b) { return (float)((double)a / (double)b); }
__device__ __forceinline__ bool env_on(const RenderScene &sc, const FrameArgs &f) {
    return sc.env != nullptr && f.env_strength > 0.0f;
}
// The frame a kernel shades with: with kEnv false the environment is off at compile time, so every
// environment branch folds away and kernels for frames without a map keep their register budget.
template <bool kEnv>
__device__ __forceinline__ FrameArgs frame_env(const FrameArgs &f) {
    FrameArgs r = f;
    if (!kEnv) r.env_strength = 0.0f;
    return r;
}
__host__ __forceinline__ bool frame_has_env(const RenderScene &sc, const FrameArgs &f) {
    return sc.env != nullptr && f.env_strength > 0.0f;
}
// imported_light_probability: the chance that a light sample takes the emitters.
__device__ __forceinline__ float imported_probability(const RenderScene &sc, const FrameArgs &f) {
    if (sc.emitter_count == 0) return 0.0f;
    return env_on(sc, f) ? 0.5f : 1.0f;
}
// eval_environment: bilinear between texel centres (wrapping in longitude, latitude clamped half a texel
// inside the map), times the strength.
__device__ __forceinline__ V3 eval_environment(const RenderScene &sc, const FrameArgs &f, V3 d) {
    if (!env_on(sc, f)) return v3(0.0f, 0.0f, 0.0f);
    const float w = (float)sc.env_w, h = (float)sc.env_h;
    float u = div_rn(atan2f(d.z, d.x), kTwoPi) + 0.5f + div_rn(f.env_rotation, kTwoPi);
    // Repeat addressing: only the fraction matters (kept in [0, 1], so texel indices stay in range).
    u = fminf(fmaxf(u - floorf(u), 0.0f), 1.0f);
    float v = div_rn(acosf(fminf(fmaxf(d.y, -1.0f), 1.0f)), kPi);
    v = fminf(fmaxf(v, div_rn(0.5f, h)), 1.0f - div_rn(0.5f, h));
    // Products and sums rounded one by one (no fused multiply-add): next to the sun, texels differ by
    // orders of magnitude, so a fused rounding of a filter weight shows in the image.  The row
    // coordinate is kept inside the map (the clamp above in texel units, whatever its rounding), so a
    // lookup at a pole reads the pole row alone.
    const float fx = __fmul_rn(u, w) - 0.5f, fy = fminf(fmaxf(__fmul_rn(v, h) - 0.5f, 0.0f), h - 1.0f);
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float ax = fx - x0f, ay = fy - y0f;
    int xa = (int)x0f % (int)sc.env_w;
    if (xa < 0) xa += (int)sc.env_w;
    const int xb = xa + 1 == (int)sc.env_w ? 0 : xa + 1;
    const int ya = min(max((int)y0f, 0), (int)sc.env_h - 1), yb = min(ya + 1, (int)sc.env_h - 1);
    const float4 t00 = sc.env[(size_t)ya * sc.env_w + xa], t10 = sc.env[(size_t)ya * sc.env_w + xb];
    const float4 t01 = sc.env[(size_t)yb * sc.env_w + xa], t11 = sc.env[(size_t)yb * sc.env_w + xb];
    const float bx = 1.0f - ax, by = 1.0f - ay;
    auto lerp = [](float a, float wa, float b, float wb) { return __fadd_rn(__fmul_rn(a, wa), __fmul_rn(b, wb)); };
    const V3 top = v3(lerp(t00.x, bx, t10.x, ax), lerp(t00.y, bx, t10.y, ax), lerp(t00.z, bx, t10.z, ax));
    const V3 bot = v3(lerp(t01.x, bx, t11.x, ax), lerp(t01.y, bx, t11.y, ax), lerp(t01.z, bx, t11.z, ax));
    return v3(lerp(top.x, by, bot.x, ay), lerp(top.y, by, bot.y, ay), lerp(top.z, by, bot.z, ay)) * f.env_strength;
}
// environment_cell_pdf: the solid-angle density of texel (x, y) in `table`.
__device__ __forceinline__ float env_cell_pdf(const RenderScene &sc, uint32_t x, uint32_t y, float sin_theta,
                                              uint32_t table) {
    const uint32_t cond = y + table * sc.env_h;
    const float row = sc.env_rows[2 * y + table], prev_row = y == 0 ? 0.0f : sc.env_rows[2 * (y - 1) + table];
    const float *c = sc.env_cols + (size_t)cond * sc.env_w;
    const float col = c[x], prev_col = x == 0 ? 0.0f : c[x - 1];
    const float p = fmaxf(0.0f, row - prev_row) * fmaxf(0.0f, col - prev_col);
    return div_rn(p * (float)(sc.env_w * sc.env_h), 2.0f * kPi * kPi * fmaxf(1e-5f, sin_theta));
}
// eval_environment_pdf: the density of a light sample's environment proposal (importedProbability applied).
__device__ __forceinline__ float env_pdf(const RenderScene &sc, const FrameArgs &f, V3 wi, uint32_t table) {
    if (!env_on(sc, f)) return 0.0f;
    const float theta = acosf(fminf(fmaxf(wi.y, -1.0f), 1.0f));
    float u = div_rn(atan2f(wi.z, wi.x), kTwoPi) + 0.5f + div_rn(f.env_rotation, kTwoPi);
    u = fminf(u - floorf(u), 0x1.fffffep-1f);
    const float v = fminf(fmaxf(div_rn(theta, kPi), 0.0f), 1.0f - 1e-7f);
    const uint32_t x = min((uint32_t)(u * (float)sc.env_w), sc.env_w - 1u);
    const uint32_t y = min((uint32_t)(v * (float)sc.env_h), sc.env_h - 1u);
    return (1.0f - imported_probability(sc, f)) * env_cell_pdf(sc, x, y, sinf(theta), table);
}
// environment_cdf_index: the first of `count` CDF entries (entry i at cdf[i * step]) >= value.
__device__ __forceinline__ uint32_t env_cdf_index(const float *cdf, uint32_t count, uint32_t step, float value) {
    uint32_t low = 0, high = max(count, 1u) - 1u;
    for (uint32_t it = 0; it < 16u && low < high; ++it) {
        const uint32_t mid = (low + high) / 2;
        if (cdf[(size_t)mid * step] >= value) high = mid;
        else low = mid + 1;
    }
    return low;
}
// The largest float below x (> 0): nextafter(x, 0).
__device__ __forceinline__ float float_below(float x) { return __uint_as_float(__float_as_uint(x) - 1u); }
// The environment branch of sample_direct_light (4 draws; two 2D constituents in Z mode).
template <class R>
__device__ __forceinline__ LightSample sample_environment(const RenderScene &sc, const FrameArgs &f, V3 p, R &seed,
                                                          uint32_t table) {
    const uint32_t w = sc.env_w, h = sc.env_h;
    float cx, cy, ox, oy;
    rand_f2(seed, cx, cy);
    const uint32_t y = env_cdf_index(sc.env_rows + table, h, 2u, cx);
    const uint32_t x = env_cdf_index(sc.env_cols + (size_t)(y + table * h) * w, w, 1u, cy);
    rand_f2(seed, ox, oy);
    // The in-texel offset stays inside the chosen cell, whose density is the sample's.
    const float fw = (float)w, fh = (float)h;
    const float ux = fminf(div_rn((float)x, fw) + div_rn(ox, fw), float_below(div_rn((float)x + 1.0f, fw)));
    const float uy = fminf(div_rn((float)y, fh) + div_rn(oy, fh), float_below(div_rn((float)y + 1.0f, fh)));
    const float theta = uy * kPi, phi = (ux - 0.5f - div_rn(f.env_rotation, kTwoPi)) * kTwoPi;
    const float st = sinf(theta);
    LightSample ls;
    ls.wi = normalize(v3(st * cosf(phi), cosf(theta), st * sinf(phi)));
    ls.position = p + ls.wi * 1e6f;
    ls.dist = 1e6f;
    ls.light = 1u;
    ls.slot = 0u;
    ls.geometry = 1.0f;
    ls.pdf = (1.0f - imported_probability(sc, f)) * env_cell_pdf(sc, x, y, st, table);
    ls.emission = eval_environment(sc, f, ls.wi);
    return ls;
}

// walk(u, pmf): the light-tree walk for uniform u (lt_sample, or lt_sample_top with a pixel's cache).
template <class Walk, class R>
__device__ __forceinline__ LightSample sample_emitter_via(const RenderScene &sc, const FrameArgs &f, V3 p, R &seed,
                                                          uint32_t table, const Walk &walk) {
    LightSample ls;
    ls.pdf = 0.0f;
    ls.slot = 0;
    ls.light = 0;
    ls.geometry = 0.0f;
    ls.emission = v3(0.0f, 0.0f, 0.0f);
    ls.position = p;
    ls.wi = v3(0.0f, 0.0f, 1.0f);
    ls.dist = 0.0f;
    // The emitters with probability importedProbability (its draw is taken whenever it is positive,
    // even at 1), the environment otherwise.
    const float pimp = imported_probability(sc, f);
    if (!(pimp > 0.0f && rand_f(seed) < pimp)) return env_on(sc, f) ? sample_environment(sc, f, p, seed, table) : ls;
    // The emitter choice and the point on it: one 3D constituent in Z mode (main.swift:2886-2894).
    float value, ry, rz;
    rand_f3(seed, value, ry, rz);
    uint32_t face;
    float tree_pmf = 0.0f;
    if (sc.tree_node_count) {
        uint32_t position = walk(value, tree_pmf);
        if (position == 0xffffffffu) return ls;
        face = sc.tree_leaves[3 * (size_t)position];
    } else {
        uint32_t low = 0, high = sc.emitter_count - 1;
        while (low < high) {
            uint32_t mid = (low + high) / 2;
            if (sc.emitter_cdf[mid] >= value) high = mid;
            else low = mid + 1;
        }
        face = sc.emitter_face[low];
    }
    V3 a, b, c;
    uint32_t tag;
    const float *va, *vb, *vc;
    face_corners(sc, face, a, b, c, tag, &va, &vb, &vc);
    float root = sqrtf(ry);
    ls.position = a * (1.0f - root) + b * (root * (1.0f - rz)) + c * (root * rz);
    V3 delta = ls.position - p;
    ls.dist = length(delta);
    ls.wi = delta * (1.0f / fmaxf(ls.dist, 1e-8f));
    float geometry = imported_geometry(p, ls.position, a, b, c);
    Mat m = material(sc, tag);
    ls.emission = m.emission;
    ls.slot = tag;
    ls.light = face + 2u;
    ls.geometry = geometry;
    float area_pdf;
    if (sc.tree_node_count) {
        float area = 0.5f * length(cross(b - a, c - a));
        area_pdf = area > 0.0f ? tree_pmf / area : 0.0f;
    } else {
        area_pdf = sc.emitter_total > 0.0f ? m.weight / sc.emitter_total : 0.0f;
    }
    ls.pdf = geometry > 1e-12f ? pimp * area_pdf / geometry : 0.0f;
    return ls;
}
template <class R>
__device__ __forceinline__ LightSample sample_emitter(const RenderScene &sc, const FrameArgs &f, V3 p, R &seed,
                                                      uint32_t table) {
    return sample_emitter_via(sc, f, p, seed, table, [&](float u, float &pmf) { return lt_sample(sc, p, u, pmf); });
}
// eval_light_pdf of an imported emitter face reached from p (main.swift:3070-3075): importedProbability
// x the area density over the geometry term.
__device__ __forceinline__ float emitter_pdf(const RenderScene &sc, const FrameArgs &f, V3 p, V3 position, uint32_t face,
                                             float weight) {
    if (sc.emitter_count == 0) return 0.0f;
    const float pimp = imported_probability(sc, f);
    V3 a, b, c;
    uint32_t tag;
    const float *va, *vb, *vc;
    face_corners(sc, face, a, b, c, tag, &va, &vb, &vc);
    float g = imported_geometry(p, position, a, b, c);
    if (!(g > 1e-12f)) return 0.0f;
    if (sc.tree_node_count) {
        // imported_emitter_position: binary search of the ascending emitter faces (main.swift:2846-2850).
        uint32_t lo = 0, hi = sc.emitter_count;
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            if (sc.tree_ids[mid] < face) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= sc.emitter_count || sc.tree_ids[lo] != face) return 0.0f;
        float area = 0.5f * length(cross(b - a, c - a));
        return area > 0.0f ? pimp * (lt_pmf(sc, p, sc.tree_pos[lo]) / area) / g : 0.0f;
    }
    if (!(sc.emitter_total > 0.0f) || !(weight > 0.0f)) return 0.0f;
    return pimp * (weight / sc.emitter_total) / g;
}

// ---------------------------------------------------------------- ReSTIR DI (main.swift:3165-3196, 4618-4716, 4812-4835, 5015-5033)
struct DiSample {
    V3 position, emission;
    uint32_t light;  // face + 2, 0 = none
    float pdf, u;
};
__device__ __forceinline__ DiSample di_empty() {
    DiSample d;
    d.position = v3(0.0f, 0.0f, 0.0f);
    d.emission = d.position;
    d.light = 0;
    d.pdf = 0.0f;
    d.u = 0.0f;
    return d;
}
// restir_di_pack_light / unpack (main.swift:4674-4675): id bits offset into the normal floats.
__device__ __forceinline__ float pack_light(uint32_t id) { return id == 0u ? 0.0f : __uint_as_float(id + 0x00800000u); }
__device__ __forceinline__ uint32_t unpack_light(float w) {
    uint32_t bits = __float_as_uint(w);
    return bits == 0u ? 0u : bits - 0x00800000u;
}
__device__ __forceinline__ uint2 pack_half4(float x, float y, float z, float w) {
    uint32_t a = __half_as_ushort(__float2half_rn(x)), b = __half_as_ushort(__float2half_rn(y));
    uint32_t c = __half_as_ushort(__float2half_rn(z)), d = __half_as_ushort(__float2half_rn(w));
    return make_uint2(a | (b << 16), c | (d << 16));
}
__device__ __forceinline__ float4 unpack_half4(uint2 v) {
    return make_float4(__half2float(__ushort_as_half((unsigned short)(v.x & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(v.x >> 16))),
                       __half2float(__ushort_as_half((unsigned short)(v.y & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(v.y >> 16))));
}
// restir_di_stored_sample (main.swift:4679-4688).
__device__ __forceinline__ DiSample di_stored(float4 posdir, float4 emitpdf, float4 weights) {
    DiSample d;
    d.position = v3(posdir.x, posdir.y, posdir.z);
    d.light = unpack_light(posdir.w);
    d.emission = v3(emitpdf.x, emitpdf.y, emitpdf.z);
    d.pdf = emitpdf.w;
    d.u = weights.w;
    return d;
}
// light_geometry of an imported emitter (main.swift:3165-3166).
__device__ __forceinline__ float light_geometry(const RenderScene &sc, V3 p, const DiSample &d) {
    if (d.light == 1u) return 1.0f;  // the environment: a direction
    if (d.light < 2u || d.light - 2u >= sc.face_count) return 0.0f;
    V3 a, b, c;
    uint32_t tag;
    const float *va, *vb, *vc;
    face_corners(sc, d.light - 2u, a, b, c, tag, &va, &vb, &vc);
    return imported_geometry(p, d.position, a, b, c);
}
// eval_bsdf for DIFFUSE (main.swift:2659-2665), given rho/pi at the lanes.
__device__ __forceinline__ S4 diffuse_f(const Surface &s, S4 rho_pi, V3 wo, V3 wi) {
    if (dot(wi, s.ng) <= 0.0f) return s4(0.0f);
    float c = fabsf(dot(s.n, wi));
    if (c < 1e-7f || dot(s.n, wo) <= 0.0f) return s4(0.0f);
    return dot(s.n, wi) > 0.0f ? rho_pi : s4(0.0f);
}
// eval_restir_target_pdf (main.swift:3187-3196): |rgb(f L cos)| G at the sample's wavelengths
// (lanes and rho_pi belong to d.u; g is light_geometry(p, d)).  An environment sample (light 1) holds
// its direction in `position` (restir_di_encode, main.swift:4714).
__device__ __forceinline__ float target_pdf(const Spectral &sp, const Lanes &l, S4 rho_pi, const Surface &s, V3 raydir,
                                            const DiSample &d, float g) {
    if (d.pdf <= 0.0f) return 0.0f;
    V3 dir = d.light == 1u ? d.position : normalize(d.position - s.p);
    float ct = fmaxf(0.0f, dot(s.n, dir));
    if (ct <= 0.0f) return 0.0f;
    S4 f = diffuse_f(s, rho_pi, -raydir, dir);
    return length(spectrum_rgb(sp, l, scale(mul(f, emission_spectrum(sp, l, d.emission)), ct))) * g;
}
// The DI sample's emission as light_spectrum reads it (main.swift:2181-2186): the environment re-reads
// the map in the sample's direction.
__device__ __forceinline__ V3 di_emission(const RenderScene &sc, const FrameArgs &f, const DiSample &d) {
    return d.light == 1u ? eval_environment(sc, f, d.position) : d.emission;
}
__device__ __forceinline__ float target_pdf_at(const RenderScene &sc, const FrameArgs &f, const Spectral &sp, const Mat &m,
                                               const Surface &s, V3 raydir, DiSample d) {
    if (d.pdf <= 0.0f) return 0.0f;
    d.emission = di_emission(sc, f, d);
    const Lanes l = lanes_of(sp, d.u);
    return target_pdf(sp, l, scale(reflectance(sp, l, m.lagrange), 1.0f / kPi), s, raydir, d, light_geometry(sc, s.p, d));
}

// ---------------------------------------------------------------- ReSTIR GI (main.swift:3200-3251)
// gi_geometry: cosines at both ends over the squared distance.
__device__ __forceinline__ float gi_geometry(V3 x1, V3 n1, V3 x2, V3 n2) {
    const V3 delta = x2 - x1;
    const float d2 = dot(delta, delta);
    if (d2 < 1e-10f) return 0.0f;
    const V3 dir = delta * (1.0f / sqrtf(d2));
    return fmaxf(0.0f, dot(n1, dir)) * fmaxf(0.0f, dot(n2, -dir)) / d2;
}
// restir_gi_jacobian / restir_gi_accepts_shift: the reconnection's solid-angle Jacobian in [0.1, 10].
__device__ __forceinline__ bool gi_accepts_shift(V3 x1r, V3 x1q, V3 x2, V3 n2) {
    const V3 cur = x1r - x2, src = x1q - x2;
    const float cd2 = dot(cur, cur), sd2 = dot(src, src);
    if (cd2 < 1e-10f || sd2 < 1e-10f) return false;
    const float cc = fmaxf(0.0f, dot(n2, cur * (1.0f / sqrtf(cd2))));
    const float sco = fmaxf(0.0f, dot(n2, src * (1.0f / sqrtf(sd2))));
    float j = cc * sd2 / (sco * cd2);
    if (!isfinite(j)) j = 0.0f;
    return j >= 0.1f && j <= 10.0f;
}
// eval_restir_gi_target at a diffuse x1: |rgb(f rad G)| at the sample's wavelengths (number u).
__device__ __forceinline__ float gi_target(const Spectral &sp, const Mat &m, const Surface &s, V3 raydir, V3 x2, V3 n2,
                                           S4 rad, float u) {
    const float g = gi_geometry(s.p, s.n, x2, n2);
    if (g <= 0.0f || !(isfinite(rad.v[0]) && isfinite(rad.v[1]) && isfinite(rad.v[2]) && isfinite(rad.v[3]))) return 0.0f;
    const V3 dir = normalize(x2 - s.p);
    const Lanes l = lanes_of(sp, u);
    const S4 f = diffuse_f(s, scale(reflectance(sp, l, m.lagrange), 1.0f / kPi), -raydir, dir);
    return length(spectrum_rgb(sp, l, scale(mul(f, rad), g)));
}
__device__ __forceinline__ S4 s4_of(float4 v) { return S4{{v.x, v.y, v.z, v.w}}; }

// ---------------------------------------------------------------- wave-aggregated queue append
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t lane = __lane_id();
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
    }
    return base + rank;
}

constexpr uint32_t kPrevDelta = 1u << 16, kPrevNee = 1u << 17, kPrevMis = 1u << 18;
// BSDF-only paths (mode 3) past the scattering budget: one last continuation that may only reach
// emission (main.swift:7716-7722, 7785).
constexpr uint32_t kEmissionOnly = 1u << 19;
// Training paths of the radiance cache: the path has recorded its cache vertex.
constexpr uint32_t kRecorded = 1u << 20;
// ReSTIR GI: the path left a diffuse primary whose GI reservoir holds the light of its secondary vertex.
constexpr uint32_t kGiSecondary = 1u << 21;

// ---------------------------------------------------------------- learned radiance cache
// Training paths: 1 in f.train_period, by a hash of pixel and sample index.
__device__ __forceinline__ bool nrc_training(const FrameArgs &f, uint32_t id) {
    const uint32_t smp = id / f.npix, pix = id - smp * f.npix;
    return pcg_hash((pix * 0x9E3779B9u) ^ ((f.sample_index0 + smp) * 0x85EBCA6Bu)) % f.train_period == 0u;
}
// The 11 cache inputs of a vertex: position in the scene box, direction toward the previous vertex,
// shading normal (both mapped to [0, 1]), roughness (1 for Lambert), wavelength number.
__device__ __forceinline__ void nrc_inputs(const FrameArgs &f, const Surface &s, const Mat &m, V3 dir, float wl, float *out) {
    out[0] = fminf(fmaxf((s.p.x - f.box_lo[0]) * f.box_scale[0], 0.0f), 1.0f);
    out[1] = fminf(fmaxf((s.p.y - f.box_lo[1]) * f.box_scale[1], 0.0f), 1.0f);
    out[2] = fminf(fmaxf((s.p.z - f.box_lo[2]) * f.box_scale[2], 0.0f), 1.0f);
    out[3] = (1.0f - dir.x) * 0.5f;
    out[4] = (1.0f - dir.y) * 0.5f;
    out[5] = (1.0f - dir.z) * 0.5f;
    out[6] = (s.n.x + 1.0f) * 0.5f;
    out[7] = (s.n.y + 1.0f) * 0.5f;
    out[8] = (s.n.z + 1.0f) * 0.5f;
    out[9] = m.type == kDiffuse ? 1.0f : fminf(fmaxf(m.roughness, 0.0f), 1.0f);
    out[10] = wl;
}
__device__ __forceinline__ void add_f4(float4 *p, S4 v) {
    float4 a = *p;
    *p = make_float4(a.x + v.v[0], a.y + v.v[1], a.z + v.v[2], a.w + v.v[3]);
}

// ---------------------------------------------------------------- kernels
// The camera path of pixel `pix` for sample `sample_index`, written to queue slot i (path id i).
__device__ __forceinline__ void generate_path(const FrameArgs &f, uint32_t pix, uint32_t sample_index, const PathQueue &q,
                                              uint32_t i, float *radiance) {
    const uint32_t gx = pix % f.width, gy = pix / f.width;
    // pixel_sampler (main.swift:508-515) and wavelength_numbers (main.swift:2218-2230).
    uint32_t seed = (gy * f.width + gx) ^ (sample_index * 1999999973u);
    const uint32_t h = pcg_hash(seed ^ 0x5bd1e995u);
    float wl = (float)(h >> 8u) * (1.0f / 16777216.0f);
    if (f.sampler == 1u) {
        // Z mode: dimension 0 of the vertex-0 Z_WAVELENGTH event under the pixel's key (main.swift:2221-2226).
        ZSampler zs{z_pixel_key(gx, gy, f.frame0 + (sample_index - f.sample_index0) + 1u, sample_index, f.z_temporal), 0u, 0u, 0u};
        zs.event(0u, kZWavelength);
        wl = zs.rand1();
    }
    // frameJitter: Halton(2, 3) - 0.5 with a 1024-frame period (main.swift:8773-8786).
    uint32_t jn = (max(1u, sample_index) - 1u) % 1024u + 1u;
    float jx = 0.0f, jy = 0.0f;
    {
        uint32_t m = jn;
        float frac = 1.0f;
        while (m > 0u) { frac /= 2.0f; jx += frac * (float)(m % 2u); m /= 2u; }
        m = jn;
        frac = 1.0f;
        while (m > 0u) { frac /= 3.0f; jy += frac * (float)(m % 3u); m /= 3u; }
        jx -= 0.5f;
        jy -= 0.5f;
    }
    // Pass-2 camera ray (main.swift:7420-7436).
    float u = (((float)gx + (jx + 0.5f)) / (float)f.width) * 2.0f - 1.0f;
    float v = (((float)gy + (jy + 0.5f)) / (float)f.height) * 2.0f - 1.0f;
    v = -v;
    u *= f.aspect * f.fov_scale;
    v *= f.fov_scale;
    V3 fw = load_v3(f.forward), rt = load_v3(f.right), up = load_v3(f.up), eye = load_v3(f.eye);
    V3 d = normalize(fw + rt * u + up * v);
    // decorrelate_shading_seed (main.swift:106-108): pass 2 draws from a rehashed stream.
    seed = pcg_hash(seed ^ 0x27d4eb2fu);
    q.ox[i] = eye.x; q.oy[i] = eye.y; q.oz[i] = eye.z;
    q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
    q.tmin[i] = ray_t_min(eye);
    q.tmax[i] = 1e20f;
    q.id[i] = i;
    q.seed[i] = seed;
    q.wl[i] = wl;
    q.t0[i] = q.t1[i] = q.t2[i] = q.t3[i] = 1.0f;
    q.state[i] = 0u;
    q.pdf[i] = 0.0f;
    q.px[i] = eye.x; q.py[i] = eye.y; q.pz[i] = eye.z;
    radiance[3 * (size_t)i] = 0.0f;
    radiance[3 * (size_t)i + 1] = 0.0f;
    radiance[3 * (size_t)i + 2] = 0.0f;
}

__global__ __launch_bounds__(256) void generate_kernel(FrameArgs f, PathQueue q, float *radiance, uint32_t *count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = f.npix * f.samples;
    if (i == 0) *count = n;
    if (i >= n) return;
    const uint32_t s = i / f.npix;
    generate_path(f, i - s * f.npix, f.sample_index0 + s, q, i, radiance);
}

// Pass 1 of ReSTIR DI for one sample (restir_temporal_kernel, main.swift:4842-4940): one thread per
// pixel over the camera wavefront (queue slot = pixel).  Writes the G-buffer and the pixel's
// reservoir: RIS over ra.candidates light samples on the pixel's un-decorrelated PCG stream
// (restir_di_initial :4637-4668), then the temporal merge of the backprojected history reservoir
// (restir_di_merge :4692-4704) and the encode (restir_di_encode :4708-4716).
// Without the environment the kernel fits 128 registers, 4 waves per SIMD (asked for explicitly: the
// light-tree walk is latency-bound and a fourth wave is worth more than the few registers it saves).
template <bool kEnv>
__global__ __launch_bounds__(256, kEnv ? 1 : 4) void restir_initial_kernel(RenderScene sc, FrameArgs fa, PathQueue q,
                                                                          HitArgs hits, RestirArgs ra) {
    const FrameArgs f = frame_env<kEnv>(fa);
    __shared__ LtTop top;
    __shared__ float top_l[kLtTopInner][256], top_r[kLtTopInner][256];
    const bool cached = sc.tree_node_count != 0u;
    if (cached && threadIdx.x < (unsigned)kLtTopAll) lt_top_fill(sc, top, threadIdx.x);
    __syncthreads();
    const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= f.npix) return;
    float4 gpos = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    uint2 gnorm = pack_half4(0.0f, 0.0f, 0.0f, -1.0f);
    float4 posdir = make_float4(0.0f, 0.0f, 0.0f, 0.0f), emitpdf = posdir, weights = posdir;
    const int face = hits.face[pix];
    if (face >= 0) {
        const V3 dir = v3(q.dx[pix], q.dy[pix], q.dz[pix]);
        const Surface s = surface_at(sc, (uint32_t)face, hits.b1[pix], hits.b2[pix], dir);
        const Mat m = material(sc, s.slot);
        const float t = hits.t[pix];
        gpos = make_float4(s.p.x, s.p.y, s.p.z, t);
        gnorm = pack_half4(s.n.x, s.n.y, s.n.z, (float)m.type);
        if (m.type == kDiffuse) {
            const Spectral sp{sc.icdf, reinterpret_cast<const float4 *>(sc.cmf), reinterpret_cast<const float4 *>(sc.basis)};
            // The PCG stream for the resampling decisions; in Z mode each light candidate i is
            // candidate i of the vertex-1 Z_DI event under the pixel's key (restir_di_initial,
            // main.swift:4635-4652), so the candidates are stratified.
            Rng rng{pix ^ (f.sample_index0 * 1999999973u), ZSampler{0u, 0u, 0u, 0u}, f.sampler == 1u};
            if (rng.z) rng.zs.key = z_pixel_key(pix % f.width, pix / f.width, f.frame0 + 1u, f.sample_index0, f.z_temporal);
            const uint32_t bits = ra.candidates > 1u ? 32u - (uint32_t)__clz(ra.candidates - 1u) : 0u;
            // The primary's BSDF continuation covers the compensated table's zeros unless the depth is 1.
            const uint32_t table = f.scattering_limit > 0 ? kEnvCompensated : kEnvOriginal;
            const float u = q.wl[pix];
            const Lanes lanes = lanes_of(sp, u);
            const S4 rho_pi = scale(reflectance(sp, lanes, m.lagrange), 1.0f / kPi);
            DiSample sel = di_empty();
            float wsum = 0.0f, M = 0.0f;
            if (cached && ra.candidates != 0u) lt_top_importances<256>(sc, top, s.p, top_l, top_r);
            const V3 sp_p = s.p;
            for (uint32_t i = 0; i < ra.candidates; ++i) {
                if (rng.z) rng.zs.candidate(1u, kZDi, i, bits);
                const LightSample ls =
                    cached ? sample_emitter_via(sc, f, sp_p, rng, table,
                                                [&](float u, float &pmf) { return lt_sample_top<256>(sc, sp_p, u, pmf, top, top_l, top_r); })
                           : sample_emitter(sc, f, sp_p, rng, table);
                DiSample c;
                c.position = ls.light == 1u ? ls.wi : ls.position;
                c.emission = ls.emission;
                c.light = ls.light;
                c.pdf = ls.pdf;
                c.u = u;
                M += 1.0f;
                if (c.pdf > 0.0f) {
                    const float ph = target_pdf(sp, lanes, rho_pi, s, dir, c, ls.geometry);
                    const float proposal = c.pdf * ls.geometry;
                    const float w = proposal > 0.0f ? ph / proposal : 0.0f;
                    wsum += w;
                    if (rand_f(rng.pcg) * wsum < w) sel = c;
                }
            }
            if (ra.temporal) {
                // restir_backproject through the previous sample's camera.
                const V3 d = s.p - load_v3(ra.prev_eye);
                const float w = dot(d, load_v3(ra.prev_forward));
                if (w > 0.0f) {
                    const float x = dot(d, load_v3(ra.prev_right)) / (w * ra.prev_aspect * ra.prev_fov_scale);
                    const float y = dot(d, load_v3(ra.prev_up)) / (w * ra.prev_fov_scale);
                    const float ux = x * 0.5f + 0.5f, uy = y * -0.5f + 0.5f;
                    if (ux >= 0.0f && uy >= 0.0f && ux < 1.0f && uy < 1.0f) {
                        const int px = (int)(ux * (float)f.width), py = (int)(uy * (float)f.height);
                        if (px >= 0 && px < (int)f.width && py >= 0 && py < (int)f.height) {
                            const uint32_t hp = (uint32_t)py * f.width + (uint32_t)px;
                            const float4 hw = ra.hist.weights[hp];
                            const float4 og = ra.hist.gpos[hp];
                            const float4 on = unpack_half4(ra.hist.gnorm[hp]);
                            // restir_same_surface (main.swift:4816-4823).
                            const bool same = og.w > 0.0f && on.w == (float)m.type &&
                                              dot(v3(on.x, on.y, on.z), s.n) > 0.95f &&
                                              length(v3(og.x, og.y, og.z) - s.p) < fmaxf(0.01f, t * 0.01f);
                            if (same && hw.y > 0.0f && hw.z > 0.0f) {
                                const DiSample hs = di_stored(ra.hist.posdir[hp], ra.hist.emitpdf[hp], hw);
                                const float ph = target_pdf_at(sc, f, sp, m, s, dir, hs);
                                const float cm = fminf(hw.y, 20.0f);
                                const float wt = ph * hw.z * cm;
                                M += cm;
                                wsum += wt;
                                if (rand_f(rng.pcg) * wsum < wt) sel = hs;
                            }
                        }
                    }
                }
            }
            const float ph = target_pdf_at(sc, f, sp, m, s, dir, sel);
            const float W = (M > 0.0f && ph > 0.0f) ? wsum / (M * ph) : 0.0f;
            posdir = make_float4(sel.position.x, sel.position.y, sel.position.z, pack_light(sel.light));
            emitpdf = make_float4(sel.emission.x, sel.emission.y, sel.emission.z, sel.pdf);
            weights = make_float4(wsum, M, W, sel.u);
            if (ra.gi) ra.gw.seed[pix] = rng.pcg;  // ReSTIR GI's draws continue this stream
        }
    }
    ra.cur.gpos[pix] = gpos;
    ra.cur.gnorm[pix] = gnorm;
    ra.cur.posdir[pix] = posdir;
    ra.cur.emitpdf[pix] = emitpdf;
    ra.cur.weights[pix] = weights;
}

// ReSTIR GI stage 1 (restir_gi_initial, main.swift:4733-4743): at a diffuse primary hit, the x1 -> x2
// direction, a BSDF sample drawn after ReSTIR DI's draws (the PCG state pass 1 left in gw.seed), and
// its ray, traced by the next stage.  Pixels without one get a ray that ends before it starts.
__global__ __launch_bounds__(256) void restir_gi_rays_kernel(RenderScene sc, FrameArgs f, PathQueue q, HitArgs hits,
                                                             RestirArgs ra) {
    const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= f.npix) return;
    const GiWork &g = ra.gw;
    bool ray = false;
    V3 o = v3(0.0f, 0.0f, 0.0f), gd = v3(0.0f, 0.0f, 1.0f);
    const int face = hits.face[pix];
    if (face >= 0) {
        const V3 dir = v3(q.dx[pix], q.dy[pix], q.dz[pix]);
        const Surface s = surface_at(sc, (uint32_t)face, hits.b1[pix], hits.b2[pix], dir);
        const Mat m = material(sc, s.slot);
        if (m.type == kDiffuse) {
            const Spectral sp{sc.icdf, reinterpret_cast<const float4 *>(sc.cmf), reinterpret_cast<const float4 *>(sc.basis)};
            Rng rng{g.seed[pix], ZSampler{0u, 0u, 0u, 0u}, f.sampler == 1u};
            if (rng.z) rng.zs.key = z_pixel_key(pix % f.width, pix / f.width, f.frame0 + 1u, f.sample_index0, f.z_temporal);
            rng_event(rng, 1u, kZGiBsdf);
            S4 bw;
            float gpdf;
            if (sample_bsdf<false>(sc, sp, lanes_of(sp, q.wl[pix]), s, m, dir, rng, gd, bw, gpdf) && gpdf > 0.0f) {
                o = ray_origin(s.p, s.ng, gd, s.error);
                g.bpdf[pix] = gpdf;
                ray = true;
            } else {
                gd = v3(0.0f, 0.0f, 1.0f);
            }
            g.seed[pix] = rng.pcg;
        }
    }
    g.stage[pix] = ray ? 1u : 0u;
    g.ox[pix] = o.x; g.oy[pix] = o.y; g.oz[pix] = o.z;
    g.dx[pix] = gd.x; g.dy[pix] = gd.y; g.dz[pix] = gd.z;
    g.tmin[pix] = ray ? ray_t_min(o) : 1.0f;
    g.tmax[pix] = ray ? 1e20f : 0.0f;
}

// ReSTIR GI stage 2 (restir_gi_initial at x2, main.swift:4744-4756): at a diffuse secondary hit, the
// light sample of the vertex-2 Z_GI_NEE event and its shadow ray (light_visible).  Stage 2 marks a
// sample whose visibility the shadow ray decides, 3 one with no pdf or no room between the offset
// origin and the light (not visible).
__global__ __launch_bounds__(256) void restir_gi_secondary_kernel(RenderScene sc, FrameArgs f, RestirArgs ra) {
    const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= f.npix) return;
    const GiWork &g = ra.gw;
    bool ray = false;
    V3 so = v3(0.0f, 0.0f, 0.0f), sd = v3(0.0f, 0.0f, 1.0f);
    float stmin = 1.0f, stmax = 0.0f;
    if (g.stage[pix] == 1u && g.face[pix] >= 0) {
        const V3 gd = v3(g.dx[pix], g.dy[pix], g.dz[pix]);
        const Surface x2 = surface_at(sc, (uint32_t)g.face[pix], g.b1[pix], g.b2[pix], gd);
        const Mat m2 = material(sc, x2.slot);
        if (m2.type == kDiffuse) {
            Rng rng{g.seed[pix], ZSampler{0u, 0u, 0u, 0u}, f.sampler == 1u};
            if (rng.z) rng.zs.key = z_pixel_key(pix % f.width, pix / f.width, f.frame0 + 1u, f.sample_index0, f.z_temporal);
            rng_event(rng, 2u, kZGiNee);
            const LightSample ls =
                sample_emitter(sc, f, x2.p, rng, f.scattering_limit >= 2 ? kEnvCompensated : kEnvOriginal);
            g.seed[pix] = rng.pcg;
            g.light_wi[pix] = make_float4(ls.wi.x, ls.wi.y, ls.wi.z, ls.pdf);
            g.light_le[pix] = make_float4(ls.emission.x, ls.emission.y, ls.emission.z, 0.0f);
            uint32_t stage = 3u;
            if (ls.pdf > 0.0f) ray = shadow_ray(x2, ls.light, ls.wi, ls.position, so, sd, stmin, stmax);
            if (ray) stage = 2u;
            g.stage[pix] = stage;
        }
    }
    g.sox[pix] = so.x; g.soy[pix] = so.y; g.soz[pix] = so.z;
    g.sdx[pix] = sd.x; g.sdy[pix] = sd.y; g.sdz[pix] = sd.z;
    g.stmin[pix] = ray ? stmin : 1.0f;
    g.stmax[pix] = ray ? stmax : 0.0f;
}

// ReSTIR GI stage 3: the initial reservoir (restir_gi_initial :4757-4774), the temporal merge of the
// backprojected history (restir_gi_merge :4784-4801) on the PCG stream continued from the stages
// before, and the encode (restir_gi_encode :4803-4811) into ra.gcur.
__global__ __launch_bounds__(256) void restir_gi_finish_kernel(RenderScene sc, FrameArgs f, PathQueue q, HitArgs hits,
                                                               RestirArgs ra) {
    const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= f.npix) return;
    const GiWork &g = ra.gw;
    float4 opos = make_float4(0.0f, 0.0f, 0.0f, 0.0f), orad = opos, ow = opos;
    uint2 onrm = pack_half4(0.0f, 0.0f, 0.0f, 0.0f);
    const int face = hits.face[pix];
    if (face >= 0) {
        const V3 dir = v3(q.dx[pix], q.dy[pix], q.dz[pix]);
        const Surface s = surface_at(sc, (uint32_t)face, hits.b1[pix], hits.b2[pix], dir);
        const Mat m = material(sc, s.slot);
        if (m.type == kDiffuse) {
            const Spectral sp{sc.icdf, reinterpret_cast<const float4 *>(sc.cmf), reinterpret_cast<const float4 *>(sc.basis)};
            const float u = q.wl[pix];
            V3 rpos = v3(0.0f, 0.0f, 0.0f), rnrm = rpos;
            S4 rrad = s4(0.0f);
            float rsrc = 0.0f, wsum = 0.0f, M = 1.0f, ru = u;
            if (g.stage[pix] == 2u && !g.occluded[pix]) {
                const V3 gd = v3(g.dx[pix], g.dy[pix], g.dz[pix]);
                const Surface x2 = surface_at(sc, (uint32_t)g.face[pix], g.b1[pix], g.b2[pix], gd);
                const Mat m2 = material(sc, x2.slot);
                const Lanes lanes = lanes_of(sp, u);
                const float4 lw = g.light_wi[pix], le = g.light_le[pix];
                const V3 wi = v3(lw.x, lw.y, lw.z);
                const float c2 = fmaxf(0.0f, dot(x2.n, wi));
                float pdf2;
                const S4 f2 = eval_bsdf_pdf<false>(sc, sp, lanes, x2, m2, -gd, wi, pdf2);
                const float mis = f.scattering_limit >= 2 ? power_heuristic(lw.w, pdf2) : 1.0f;
                const S4 rad = scale(mul(f2, emission_spectrum(sp, lanes, v3(le.x, le.y, le.z))), c2 * (mis / lw.w));
                const V3 dd = x2.p - s.p;
                const float d2 = dot(dd, dd);
                const float cs = fmaxf(0.0f, dot(x2.n, -gd));
                const float src = d2 > 1e-10f ? g.bpdf[pix] * cs / d2 : 0.0f;
                const float ph = gi_target(sp, m, s, dir, x2.p, x2.n, rad, u);
                if (src > 0.0f && ph > 0.0f) {
                    rpos = x2.p;
                    rnrm = x2.n;
                    rrad = rad;
                    rsrc = src;
                    wsum = ph / src;
                }
            }
            if (ra.gi_temporal) {
                // The history pixel and the same-surface test of the DI merge (restir_backproject,
                // restir_same_surface), then the reconnection's Jacobian test.
                uint32_t seed = g.seed[pix];
                const V3 d = s.p - load_v3(ra.prev_eye);
                const float w = dot(d, load_v3(ra.prev_forward));
                if (w > 0.0f) {
                    const float x = dot(d, load_v3(ra.prev_right)) / (w * ra.prev_aspect * ra.prev_fov_scale);
                    const float y = dot(d, load_v3(ra.prev_up)) / (w * ra.prev_fov_scale);
                    const float ux = x * 0.5f + 0.5f, uy = y * -0.5f + 0.5f;
                    if (ux >= 0.0f && uy >= 0.0f && ux < 1.0f && uy < 1.0f) {
                        const int px = (int)(ux * (float)f.width), py = (int)(uy * (float)f.height);
                        if (px >= 0 && px < (int)f.width && py >= 0 && py < (int)f.height) {
                            const uint32_t hp = (uint32_t)py * f.width + (uint32_t)px;
                            const float4 og = ra.hist.gpos[hp];
                            const float4 on = unpack_half4(ra.hist.gnorm[hp]);
                            const bool same = og.w > 0.0f && on.w == (float)m.type &&
                                              dot(v3(on.x, on.y, on.z), s.n) > 0.95f &&
                                              length(v3(og.x, og.y, og.z) - s.p) < fmaxf(0.01f, hits.t[pix] * 0.01f);
                            const float4 hw = ra.ghist.weights[hp], hpos = ra.ghist.pos[hp];
                            const float4 hn = unpack_half4(ra.ghist.nrm[hp]);
                            if (same && hw.y > 0.0f && hw.z > 0.0f && hpos.w > 0.0f && hn.w > 0.0f) {
                                const V3 op = v3(hpos.x, hpos.y, hpos.z), onr = v3(hn.x, hn.y, hn.z);
                                if (gi_accepts_shift(s.p, v3(og.x, og.y, og.z), op, onr)) {
                                    const S4 hrad = s4_of(ra.ghist.rad[hp]);
                                    const float ph = gi_target(sp, m, s, dir, op, onr, hrad, hw.w);
                                    const float cm = fminf(hw.y, 20.0f);
                                    const float tw = ph * hw.z * cm;
                                    M += cm;
                                    wsum += tw;
                                    if (rand_f(seed) * wsum < tw) {
                                        rpos = op;
                                        rnrm = onr;
                                        rrad = hrad;
                                        rsrc = hpos.w;
                                        ru = hw.w;
                                    }
                                }
                            }
                        }
                    }
                }
            }
            const float gt = gi_target(sp, m, s, dir, rpos, rnrm, rrad, ru);
            const float W = M > 0.0f && gt > 0.0f ? wsum / (M * gt) : 0.0f;
            opos = make_float4(rpos.x, rpos.y, rpos.z, rsrc);
            onrm = pack_half4(rnrm.x, rnrm.y, rnrm.z, rsrc > 0.0f ? 1.0f : 0.0f);
            orad = make_float4(rrad.v[0], rrad.v[1], rrad.v[2], rrad.v[3]);
            ow = make_float4(wsum, M, W, ru);
        }
    }
    ra.gcur.pos[pix] = opos;
    ra.gcur.nrm[pix] = onrm;
    ra.gcur.rad[pix] = orad;
    ra.gcur.weights[pix] = ow;
    // The visibility arrays now hold pass 2's connection rays: none until the shading writes one.
    g.sox[pix] = g.soy[pix] = g.soz[pix] = 0.0f;
    g.sdx[pix] = g.sdy[pix] = 0.0f;
    g.sdz[pix] = 1.0f;
    g.stmin[pix] = 1.0f;
    g.stmax[pix] = 0.0f;
    g.rgb[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// The GI connections of pass 2 that x1 sees add their contribution (gi_connection_visible,
// main.swift:7640-7657).  Runs after the DI shadow resolve, so each pixel has one writer at a time.
__global__ __launch_bounds__(256) void restir_gi_resolve_kernel(FrameArgs f, RestirArgs ra, float *radiance) {
    const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= f.npix) return;
    if (ra.gw.occluded[pix]) return;  // pixels without a connection hold a zero contribution
    const float4 v = ra.gw.rgb[pix];
    radiance[3 * (size_t)pix] += v.x;
    radiance[3 * (size_t)pix + 1] += v.y;
    radiance[3 * (size_t)pix + 2] += v.z;
}

// Shading of one wavefront vertex per path.  kRestir: sampling mode 0, where a diffuse primary vertex
// takes its direct light from ReSTIR DI's spatial reuse (main.swift:7484-7580) instead of NEE.
// Shading of queue slot i (active: i holds a path) by every lane of the wave together: the queue
// appends are wave-wide.  `job` is carried to the continuation and shadow queues when they have a job
// array (batched rendering).
template <bool kRestir, bool kLayered>
__device__ __forceinline__ void shade_path(const RenderScene &sc, const FrameArgs &f, const PathQueue &in,
                                           const HitArgs &hits, const PathQueue &out, const ShadowQueue &sh,
                                           const Counters &c, uint32_t capacity, float *radiance, const RestirArgs &ra,
                                           const NrcArgs &nrc, uint32_t i, bool active, uint32_t job) {
    bool cont = false, shadow = false;
    // Continuation and shadow-ray values produced below (written after the wave-wide appends).
    V3 new_o = v3(0.0f, 0.0f, 0.0f), new_d = v3(0.0f, 0.0f, 1.0f), prev_p = v3(0.0f, 0.0f, 0.0f);
    V3 sh_o = new_o, sh_d = new_d, sh_c = new_o;
    S4 sh_s = s4(0.0f);
    float sh_tmin = 0.0f, sh_tmax = 0.0f, new_pdf = 0.0f;
    // Cache query (rendering paths) or training record (training paths) of this vertex.
    bool query = false, record = false, training = false;
    float cin[kNrcInputCount];
    uint32_t id = 0, state = 0;
    Rng rng{0u, ZSampler{0u, 0u, 0u, 0u}, false};
    float wl = 0.0f;
    S4 thr = s4(0.0f);
    if (active) {
        id = in.id[i];
        rng.pcg = in.seed[i];
        if (f.sampler == 1u) {
            // Z mode: the pixel's key; path id = sample * npix + pixel, and sample k is frame
            // frame0 + k + 1 of the accumulation.
            const uint32_t k = id / f.npix, pixel = id - k * f.npix;
            rng.zs.key = z_pixel_key(pixel % f.width, pixel / f.width, f.frame0 + k + 1u, f.sample_index0 + k, f.z_temporal);
            rng.z = true;
        }
        wl = in.wl[i];
        thr = S4{{in.t0[i], in.t1[i], in.t2[i], in.t3[i]}};
        state = in.state[i];
        const uint32_t bounce = state & 0xffu;
        int sdepth = (int)((state >> 8) & 0xffu);
        const int face = hits.face[i];
        const V3 dir = v3(in.dx[i], in.dy[i], in.dz[i]);
        const Spectral sp{sc.icdf, reinterpret_cast<const float4 *>(sc.cmf), reinterpret_cast<const float4 *>(sc.basis)};
        const Lanes lanes = lanes_of(sp, wl);
        V3 add = v3(0.0f, 0.0f, 0.0f);
        bool alive = face >= 0;
        Surface s;
        Mat m;
        if (!alive && env_on(sc, f)) {
            // A ray that leaves the scene: the camera sees the environment's exact colour
            // (environment_rgb); a continuation gathers it with the MIS weight of its light-sample
            // density in the compensated table (main.swift:7454, 7751-7758).
            if (bounce == 0u) {
                add = eval_environment(sc, f, dir);
            } else {
                const float w = emission_weight((state & kPrevDelta) != 0u, (state & kPrevNee) != 0u,
                                                (state & kPrevMis) != 0u, in.pdf[i], env_pdf(sc, f, dir, kEnvCompensated));
                const S4 se = scale(mul(thr, emission_spectrum(sp, lanes, eval_environment(sc, f, dir))), w);
                add = spectrum_rgb(sp, lanes, se);
                if (f.nrc && nrc_training(f, id)) add_f4(nrc.sacc + id, se);
            }
        }
        if (alive) {
            s = surface_at(sc, (uint32_t)face, hits.b1[i], hits.b2[i], dir);
            m = material(sc, s.slot);
            if (m.type == kEmissive) {
                V3 e = s.front ? m.emission : v3(0.0f, 0.0f, 0.0f);
                if (bounce == 0u) {
                    add = e;  // camera rays see the exact RGB (emitter_rgb, D65)
                } else {
                    const V3 prev = v3(in.px[i], in.py[i], in.pz[i]);
                    float light_pdf = emitter_pdf(sc, f, prev, s.p, (uint32_t)face, m.weight);
                    float w = emission_weight((state & kPrevDelta) != 0u, (state & kPrevNee) != 0u,
                                              (state & kPrevMis) != 0u, in.pdf[i], light_pdf);
                    const S4 se = scale(mul(thr, emission_spectrum(sp, lanes, e)), w);
                    add = spectrum_rgb(sp, lanes, se);
                    if (f.nrc && nrc_training(f, id)) add_f4(nrc.sacc + id, se);
                }
                alive = false;
            }
        }
        if (alive && (state & kEmissionOnly) != 0u) alive = false;
        if (f.nrc && alive && bounce > 0u && (state & (kPrevDelta | kRecorded)) == 0u) {
            // The first vertex after a non-delta scatter: rendering paths end here with the cache's
            // answer; training paths go on and record this vertex only, so that the records see the
            // same remaining depth as the queries the cache answers.
            training = nrc_training(f, id);
            nrc_inputs(f, s, m, dir, wl, cin);
            if (training) {
                record = true;
            } else {
                query = true;
                alive = false;
            }
        }
        if (kRestir && alive && bounce == 0u && m.type == kDiffuse) {
            // Spatial reuse: the pixel's reservoir and four uniform taps in a 16-pixel box
            // (uniform_neighbor, restir2020_compatible), combined with w = p_hat W M.
            const uint32_t pix = id;
            const uint32_t gx = pix % f.width, gy = pix / f.width;
            const float4 cg = ra.cur.gpos[pix], cw = ra.cur.weights[pix];
            DiSample sel = di_stored(ra.cur.posdir[pix], ra.cur.emitpdf[pix], cw);
            float wsum = target_pdf_at(sc, f, sp, m, s, dir, sel) * cw.z * cw.y;
            float M = cw.y;
            for (int k = 0; k < 4; ++k) {
                const float rx = rand_f(rng.pcg), ry = rand_f(rng.pcg);
                const int nx = (int)gx + (int)((rx * 2.0f - 1.0f) * 16.0f);
                const int ny = (int)gy + (int)((ry * 2.0f - 1.0f) * 16.0f);
                if (nx < 0 || nx >= (int)f.width || ny < 0 || ny >= (int)f.height) continue;
                const uint32_t np = (uint32_t)ny * f.width + (uint32_t)nx;
                const float4 ng = ra.cur.gpos[np];
                const float4 nn = unpack_half4(ra.cur.gnorm[np]);
                const bool compatible = ng.w > 0.0f && nn.w == (float)m.type && dot(s.n, v3(nn.x, nn.y, nn.z)) > 0.95f &&
                                        fabsf(cg.w - ng.w) < 0.05f * cg.w;
                if (!compatible) continue;
                const float4 nw = ra.cur.weights[np];
                if (!(nw.y > 0.0f && nw.z > 0.0f)) continue;
                const DiSample ns = di_stored(ra.cur.posdir[np], ra.cur.emitpdf[np], nw);
                const float w = target_pdf_at(sc, f, sp, m, s, dir, ns) * nw.z * nw.y;
                wsum += w;
                M += nw.y;
                if (rand_f(rng.pcg) * wsum < w) sel = ns;
            }
            const float fp = target_pdf_at(sc, f, sp, m, s, dir, sel);
            const float W = (M > 0.0f && fp > 0.0f) ? wsum / (M * fp) : 0.0f;
            if (W > 0.0f) {
                // Shading with visibility (light_visible): bsdf cos L W G at the sample's wavelengths.
                const V3 wi = sel.light == 1u ? sel.position : normalize(sel.position - s.p);
                if (shadow_ray(s, sel.light, wi, sel.position, sh_o, sh_d, sh_tmin, sh_tmax)) {
                    const float ct = fmaxf(0.0f, dot(s.n, wi));
                    const Lanes sl = lanes_of(sp, sel.u);
                    const S4 fb = diffuse_f(s, scale(reflectance(sp, sl, m.lagrange), 1.0f / kPi), -dir, wi);
                    const S4 v = scale(scale(mul(scale(fb, ct), emission_spectrum(sp, sl, di_emission(sc, f, sel))), W),
                                       light_geometry(sc, s.p, sel));
                    sh_c = spectrum_rgb(sp, sl, v);  // at the sample's own lanes; not a cache target (primary)
                    shadow = true;
                }
            }
            if (ra.gi) {
                // ReSTIR GI spatial reuse (main.swift:7586-7657) on the same PCG stream: the pixel's GI
                // reservoir and four uniform taps, each tap's x2 reconnected to this x1 when the
                // neighbour's surface is compatible and the shift's Jacobian is in range.
                const float4 g0 = ra.gcur.pos[pix], w0 = ra.gcur.weights[pix];
                const float4 n0 = unpack_half4(ra.gcur.nrm[pix]);
                V3 spos = v3(g0.x, g0.y, g0.z), snrm = v3(n0.x, n0.y, n0.z);
                S4 srad = s4_of(ra.gcur.rad[pix]);
                float su = w0.w;
                bool sflag = n0.w > 0.0f;
                float gws = (sflag ? gi_target(sp, m, s, dir, spos, snrm, srad, su) : 0.0f) * w0.z * w0.y;
                float gM = w0.y;
                for (int k = 0; k < 4; ++k) {
                    const float rx = rand_f(rng.pcg), ry = rand_f(rng.pcg);
                    const int nx = (int)gx + (int)((rx * 2.0f - 1.0f) * 16.0f);
                    const int ny = (int)gy + (int)((ry * 2.0f - 1.0f) * 16.0f);
                    if (nx < 0 || nx >= (int)f.width || ny < 0 || ny >= (int)f.height) continue;
                    const uint32_t np = (uint32_t)ny * f.width + (uint32_t)nx;
                    const float4 ng = ra.cur.gpos[np];
                    const float4 nn = unpack_half4(ra.cur.gnorm[np]);
                    const bool compatible = ng.w > 0.0f && nn.w == (float)m.type && dot(s.n, v3(nn.x, nn.y, nn.z)) > 0.95f &&
                                            fabsf(cg.w - ng.w) < 0.05f * cg.w;
                    if (!compatible) continue;
                    const float4 qw = ra.gcur.weights[np], qp = ra.gcur.pos[np];
                    const float4 qn = unpack_half4(ra.gcur.nrm[np]);
                    if (qw.y <= 0.0f || qw.z <= 0.0f || qp.w <= 0.0f || qn.w <= 0.0f) continue;
                    const V3 q2 = v3(qp.x, qp.y, qp.z), qn2 = v3(qn.x, qn.y, qn.z);
                    if (!gi_accepts_shift(s.p, v3(ng.x, ng.y, ng.z), q2, qn2)) continue;
                    const S4 qrad = s4_of(ra.gcur.rad[np]);
                    const float w = gi_target(sp, m, s, dir, q2, qn2, qrad, qw.w) * qw.z * qw.y;
                    gws += w;
                    gM += qw.y;
                    if (rand_f(rng.pcg) * gws < w) {
                        spos = q2;
                        snrm = qn2;
                        srad = qrad;
                        su = qw.w;
                        sflag = true;
                    }
                }
                const float ft = sflag ? gi_target(sp, m, s, dir, spos, snrm, srad, su) : 0.0f;
                const float gW = gM > 0.0f && ft > 0.0f ? gws / (gM * ft) : 0.0f;
                const V3 delta = spos - s.p;
                const float dist = length(delta);
                if (gW > 0.0f && dist > 2.0f * ray_epsilon(s.p)) {
                    // The connection ray of gi_connection_visible, traced after this pass; its
                    // contribution f rad G W at the sample's wavelengths.
                    const V3 o = ray_origin(s.p, s.ng, delta * (1.0f / dist), s.error);
                    const V3 ed = spos - o;
                    const float d = length(ed);
                    const V3 cd = ed * (1.0f / d);
                    const V3 wi = normalize(delta);
                    const Lanes sl = lanes_of(sp, su);
                    const S4 fb = diffuse_f(s, scale(reflectance(sp, sl, m.lagrange), 1.0f / kPi), -dir, wi);
                    const V3 rgb = spectrum_rgb(sp, sl, scale(mul(fb, srad), gi_geometry(s.p, s.n, spos, snrm) * gW));
                    const GiWork &g = ra.gw;
                    g.sox[pix] = o.x; g.soy[pix] = o.y; g.soz[pix] = o.z;
                    g.sdx[pix] = cd.x; g.sdy[pix] = cd.y; g.sdz[pix] = cd.z;
                    g.stmin[pix] = ray_t_min(o);
                    g.stmax[pix] = d - endpoint_tolerance(o, spos, d);
                    g.rgb[pix] = make_float4(rgb.x, rgb.y, rgb.z, 0.0f);
                }
            }
        } else if (alive && !is_delta(m) && f.sampling_mode != 3u &&
                   !((state & kGiSecondary) != 0u && bounce == 1u && m.type == kDiffuse)) {
            // (A diffuse secondary vertex of a ReSTIR GI path takes no NEE: the GI reservoir holds
            // that light, restirGISecondary, main.swift:7801-7803.)
            // Next-event estimation at this non-delta vertex (main.swift:7655-7672, 7790-7805).
            const bool use_mis = f.sampling_mode <= 1u && (bounce == 0u ? f.scattering_limit > 0 : sdepth < f.scattering_limit);
            rng_event(rng, bounce + 1u, kZNee);  // this hit is vertex bounce + 1 (main.swift:7658, 7804)
            // The environment's compensated table where this vertex's BSDF continuation follows.
            const bool continues = bounce == 0u ? f.scattering_limit > 0 : sdepth < f.scattering_limit;
            LightSample ls = sample_emitter(sc, f, s.p, rng, continues ? kEnvCompensated : kEnvOriginal);
            V3 o, d;
            float tmin, tmax;
            if (ls.pdf > 0.0f && shadow_ray(s, ls.light, ls.wi, ls.position, o, d, tmin, tmax)) {
                float pdf;
                const S4 fbsdf = eval_bsdf_pdf<kLayered>(sc, sp, lanes, s, m, -dir, ls.wi, pdf);
                if (smax(fbsdf) > 0.0f) {
                    const float cosine = fabsf(dot(s.n, ls.wi));
                    const float w = use_mis ? power_heuristic(ls.pdf, pdf) : 1.0f;
                    const S4 le = emission_spectrum(sp, lanes, ls.emission);
                    sh_s = scale(mul(mul(thr, fbsdf), le), cosine * w / ls.pdf);
                    sh_c = spectrum_rgb(sp, lanes, sh_s);
                    sh_o = o;
                    sh_d = d;
                    sh_tmin = tmin;
                    sh_tmax = tmax;
                    shadow = true;
                }
            }
        }
        if (alive && bounce > 3u) {
            // Russian roulette after the third bounce (main.swift:7819-7827).
            const float survival = fminf(fmaxf(smax(thr), 0.05f), is_delta(m) ? 0.99f : 0.95f);
            rng_event(rng, bounce + 1u, kZRoulette);
            if (rand_f(rng) >= survival) alive = false;
            else thr = scale(thr, 1.0f / survival);
        }
        const bool prev_delta = alive && is_delta(m);
        bool emission_only = false;
        if (alive) {
            // Departure: the scattering budget counts non-delta events (main.swift:7712-7742).
            if (!prev_delta && sdepth >= f.scattering_limit) {
                if (f.sampling_mode != 3u) alive = false;
                else emission_only = true;
            }
        }
        if (alive) {
            if (!prev_delta) ++sdepth;
            V3 nd;
            S4 bw;
            float pdf;
            rng_event(rng, bounce + 1u, kZBsdf);  // vertex bounce + 1 scatters (main.swift:7730)
            if (sample_bsdf<kLayered>(sc, sp, lanes, s, m, dir, rng, nd, bw, pdf)) {
                thr = mul(thr, bw);
                bool finite = isfinite(thr.v[0]) && isfinite(thr.v[1]) && isfinite(thr.v[2]) && isfinite(thr.v[3]);
                if (finite && smax(thr) > 0.0f) {
                    new_o = ray_origin(s.p, s.ng, nd, s.error);
                    new_d = nd;
                    new_pdf = pdf;
                    prev_p = s.p;
                    // previousMIS (main.swift:7725-7728): a diffuse primary's DI reservoir estimated the
                    // whole direct integral, so in ReSTIR mode its continuation is not MIS-weighted.
                    const bool mis = f.sampling_mode == 1u || (f.sampling_mode == 0u && (bounce > 0u || m.type != kDiffuse));
                    const bool nee = !prev_delta && f.sampling_mode != 3u;
                    const bool gi_path = kRestir && ra.gi && bounce == 0u && m.type == kDiffuse;
                    uint32_t flags = (prev_delta ? kPrevDelta : 0u) | (nee ? kPrevNee : 0u) | (mis ? kPrevMis : 0u) |
                                     (emission_only ? kEmissionOnly : 0u) | ((record || (state & kRecorded)) ? kRecorded : 0u) |
                                     (gi_path ? kGiSecondary : 0u);
                    state = min(bounce + 1u, 255u) | ((uint32_t)min(sdepth, 255) << 8) | flags;
                    cont = true;
                }
            }
        }
        if (add.x != 0.0f || add.y != 0.0f || add.z != 0.0f) {
            radiance[3 * (size_t)id] += add.x;
            radiance[3 * (size_t)id + 1] += add.y;
            radiance[3 * (size_t)id + 2] += add.z;
        }
    }
    const uint32_t slot = wave_append(c.out, cont);
    const uint32_t sslot = wave_append(c.shadow, shadow);
    if (f.nrc) {
        const uint32_t qslot = wave_append(nrc.q_count, query);
        const uint32_t rslot = wave_append(nrc.r_count, record);
        if (query && qslot < capacity) {
            for (int k = 0; k < kNrcInputCount; ++k) nrc.q_in[(size_t)k * nrc.q_stride + qslot] = cin[k];
            nrc.q_thr[qslot] = make_float4(thr.v[0], thr.v[1], thr.v[2], thr.v[3]);  // unchanged: the path ended
            nrc.q_wl[qslot] = wl;
            nrc.q_id[qslot] = id;
        }
        if (record && rslot < nrc.r_cap) {
            for (int k = 0; k < kNrcInputCount; ++k) nrc.r_in[(size_t)k * nrc.r_cap + rslot] = cin[k];
            // The throughput that reached the vertex (thr now carries this vertex's roulette and BSDF
            // weight for the continuation): reloaded rather than kept live through the whole body.
            nrc.r_thr[rslot] = make_float4(in.t0[i], in.t1[i], in.t2[i], in.t3[i]);
            nrc.r_snap[rslot] = nrc.sacc[id];
            nrc.r_id[rslot] = id;
        }
    }
    if (cont && slot < capacity) {
        out.ox[slot] = new_o.x; out.oy[slot] = new_o.y; out.oz[slot] = new_o.z;
        out.dx[slot] = new_d.x; out.dy[slot] = new_d.y; out.dz[slot] = new_d.z;
        out.tmin[slot] = ray_t_min(new_o);
        out.tmax[slot] = 1e20f;
        out.id[slot] = id;
        out.seed[slot] = rng.pcg;
        out.wl[slot] = wl;
        out.t0[slot] = thr.v[0]; out.t1[slot] = thr.v[1]; out.t2[slot] = thr.v[2]; out.t3[slot] = thr.v[3];
        out.state[slot] = state;
        out.pdf[slot] = new_pdf;
        out.px[slot] = prev_p.x; out.py[slot] = prev_p.y; out.pz[slot] = prev_p.z;
        if (out.job) out.job[slot] = job;
    }
    if (shadow && sslot < capacity) {
        sh.ox[sslot] = sh_o.x; sh.oy[sslot] = sh_o.y; sh.oz[sslot] = sh_o.z;
        sh.dx[sslot] = sh_d.x; sh.dy[sslot] = sh_d.y; sh.dz[sslot] = sh_d.z;
        sh.tmin[sslot] = sh_tmin;
        sh.tmax[sslot] = sh_tmax;
        sh.id[sslot] = id;
        sh.r[sslot] = sh_c.x; sh.g[sslot] = sh_c.y; sh.b[sslot] = sh_c.z;
        if (f.nrc) sh.s[sslot] = make_float4(sh_s.v[0], sh_s.v[1], sh_s.v[2], sh_s.v[3]);  // read for cache training only
        if (sh.job) sh.job[sslot] = job;
    }
}

template <bool kRestir, bool kLayered, bool kEnv>
__global__ __launch_bounds__(256) void shade_kernel(RenderScene sc, FrameArgs f, PathQueue in, HitArgs hits,
                                                    PathQueue out, ShadowQueue sh, Counters c, uint32_t capacity,
                                                    float *radiance, RestirArgs ra, NrcArgs nrc) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = min(*c.in, capacity);
    shade_path<kRestir, kLayered>(sc, frame_env<kEnv>(f), in, hits, out, sh, c, capacity, radiance, ra, nrc, i, i < n, 0u);
}
// Material sorting: paths whose hit needs the layered BSDF (GLOSSY above the delta threshold, OPENPBR)
// are listed from the end of `perm`, all others (Lambert, delta, emitters, misses) from its start, both
// by wave-aggregated appends; each class is then shaded by its own kernel (shade_sorted_kernel), so a
// wave runs one BSDF branch and the Lambert class keeps the lean kernel's register budget.
__device__ __forceinline__ bool needs_layered(const Mat &m) {
    return m.type == kOpenPbr || (m.type == kGlossy && !is_delta(m));
}
__global__ __launch_bounds__(256) void classify_kernel(RenderScene sc, HitArgs hits, const uint32_t *count, uint32_t capacity,
                                                       uint32_t *perm, uint32_t *classes) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = min(*count, capacity);
    const bool active = i < n;
    bool layered = false;
    if (active) {
        const int face = hits.face[i];
        if (face >= 0) layered = needs_layered(material(sc, sc.faces[4 * (size_t)face + 3]));
    }
    const uint32_t a = wave_append(classes, active && !layered);
    const uint32_t b = wave_append(classes + 1, active && layered);
    if (active && !layered) perm[a] = i;
    if (active && layered) perm[capacity - 1u - b] = i;
}
template <bool kLayered>
__global__ __launch_bounds__(256) void shade_sorted_kernel(RenderScene sc, FrameArgs f, PathQueue in, HitArgs hits,
                                                           PathQueue out, ShadowQueue sh, Counters c, uint32_t capacity,
                                                           float *radiance, NrcArgs nrc, const uint32_t *perm,
                                                           const uint32_t *class_count) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    const bool active = k < min(*class_count, capacity);
    const uint32_t i = active ? perm[kLayered ? capacity - 1u - k : k] : 0u;
    shade_path<false, kLayered>(sc, f, in, hits, out, sh, c, capacity, radiance, RestirArgs{}, nrc, i, active, 0u);
}

// The Lambert / delta shading without ReSTIR (a kernel of its own so its register allocation is not
// shaped by the layered BSDF's).
template <bool kEnv>
__global__ __launch_bounds__(256) void shade_lean_kernel(
    RenderScene sc, FrameArgs f, PathQueue in, HitArgs hits, PathQueue out, ShadowQueue sh, Counters c,
    uint32_t capacity, float *radiance, NrcArgs nrc) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = min(*c.in, capacity);
    shade_path<false, false>(sc, frame_env<kEnv>(f), in, hits, out, sh, c, capacity, radiance, RestirArgs{}, nrc, i,
                             i < n, 0u);
}

__global__ __launch_bounds__(256) void resolve_kernel(FrameArgs f, ShadowQueue sh, const uint8_t *occluded,
                                                      const uint32_t *count, uint32_t capacity, float *radiance,
                                                      NrcArgs nrc) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= min(*count, capacity)) return;
    if (occluded[i]) return;
    const uint32_t id = sh.id[i];
    radiance[3 * (size_t)id] += sh.r[i];
    radiance[3 * (size_t)id + 1] += sh.g[i];
    radiance[3 * (size_t)id + 2] += sh.b[i];
    if (f.nrc && nrc_training(f, id)) {
        const float4 v = sh.s[i];
        add_f4(nrc.sacc + id, S4{{v.x, v.y, v.z, v.w}});
    }
}

// radiance[id] += rgb(thr * max(y, 0) / unit) for every cache query (its answer y at the path's lanes).
__global__ __launch_bounds__(256) void cache_resolve_kernel(RenderScene sc, NrcArgs nrc, uint32_t capacity, float *radiance) {
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= min(*nrc.q_count, capacity)) return;
    const Spectral sp{sc.icdf, reinterpret_cast<const float4 *>(sc.cmf), reinterpret_cast<const float4 *>(sc.basis)};
    const Lanes lanes = lanes_of(sp, nrc.q_wl[q]);
    const float4 t = nrc.q_thr[q];
    const float *y = nrc.q_out + 4 * (size_t)q;
    const float iu = 1.0f / nrc.unit;
    const S4 v{{t.x * fmaxf(y[0], 0.0f) * iu, t.y * fmaxf(y[1], 0.0f) * iu, t.z * fmaxf(y[2], 0.0f) * iu,
                t.w * fmaxf(y[3], 0.0f) * iu}};
    const V3 rgb = spectrum_rgb(sp, lanes, v);
    if (!(isfinite(rgb.x) && isfinite(rgb.y) && isfinite(rgb.z))) return;
    const uint32_t id = nrc.q_id[q];
    radiance[3 * (size_t)id] += rgb.x;
    radiance[3 * (size_t)id + 1] += rgb.y;
    radiance[3 * (size_t)id + 2] += rgb.z;
}

// Training target of each record: the spectral radiance its path gathered after reaching the vertex,
// per unit throughput, in the network's unit.
__global__ __launch_bounds__(256) void build_targets_kernel(NrcArgs nrc) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r >= min(*nrc.r_count, nrc.r_cap)) return;
    const float4 fin = nrc.sacc[nrc.r_id[r]], snap = nrc.r_snap[r], t = nrc.r_thr[r];
    float *out = nrc.r_target + 4 * (size_t)r;
    out[0] = t.x > 0.0f ? (fin.x - snap.x) / t.x * nrc.unit : 0.0f;
    out[1] = t.y > 0.0f ? (fin.y - snap.y) / t.y * nrc.unit : 0.0f;
    out[2] = t.z > 0.0f ? (fin.z - snap.z) / t.z * nrc.unit : 0.0f;
    out[3] = t.w > 0.0f ? (fin.w - snap.w) / t.w * nrc.unit : 0.0f;
    for (int k = 0; k < 4; ++k)
        if (!isfinite(out[k])) out[k] = 0.0f;
}

// The decoded light tree (LtDec, 5 float4 per node) from the quantised node words.
__global__ __launch_bounds__(256) void decode_light_tree_kernel(RenderScene sc, float4 *out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= sc.tree_node_count) return;
    const LtNode n = lt_node(sc, i);
    const V3 pc = (n.lo + n.hi) * 0.5f, dh = pc - n.hi;
    const uint32_t links = reinterpret_cast<const uint4 *>(sc.tree_nodes)[2 * i + 1].w;
    float4 *q = out + kLtDecWords * (size_t)i;
    q[0] = make_float4(n.lo.x, n.lo.y, n.lo.z, n.phi);
    q[1] = make_float4(n.hi.x, n.hi.y, n.hi.z, fmaxf(0.5f * length(n.hi - n.lo), 1e-20f));
    q[2] = make_float4(n.w.x, n.w.y, n.w.z, n.cos_o);
    q[3] = make_float4(n.cos_e, __uint_as_float(n.child), __uint_as_float(links), dot(dh, dh));
}

__global__ __launch_bounds__(256) void zero_f4_kernel(float4 *p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- energy-table bakes
__device__ __forceinline__ uint32_t reverse_bits32(uint32_t v) { return __builtin_bitreverse32(v); }
// scripts/generate_metal_energy.py:12-33: node (roughness i/128, cosine max(1e-5, (j/128)^2)), mean of
// 1 - G1(l) over 16384 Hammersley VNDF samples; one thread per node, sequential float sum.
__global__ __launch_bounds__(64) void bake_metal_kernel(float *out) {
    const uint32_t node = blockIdx.x * 64 + threadIdx.x;
    if (node >= 129u * 129u) return;
    const uint32_t i = node / 129u, j = node % 129u;
    if (i == 0u) {
        out[node] = 0.0f;
        return;
    }
    const float roughness = (float)i / 128.0f;
    const float cosine = fmaxf(1e-5f, powf((float)j / 128.0f, 2.0f));
    const V3 v = v3(sqrtf(fmaxf(0.0f, 1.0f - cosine * cosine)), 0.0f, cosine);
    const float alpha = roughness * roughness;
    float sum = 0.0f;
    for (uint32_t k = 0; k < 16384u; ++k) {
        const float r0 = ((float)k + 0.5f) / 16384.0f, r1 = (float)reverse_bits32(k) * 2.3283064365386963e-10f;
        const V3 h = sample_vndf(alpha, alpha, v, r0, r1);
        const float d = dot(v, h);
        const V3 l = -v + h * (2.0f * d);
        const float e = l.z > 0.0f ? smith_g1(l, alpha, alpha) : 0.0f;
        sum += 1.0f - e;
    }
    out[node] = sum / 16384.0f;
}
__global__ void reset_kernel(uint32_t *counters, uint32_t n) {
    if (threadIdx.x < n) counters[threadIdx.x] = 0u;
}

// Running mean per pixel over this wavefront's samples in sample order (main.swift:7862-7913):
// NaN/Inf samples count as 0; negative channels stay (spectral transport).
__global__ __launch_bounds__(256) void accumulate_kernel(FrameArgs f, const float *radiance, float *accum,
                                                         uint32_t frame_index) {
    const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= f.npix) return;
    float4 a = reinterpret_cast<float4 *>(accum)[pix];
    for (uint32_t s = 0; s < f.samples; ++s) {
        const size_t id = (size_t)s * f.npix + pix;
        float r = radiance[3 * id], g = radiance[3 * id + 1], b = radiance[3 * id + 2];
        if (!(isfinite(r) && isfinite(g) && isfinite(b))) r = g = b = 0.0f;
        const uint32_t k = frame_index + s;
        if (k <= 1u) {
            a = make_float4(r, g, b, 1.0f);
        } else {
            a.x = a.x + (r - a.x) / (float)k;
            a.y = a.y + (g - a.y) / (float)k;
            a.z = a.z + (b - a.z) / (float)k;
            a.w = 1.0f;
        }
    }
    reinterpret_cast<float4 *>(accum)[pix] = a;
}


// ---------------------------------------------------------------- batched rendering
// Persistent kernels: a fixed grid (chosen by the autotuner) strides over the device-side count, so
// a captured graph replays them unchanged whatever the queue lengths.  Work item g of a sample is
// pixel g - first[j] of job j (first[j] <= g < first[j + 1]); its path id is its slot in the chunk.
__device__ __forceinline__ uint32_t job_of(const BatchTables &bt, uint32_t g) {
    uint32_t lo = 0, hi = bt.jobs;  // largest j with first[j] <= g
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bt.first[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void batch_generate_kernel(BatchTables bt, uint32_t chunk_base, uint32_t chunk_n,
                                                             PathQueue q, float *radiance, uint32_t *count) {
    const uint32_t stride = gridDim.x * 256;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = chunk_n;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < chunk_n; i += stride) {
        const uint32_t g = chunk_base + i, j = job_of(bt, g);
        const FrameArgs &f = bt.frames[j];
        generate_path(f, g - bt.first[j], f.sample_index0, q, i, radiance);
        q.job[i] = j;
    }
}

// Waves stride together (the loop bound is wave-uniform) so the queue appends see every lane.
template <bool kLayered>
__global__ __launch_bounds__(256) void batch_shade_kernel(BatchTables bt, PathQueue in, HitArgs hits, PathQueue out,
                                                          ShadowQueue sh, Counters c, uint32_t capacity, float *radiance) {
    const uint32_t n = min(*c.in, capacity);
    const uint32_t lane = threadIdx.x & 63u, stride = gridDim.x * 256;
    for (uint32_t base = blockIdx.x * 256 + (threadIdx.x & ~63u); base < n; base += stride) {
        const uint32_t i = base + lane;
        const bool active = i < n;
        const uint32_t j = active ? in.job[i] : 0u;
        // Batched jobs render without environment lighting.
        shade_path<false, kLayered>(bt.scenes[bt.job_scene[j]], frame_env<false>(bt.frames[j]), in, hits, out, sh, c,
                                    capacity, radiance, RestirArgs{}, NrcArgs{}, i, active, j);
    }
}

__global__ __launch_bounds__(256) void batch_resolve_kernel(ShadowQueue sh, const uint8_t *occluded, const uint32_t *count,
                                                            uint32_t capacity, float *radiance) {
    const uint32_t n = min(*count, capacity), stride = gridDim.x * 256;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        if (occluded[i]) continue;
        const uint32_t id = sh.id[i];
        radiance[3 * (size_t)id] += sh.r[i];
        radiance[3 * (size_t)id + 1] += sh.g[i];
        radiance[3 * (size_t)id + 2] += sh.b[i];
    }
}

// accum[g] (rgba32f, one pixel per work item) folds in the chunk's radiance as sample *frame_count + 1.
__global__ __launch_bounds__(256) void batch_accumulate_kernel(BatchTables bt, uint32_t chunk_base, uint32_t chunk_n,
                                                               const float *radiance, float *accum) {
    const uint32_t stride = gridDim.x * 256, k = *bt.frame_count + 1u;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < chunk_n; i += stride) {
        float r = radiance[3 * (size_t)i], g = radiance[3 * (size_t)i + 1], b = radiance[3 * (size_t)i + 2];
        if (!(isfinite(r) && isfinite(g) && isfinite(b))) r = g = b = 0.0f;
        float4 *px = reinterpret_cast<float4 *>(accum) + chunk_base + i;
        float4 a = *px;
        if (k <= 1u) {
            a = make_float4(r, g, b, 1.0f);
        } else {
            a.x = a.x + (r - a.x) / (float)k;
            a.y = a.y + (g - a.y) / (float)k;
            a.z = a.z + (b - a.z) / (float)k;
            a.w = 1.0f;
        }
        *px = a;
    }
}

// After every chunk of a sample: each job's next sample index, and the accumulated sample count.
__global__ __launch_bounds__(256) void batch_advance_kernel(BatchTables bt) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j < bt.jobs) bt.frames[j].sample_index0 += 1u;
    if (j == 0) *bt.frame_count += 1u;
}
}  // namespace

hipError_t launch_generate(const FrameArgs &f, const PathQueue &q, float *radiance, uint32_t *count, hipStream_t s) {
    const uint32_t n = f.npix * f.samples;
    hipLaunchKernelGGL(generate_kernel, dim3((n + 255) / 256), dim3(256), 0, s, f, q, radiance, count);
    return hipGetLastError();
}

hipError_t launch_shade(const RenderScene &sc, const FrameArgs &f, const PathQueue &in, const HitArgs &hits,
                        const PathQueue &out, const ShadowQueue &sh, const Counters &c, uint32_t capacity,
                        float *radiance, const NrcArgs &nrc, hipStream_t s) {
    const dim3 grid((capacity + 255) / 256);
    const bool env = frame_has_env(sc, f);
    if (sc.layered && env)
        hipLaunchKernelGGL((shade_kernel<false, true, true>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c,
                           capacity, radiance, RestirArgs{}, nrc);
    else if (sc.layered)
        hipLaunchKernelGGL((shade_kernel<false, true, false>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c,
                           capacity, radiance, RestirArgs{}, nrc);
    else if (env)
        hipLaunchKernelGGL(shade_lean_kernel<true>, grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c, capacity,
                           radiance, nrc);
    else
        hipLaunchKernelGGL(shade_lean_kernel<false>, grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c, capacity,
                           radiance, nrc);
    return hipGetLastError();
}

hipError_t launch_shade_sorted(const RenderScene &sc, const FrameArgs &f, const PathQueue &in, const HitArgs &hits,
                               const PathQueue &out, const ShadowQueue &sh, const Counters &c, uint32_t capacity,
                               float *radiance, const NrcArgs &nrc, uint32_t *perm, uint32_t *classes, hipStream_t s) {
    const dim3 grid((capacity + 255) / 256);
    hipLaunchKernelGGL(classify_kernel, grid, dim3(256), 0, s, sc, hits, c.in, capacity, perm, classes);
    hipLaunchKernelGGL((shade_sorted_kernel<false>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c, capacity, radiance,
                       nrc, perm, classes);
    hipLaunchKernelGGL((shade_sorted_kernel<true>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c, capacity, radiance,
                       nrc, perm, classes + 1);
    return hipGetLastError();
}

hipError_t launch_shade_restir(const RenderScene &sc, const FrameArgs &f, const PathQueue &in, const HitArgs &hits,
                               const PathQueue &out, const ShadowQueue &sh, const Counters &c, uint32_t capacity,
                               float *radiance, const RestirArgs &ra, const NrcArgs &nrc, hipStream_t s) {
    const dim3 grid((capacity + 255) / 256);
    const bool env = frame_has_env(sc, f);
    if (sc.layered && env)
        hipLaunchKernelGGL((shade_kernel<true, true, true>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c,
                           capacity, radiance, ra, nrc);
    else if (sc.layered)
        hipLaunchKernelGGL((shade_kernel<true, true, false>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c,
                           capacity, radiance, ra, nrc);
    else if (env)
        hipLaunchKernelGGL((shade_kernel<true, false, true>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c,
                           capacity, radiance, ra, nrc);
    else
        hipLaunchKernelGGL((shade_kernel<true, false, false>), grid, dim3(256), 0, s, sc, f, in, hits, out, sh, c,
                           capacity, radiance, ra, nrc);
    return hipGetLastError();
}

hipError_t launch_restir_initial(const RenderScene &sc, const FrameArgs &f, const PathQueue &q, const HitArgs &hits,
                                 const RestirArgs &ra, hipStream_t s) {
    if (frame_has_env(sc, f))
        hipLaunchKernelGGL(restir_initial_kernel<true>, dim3((f.npix + 255) / 256), dim3(256), 0, s, sc, f, q, hits, ra);
    else
        hipLaunchKernelGGL(restir_initial_kernel<false>, dim3((f.npix + 255) / 256), dim3(256), 0, s, sc, f, q, hits, ra);
    return hipGetLastError();
}

hipError_t launch_restir_gi_rays(const RenderScene &sc, const FrameArgs &f, const PathQueue &q, const HitArgs &hits,
                                 const RestirArgs &ra, hipStream_t s) {
    hipLaunchKernelGGL(restir_gi_rays_kernel, dim3((f.npix + 255) / 256), dim3(256), 0, s, sc, f, q, hits, ra);
    return hipGetLastError();
}

hipError_t launch_restir_gi_secondary(const RenderScene &sc, const FrameArgs &f, const RestirArgs &ra, hipStream_t s) {
    hipLaunchKernelGGL(restir_gi_secondary_kernel, dim3((f.npix + 255) / 256), dim3(256), 0, s, sc, f, ra);
    return hipGetLastError();
}

hipError_t launch_restir_gi_finish(const RenderScene &sc, const FrameArgs &f, const PathQueue &q, const HitArgs &hits,
                                   const RestirArgs &ra, hipStream_t s) {
    hipLaunchKernelGGL(restir_gi_finish_kernel, dim3((f.npix + 255) / 256), dim3(256), 0, s, sc, f, q, hits, ra);
    return hipGetLastError();
}

hipError_t launch_restir_gi_resolve(const FrameArgs &f, const RestirArgs &ra, float *radiance, hipStream_t s) {
    hipLaunchKernelGGL(restir_gi_resolve_kernel, dim3((f.npix + 255) / 256), dim3(256), 0, s, f, ra, radiance);
    return hipGetLastError();
}

hipError_t launch_resolve(const FrameArgs &f, const ShadowQueue &sh, const uint8_t *occluded, const uint32_t *count,
                          uint32_t capacity, float *radiance, const NrcArgs &nrc, hipStream_t s) {
    hipLaunchKernelGGL(resolve_kernel, dim3((capacity + 255) / 256), dim3(256), 0, s, f, sh, occluded, count, capacity,
                       radiance, nrc);
    return hipGetLastError();
}

hipError_t launch_cache_resolve(const RenderScene &sc, const NrcArgs &nrc, uint32_t capacity, float *radiance,
                                hipStream_t s) {
    hipLaunchKernelGGL(cache_resolve_kernel, dim3((capacity + 255) / 256), dim3(256), 0, s, sc, nrc, capacity, radiance);
    return hipGetLastError();
}

hipError_t launch_build_targets(const NrcArgs &nrc, hipStream_t s) {
    hipLaunchKernelGGL(build_targets_kernel, dim3((nrc.r_cap + 255) / 256), dim3(256), 0, s, nrc);
    return hipGetLastError();
}

hipError_t launch_decode_light_tree(const RenderScene &sc, float4 *out, hipStream_t s) {
    if (sc.tree_node_count == 0) return hipSuccess;
    hipLaunchKernelGGL(decode_light_tree_kernel, dim3((sc.tree_node_count + 255) / 256), dim3(256), 0, s, sc, out);
    return hipGetLastError();
}

hipError_t launch_zero_f4(float4 *p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_f4_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_bake_metal(float *out, hipStream_t s) {
    hipLaunchKernelGGL(bake_metal_kernel, dim3((129 * 129 + 63) / 64), dim3(64), 0, s, out);
    return hipGetLastError();
}

hipError_t launch_reset_counters(uint32_t *counters, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, counters, n);
    return hipGetLastError();
}

hipError_t launch_accumulate(const FrameArgs &f, const float *radiance, float *accum, uint32_t frame_index,
                             hipStream_t s) {
    hipLaunchKernelGGL(accumulate_kernel, dim3((f.npix + 255) / 256), dim3(256), 0, s, f, radiance, accum, frame_index);
    return hipGetLastError();
}

hipError_t launch_batch_generate(const BatchTables &bt, uint32_t chunk_base, uint32_t chunk_n, const PathQueue &q,
                                 float *radiance, uint32_t *count, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(batch_generate_kernel, dim3(grid), dim3(256), 0, s, bt, chunk_base, chunk_n, q, radiance, count);
    return hipGetLastError();
}

hipError_t launch_batch_shade(const BatchTables &bt, const PathQueue &in, const HitArgs &hits, const PathQueue &out,
                              const ShadowQueue &sh, const Counters &c, uint32_t capacity, float *radiance, bool layered,
                              uint32_t grid, hipStream_t s) {
    if (layered)
        hipLaunchKernelGGL((batch_shade_kernel<true>), dim3(grid), dim3(256), 0, s, bt, in, hits, out, sh, c, capacity, radiance);
    else
        hipLaunchKernelGGL((batch_shade_kernel<false>), dim3(grid), dim3(256), 0, s, bt, in, hits, out, sh, c, capacity, radiance);
    return hipGetLastError();
}

hipError_t launch_batch_resolve(const ShadowQueue &sh, const uint8_t *occluded, const uint32_t *count, uint32_t capacity,
                                float *radiance, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(batch_resolve_kernel, dim3(grid), dim3(256), 0, s, sh, occluded, count, capacity, radiance);
    return hipGetLastError();
}

hipError_t launch_batch_accumulate(const BatchTables &bt, uint32_t chunk_base, uint32_t chunk_n, const float *radiance,
                                   float *accum, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(batch_accumulate_kernel, dim3(grid), dim3(256), 0, s, bt, chunk_base, chunk_n, radiance, accum);
    return hipGetLastError();
}

hipError_t launch_batch_advance(const BatchTables &bt, hipStream_t s) {
    hipLaunchKernelGGL(batch_advance_kernel, dim3((bt.jobs + 255) / 256), dim3(256), 0, s, bt);
    return hipGetLastError();
}

}  // namespace vt
