// openpbr_device.h -- device code of the layered BSDF subset the GLOSSY / OPENPBR materials use,
// included by render.hip inside its namespace (it uses V3, S4, kPi from there).
//
// The reference evaluates the vendored Adobe OpenPBR (pinned c91aad1) through prepare_openpbr_lanes /
// prepared_eval / prepared_pdf / prepared_sample (main.swift:2485-2645).  This is that BSDF for
// fuzz = subsurface = thin film = dispersion = anisotropy = 0, base_diffuse_roughness = 0 and the
// materials' coat (weight from the material, roughness 0.12, IOR 1.6, white, darkening 1:
// main.swift:2494-2495, Vendor/OpenPBR/openpbr_resolved_inputs.h:123-127): the coating lobe
// (impl/openpbr_coating_lobe.h; the fuzz lobe at weight 0 passes through) with its GGX reflection lobe
// (impl/openpbr_generic_minimal_microfacet_lobe.h) over
// the comprehensive GGX microfacet reflection/transmission lobe with dielectric Fresnel and F82-tint
// metal (Vendor/OpenPBR/impl/openpbr_comprehensive_microfacet_lobe.h,
// openpbr_reflection_transmission_coefficient.h), the dielectric and metal multiple-scattering lobes
// (openpbr_microfacet_multiple_scattering_lobes.h), the specular-compensated diffuse lobe
// (openpbr_diffuse_lobe.h) and the aggregate lobe's selection, pdf and sample
// (openpbr_aggregate_lobe.h), with the setup of openpbr_prepare_lobes (impl/openpbr_bsdf.h:289-860).
// The ideal-metal energy lookups are the reference's 129 x 129 table (Shaders/MetalEnergy.metal:
// 2108-2119).  Colours carry the path's four wavelength lanes; lobe and reflection/transmission
// probabilities use lanes 0-2, as the reference's `a` BSDF does (main.swift:2633-2640).
#pragma once

struct EnergyTables {
    const float *metal;  // 129 x 129 complements, then 129 averages
    const float *luts;   // 5 x 32768: dielectric, dielectric avg, ratio, opaque, opaque avg
};
enum : int { kLutDielectric = 0, kLutDielectricAvg = 1, kLutRatio = 2, kLutOpaque = 3, kLutOpaqueAvg = 4 };
constexpr float kLargestBelowOne = 0.999999940f;
constexpr float kRcpPi = 0.318309886f, kRcpTwoPi = 0.159154943f;

__device__ __forceinline__ float sqf(float x) { return x * x; }
__device__ __forceinline__ float max3(S4 a) { return fmaxf(a.v[0], fmaxf(a.v[1], a.v[2])); }
__device__ __forceinline__ S4 sadd(S4 a, S4 b) {
    return S4{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}};
}
__device__ __forceinline__ float safe_div(float a, float b, float fallback) { return b != 0.0f ? a / b : fallback; }

struct Basis {
    V3 t, b, n;
};
// openpbr_make_basis(normal): the branchless frame of Duff et al.
__device__ __forceinline__ Basis make_basis(V3 n) {
    Basis B;
    B.n = n;
    float s = n.z < 0.0f ? -1.0f : 1.0f;
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    B.t = v3(b, s + n.y * n.y * a, -n.y);
    B.b = v3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    return B;
}
__device__ __forceinline__ V3 to_local(const Basis &B, V3 d) { return v3(dot(d, B.t), dot(d, B.b), dot(d, B.n)); }
__device__ __forceinline__ V3 to_world(const Basis &B, V3 d) { return B.t * d.x + B.b * d.y + B.n * d.z; }

// openpbr_make_basis(normal, tangent, handedness 1): the tangent made perpendicular to the normal.
__device__ __forceinline__ Basis make_basis_t(V3 n, V3 t) {
    Basis B;
    B.n = n;
    B.t = normalize(t - n * dot(t, n));
    B.b = cross(n, B.t);
    return B;
}

// Anisotropic GGX and Smith G1 with alphas along the basis' tangent (ax) and bitangent (ay).
__device__ __forceinline__ float ggx_d(V3 h, float ax, float ay) {
    return 1.0f / (kPi * ax * ay * sqf(sqf(h.x / ax) + sqf(h.y / ay) + sqf(h.z)));
}
__device__ __forceinline__ float smith_g1(V3 v, float ax, float ay) {
    float z2 = sqf(v.z);
    if (z2 == 0.0f) return 0.0f;
    return 2.0f / (1.0f + sqrtf(1.0f + (sqf(ax * v.x) + sqf(ay * v.y)) / z2));
}
// openpbr_sample_aniso_ggx_smith_vndf: spherical-cap sampling of the visible normals.
__device__ __forceinline__ V3 sample_vndf(float ax, float ay, V3 wi, float r0, float r1) {
    V3 ih = normalize(v3(ax * wi.x, ay * wi.y, wi.z));
    float c = fminf(ih.z, 1.0f);
    float ang = kTwoPi * r0;
    float opc = 1.0f + c, hz = (1.0f - r1) * opc, omz = r1 * opc, omc = 1.0f - c, opz = hz + omc;
    float tc = sqrtf(opz * omz);
    float sn, cs;
    sincosf(ang, &sn, &cs);
    V3 m = v3(tc * cs + ih.x, tc * sn + ih.y, hz);
    return normalize(v3(ax * m.x, ay * m.y, m.z));
}
__device__ __forceinline__ float fresnel(float eta, float ci) {
    if (eta == 1.0f) return 0.0f;
    float st2 = (1.0f - sqf(ci)) / sqf(eta);
    if (st2 >= 1.0f) return 1.0f;
    float ct = sqrtf(1.0f - st2), eci = eta * ci, ect = eta * ct;
    float rs = sqf((ci - ect) / (ci + ect)), rp = sqf((ct - eci) / (ct + eci));
    return 0.5f * (rs + rp);
}
__device__ __forceinline__ float f0_from_ior(float eta) { return sqf((eta - 1.0f) / (eta + 1.0f)); }
__device__ __forceinline__ float apply_specular_weight_to_ior(float eta, float w) {
    float f0 = fminf(f0_from_ior(eta) * w, 0.9999f);
    float s = sqrtf(f0);
    float ext = (1.0f + s) / (1.0f - s);
    return eta < 1.0f ? 1.0f / ext : ext;
}
// F82-tint metal Fresnel with a white tint (b factor 0): Schlick from f0.
__device__ __forceinline__ S4 metal_f82(S4 f0, float c) {
    float omc = 1.0f - c, p5 = omc * omc * omc * omc * omc;
    S4 r;
    for (int k = 0; k < 4; ++k) r.v[k] = fminf(fmaxf(f0.v[k] + ((1.0f - f0.v[k]) - 0.0f * c * omc) * p5, 0.0f), 1.0f);
    return r;
}
__device__ __forceinline__ S4 metal_average_f82(S4 f0) {
    S4 r;
    for (int k = 0; k < 4; ++k)
        r.v[k] = fminf(fmaxf(f0.v[k] + (1.0f - f0.v[k]) * (1.0f / 21.0f) - 0.0f * (1.0f / 126.0f), 0.0f), 1.0f);
    return r;
}
__device__ __forceinline__ bool refract_dir(V3 wi, V3 n, float ci, float eta, V3 &wt) {
    float inv = 1.0f / eta;
    float s2t = sqf(inv) * fmaxf(0.0f, 1.0f - sqf(ci));
    if (s2t >= 1.0f) return false;
    float ct = sqrtf(1.0f - s2t);
    wt = normalize(wi * -inv + n * (inv * ci - ct));
    return true;
}

// ---------------------------------------------------------------- energy-table lookups
__device__ __forceinline__ float ior_index(float ior) {
    const float inv_range = 1.0f / (2.5f - 1.0f);
    if (ior < 1.0f) return 15.0f - ((1.0f / ior - 1.0f) * inv_range) * 15.0f;
    return 16.0f + ((ior - 1.0f) * inv_range) * 15.0f;
}
__device__ __forceinline__ float clamp_index(float x) { return fminf(fmaxf(x, 0.0f), 31.0f); }
__device__ __forceinline__ float lookup2(const EnergyTables &E, int id, float x, float y) {
    const float *t = E.luts + id * 32768;
    x = clamp_index(x);
    y = clamp_index(y);
    int x0 = (int)floorf(x), y0 = (int)floorf(y), x1 = min(x0 + 1, 31), y1 = min(y0 + 1, 31);
    float tx = x - (float)x0, ty = y - (float)y0;
    float v0 = t[x0 * 32 + y0] * (1.0f - tx) + t[x1 * 32 + y0] * tx;
    float v1 = t[x0 * 32 + y1] * (1.0f - tx) + t[x1 * 32 + y1] * tx;
    return v0 * (1.0f - ty) + v1 * ty;
}
__device__ __forceinline__ float lookup3(const EnergyTables &E, int id, float x, float y, float z) {
    const float *t = E.luts + id * 32768;
    x = clamp_index(x);
    y = clamp_index(y);
    z = clamp_index(z);
    int x0 = (int)floorf(x), y0 = (int)floorf(y), z0 = (int)floorf(z);
    int x1 = min(x0 + 1, 31), y1 = min(y0 + 1, 31), z1 = min(z0 + 1, 31);
    float tx = x - (float)x0, ty = y - (float)y0, tz = z - (float)z0;
    auto L = [&](int a, int b, int c) { return t[a * 1024 + b * 32 + c]; };
    float v00 = L(x0, y0, z0) * (1.0f - tx) + L(x1, y0, z0) * tx;
    float v01 = L(x0, y0, z1) * (1.0f - tx) + L(x1, y0, z1) * tx;
    float v10 = L(x0, y1, z0) * (1.0f - tx) + L(x1, y1, z0) * tx;
    float v11 = L(x0, y1, z1) * (1.0f - tx) + L(x1, y1, z1) * tx;
    float v0 = v00 * (1.0f - ty) + v10 * ty, v1 = v01 * (1.0f - ty) + v11 * ty;
    return v0 * (1.0f - tz) + v1 * tz;
}
__device__ __forceinline__ float alpha_index(float alpha) { return sqrtf(alpha) * 31.0f; }
__device__ __forceinline__ float extrapolate_ior(float v, float ior) {
    if (ior > 2.5f || ior < 0.4f) {
        float f0max = f0_from_ior(2.5f);
        return (1.0f - (f0_from_ior(ior) - f0max) / (1.0f - f0max)) * v;
    }
    return v;
}
__device__ __forceinline__ float diel_complement(const EnergyTables &E, float ior, float alpha, float c) {
    return lookup3(E, kLutDielectric, ior_index(ior), alpha_index(alpha), c * 31.0f);
}
__device__ __forceinline__ float opaque_complement(const EnergyTables &E, float ior, float alpha, float c) {
    return extrapolate_ior(lookup3(E, kLutOpaque, ior_index(ior), alpha_index(alpha), c * 31.0f), ior);
}
__device__ __forceinline__ float metal_complement(const EnergyTables &E, float alpha, float c) {
    const float *t = E.metal;
    float x = fminf(fmaxf(sqrtf(alpha) * 128.0f, 0.0f), 128.0f);
    float y = sqrtf(fminf(fmaxf(c, 0.0f), 1.0f)) * 128.0f;
    int ix = min((int)x, 127), iy = min((int)y, 127);
    float fy = y - (float)iy, fx = x - (float)ix;
    float a = t[ix * 129 + iy] + (t[ix * 129 + iy + 1] - t[ix * 129 + iy]) * fy;
    float b = t[(ix + 1) * 129 + iy] + (t[(ix + 1) * 129 + iy + 1] - t[(ix + 1) * 129 + iy]) * fy;
    return a + (b - a) * fx;
}
__device__ __forceinline__ float metal_avg_complement(const EnergyTables &E, float alpha) {
    const float *t = E.metal + 129 * 129;
    float x = fminf(fmaxf(sqrtf(alpha) * 128.0f, 0.0f), 128.0f);
    int ix = min((int)x, 127);
    return t[ix] + (t[ix + 1] - t[ix]) * (x - (float)ix);
}
__device__ __forceinline__ float clamp_avg(float v) { return fmaxf(v, 1e-12f); }
// mix(x, y, a) as the shading language defines it.
__device__ __forceinline__ float mixf(float x, float y, float a) { return x + (y - x) * a; }
__device__ __forceinline__ float average_fresnel(float eta) {
    if (eta > 1.0f) return (eta - 1.0f) / (4.08567f + 1.00071f * eta);
    return 0.997118f + 0.1014f * eta - 0.965241f * (eta * eta) - 0.130607f * (eta * eta) * eta;
}
constexpr float kCoatIor = 1.6f, kCoatRoughness = 0.12f;
// fuzz_color white and fuzz_roughness 0.5: OpenPBR's defaults (openpbr_resolved_inputs.h:131-133).
constexpr float kFuzzRoughness = 0.5f;

// ---------------------------------------------------------------- fuzz (Disney sheen LTC)
// openpbr_disney_sheen_fetch_coeffs, array mode (openpbr_fuzz_lobe.h:147-176): bilinear in the 32 x 32
// table over (roughness, cosine), which follows the dielectric tables in the energy buffer.
__device__ __forceinline__ void ltc_fetch(const EnergyTables &E, float cos_t, float out[3]) {
    const float *T = E.luts + kLtcOffset;
    const float row = fmaxf(0.0f, fminf(kFuzzRoughness, kLargestBelowOne)) * 31.0f;
    const float col = fmaxf(0.0f, fminf(cos_t, kLargestBelowOne)) * 31.0f;
    const float r = floorf(row), c = floorf(col), rf = row - r, cf = col - c;
    const int ri = (int)r, ci = (int)c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a = mixf(T[(ri * 32 + ci) * 3 + k], T[(ri * 32 + ci + 1) * 3 + k], cf);
        const float b = mixf(T[((ri + 1) * 32 + ci) * 3 + k], T[((ri + 1) * 32 + ci + 1) * 3 + k], cf);
        out[k] = mixf(a, b, rf);
    }
}
// openpbr_disney_sheen_rotate_vector: about z by the azimuth of (x, y).
__device__ __forceinline__ V3 ltc_rotate(V3 v, float x, float y) {
    const float r2 = x * x + y * y;
    if (r2 == 0.0f) return v;
    const float inv = rsqrtf(r2), sn = y * inv, cs = x * inv;
    return v3(cs * v.x + sn * -v.y, cs * v.y + sn * v.x, v.z);
}
// openpbr_disney_sheen_eval_ltc: the LTC's density in the aligned frame.
__device__ __forceinline__ float ltc_eval(V3 wi, float ai, float bi) {
    const V3 o = v3(ai * wi.x + bi * wi.z, ai * wi.y, wi.z);
    const float z = fmaxf(0.0f, wi.z), l2 = dot(o, o);
    if (l2 == 0.0f || ai == 0.0f || z == 0.0f) return 0.0f;
    const float k = ai / l2;
    return kRcpPi * (z * k) * k;
}

// ---------------------------------------------------------------- prepared BSDF
struct Pbr {
    V3 view, n;
    Basis sb;                // specular_anisotropy_basis_ff
    float ax, ay, iso_alpha; // the specular lobe's alphas (floored at 1e-6) and specular_alpha
    float eta_trans, eta_opaque, eta_refr;
    float scale_trans_refl, scale_opaque_refl, transmission, metal_amount;
    S4 f0;
    float d_ratio, d_ec, d_scale_refl, d_scale_trans;
    S4 m_scale;
    float m_ec;
    S4 diffuse_albedo;
    float diffuse_comp;
    float w[4], total;       // lobe weights: specular, dielectric MMS, metal MMS, diffuse
    float coat, c_prob;      // coat presence and coat reflection probability (entering a coated surface)
    float c_alpha;           // the coat's alpha (its roughness 0.12 adjusted for the fuzz over it)
    S4 c_scale;              // in_base_layer_scale: (1 - in_reflected) x darkening, per lane
    float fuzz, f_refl, f_prob;  // fuzz presence (0 off or inside), view_reflected, sheen probability
    bool d_on, m_on, c_on;
};

__device__ __forceinline__ void pbr_coefficients(const Pbr &p, float c, S4 &refl, S4 &trans) {
    float ft = fresnel(p.eta_trans, c), fo = fresnel(p.eta_opaque, c);
    S4 fm = metal_f82(p.f0, c);
    for (int k = 0; k < 4; ++k) {
        refl.v[k] = p.scale_trans_refl * ft + p.scale_opaque_refl * fo + p.metal_amount * fm.v[k];
        trans.v[k] = p.transmission * (1.0f - ft);
    }
}
__device__ __forceinline__ void pbr_probabilities(const Pbr &p, float c, S4 &refl, S4 &trans, float &pr, float &pt) {
    pbr_coefficients(p, c, refl, trans);
    float rm = max3(refl), tm = max3(trans);
    pr = safe_div(rm, rm + tm, 0.0f);
    pt = safe_div(tm, rm + tm, 0.0f);
}

// prepare_openpbr_lanes for the subset.  type: 1 GLOSSY (metalness 1) or 4 OPENPBR; n: shading normal
// on the arriving side; wo: toward the viewer; inside: hit from the back; rho: base colour per lane.
__device__ __forceinline__ void pbr_prepare(Pbr &p, const EnergyTables &E, uint32_t type, float roughness, float ior_in,
                                            float metal_in, float transw, float coat, float aniso, V3 tangent,
                                            float fuzz, V3 n, V3 wo, bool inside_flag, S4 rho) {
    const float rough = fmaxf(0.03f, roughness);
    const float metalness = type == kGlossy ? 1.0f : metal_in;
    const float ior = fmaxf(1.01f, ior_in);
    V3 outward = transw > 0.0f && inside_flag ? -n : n;
    // geometry_basis (main.swift:2508-2509): the tangent frame where the tangent is usable.  Only an
    // anisotropic lobe depends on the frame's azimuth; isotropic ones keep the normal's own frame.
    Basis gb = aniso > 0.0f && dot(tangent, tangent) > 0.5f && fabsf(dot(tangent, outward)) < 0.99f
                   ? make_basis_t(outward, tangent) : make_basis(outward);
    p.view = wo;
    const bool back = dot(wo, gb.n) < 0.0f;
    p.sb = gb;
    if (back) {
        p.sb.t = -gb.t;
        p.sb.b = -gb.b;
        p.sb.n = -gb.n;
    }
    p.n = p.sb.n;
    // IORs (openpbr_bsdf.h:339-418): entering a coated surface, the base sits under the coat, whose IOR
    // is blended by the coat weight; hit from inside, the coat is ignored.
    const bool enter_coat = coat > 0.0f && !back;
    const float wci = enter_coat ? mixf(1.0f, kCoatIor, coat) : 1.0f;
    const float wsi = wci * apply_specular_weight_to_ior(ior / wci, 1.0f);
    float r_refr, r_trans, r_opaque, rel_coat = 1.0f;
    if (back) {
        r_refr = 1.0f / wsi;
        r_trans = r_refr;
        r_opaque = wsi / 1.0f;
    } else {
        r_refr = wsi / 1.0f;
        r_trans = !enter_coat ? r_refr : wci > wsi && wsi >= 1.0f ? wci / wsi : wsi / wci;
        if (enter_coat) rel_coat = kCoatIor / 1.0f;
        r_opaque = r_trans;
    }
    r_refr = r_refr >= 1.0f ? fmaxf(r_refr, 1.0001f) : fminf(r_refr, 1.0f / 1.0001f);
    // Roughnesses (openpbr_bsdf.h:424-460): the coat's adjusted for the fuzz over it, the base's for the coat.
    const float fuzz_adjust = (3.0f * (1.0f / 3.0f)) * kFuzzRoughness * 0.005f;
    const float c_rough = mixf(
        kCoatRoughness, sqrtf(sqrtf(fminf(1.0f, sqf(sqf(kCoatRoughness)) + fuzz_adjust * sqf(sqf(kFuzzRoughness))))), fuzz);
    const float adjust = mixf(1.0f, 0.0f, rel_coat >= 1.0f ? 1.0f / rel_coat : rel_coat);
    const float adj_rough =
        mixf(rough, sqrtf(sqrtf(fminf(1.0f, sqf(sqf(rough)) + adjust * sqf(sqf(c_rough))))), coat);
    const float salpha = adj_rough * adj_rough;
    p.iso_alpha = salpha;
    // openpbr_compute_anisotropic_alpha (openpbr_lobe_utils.h:41-72), floored at 1e-6.
    p.ax = p.ay = salpha;
    if (aniso > 0.0f) {
        p.ax = salpha * sqrtf(2.0f / (1.0f + sqf(1.0f - aniso)));
        p.ay = (1.0f - aniso) * p.ax;
    }
    p.ax = fmaxf(p.ax, 1e-6f);
    p.ay = fmaxf(p.ay, 1e-6f);
    const float dielectric = 1.0f - metalness;
    const float opaque = (1.0f - transw) * dielectric, trans = dielectric * transw, dmetal = metalness * 1.0f;
    const S4 avg_f = metal_average_f82(rho);
    p.eta_trans = r_trans;
    p.eta_opaque = r_opaque;
    p.eta_refr = r_refr;
    p.scale_trans_refl = trans;
    p.scale_opaque_refl = opaque;
    p.transmission = trans;
    p.metal_amount = dmetal;
    p.f0 = rho;
    p.d_scale_refl = trans;
    p.d_scale_trans = trans;
    const float min_alpha = 0.04f * 0.04f;
    p.d_on = salpha >= min_alpha;
    p.m_on = salpha >= min_alpha;
    const float idotn = dot(wo, p.n);
    p.d_ratio = p.d_on ? lookup2(E, kLutRatio, ior_index(r_refr), alpha_index(salpha)) : 0.0f;
    p.d_ec = p.d_on ? diel_complement(E, r_refr, salpha, idotn) : 0.0f;
    for (int k = 0; k < 4; ++k) p.m_scale.v[k] = sqf(avg_f.v[k]) * dmetal;
    p.m_ec = p.m_on ? metal_complement(E, salpha, idotn) : 0.0f;
    p.diffuse_albedo = scale(rho, opaque);
    p.diffuse_comp = opaque_complement(E, r_opaque, salpha, idotn) /
                     clamp_avg(extrapolate_ior(lookup2(E, kLutOpaqueAvg, ior_index(r_opaque), alpha_index(salpha)), r_opaque));
    S4 rc, tc;
    pbr_coefficients(p, idotn, rc, tc);
    p.w[0] = max3(sadd(rc, tc));
    p.w[1] = p.d_on ? (p.d_scale_refl * p.d_ratio + p.d_scale_trans * (1.0f - p.d_ratio)) * p.d_ec : 0.0f;
    p.w[2] = p.m_on ? max3(p.m_scale) * p.m_ec : 0.0f;
    p.w[3] = max3(scale(p.diffuse_albedo, p.diffuse_comp));
    p.total = p.w[0] + p.w[1] + p.w[2] + p.w[3];
    // The coating lobe (openpbr_bsdf.h:498-552, 831-840; openpbr_coating_lobe.h:137-185): the darkening
    // Delta = (1 - K) / (1 - E_b K) lerped in by the coat weight, the view side's coat reflectance from
    // the opaque-dielectric energy table, and the coat-or-base choice probability.
    p.c_on = enter_coat;
    p.coat = coat;
    p.c_prob = 0.0f;
    p.c_scale = s4(1.0f);
    p.c_alpha = fmaxf(sqf(c_rough), 1e-6f);
    float below = p.total;  // the coating lobe's estimated contribution, for the sheen probability
    if (enter_coat) {
        const float ks = average_fresnel(rel_coat);
        const float kr = 1.0f - (1.0f - average_fresnel(rel_coat)) / sqf(rel_coat);
        const float so = average_fresnel(r_opaque), st = average_fresnel(r_trans);
        const float spec_base = (opaque * so + (0.0f + trans) * st) + metalness;
        const float kk = mixf(ks, kr, mixf(1.0f, adj_rough, spec_base));
        const float refl = idotn > 0.0f ? (1.0f - opaque_complement(E, rel_coat, sqf(c_rough), idotn)) * coat : 0.0f;
        for (int k = 0; k < 4; ++k) {
            const float eb = ((dmetal * avg_f.v[k] + opaque * mixf(rho.v[k], 1.0f, so)) + 0.0f) + trans;
            p.c_scale.v[k] = (1.0f - refl) * mixf(1.0f, (1.0f - kk) / (1.0f - eb * kk), coat * 1.0f);
        }
        p.c_prob = safe_div(refl, refl + p.total * max3(p.c_scale), 0.5f);
        below = refl + p.total * max3(p.c_scale);
    }
    // The fuzz lobe (openpbr_bsdf.h:842-858, openpbr_fuzz_lobe.h:395-430): presence `fuzz` outside, 0
    // inside; its normal blends the base and coat normals, which coincide here.
    p.fuzz = fuzz > 0.0f && !back ? fuzz : 0.0f;
    p.f_refl = 0.0f;
    p.f_prob = 0.0f;
    if (p.fuzz > 0.0f) {
        const float cos_o = idotn;  // the view's cosine in the fuzz frame (the normal's own)
        float coef[3];
        ltc_fetch(E, cos_o, coef);
        p.f_refl = cos_o <= 0.0f ? 0.0f : p.fuzz * coef[2];
        p.f_prob = safe_div(p.f_refl, p.f_refl + below * (1.0f - p.f_refl), 0.5f);
    }
}

__device__ __forceinline__ S4 spec_value(const Pbr &p, const EnergyTables &, V3 wi) {
    const V3 wo = p.view;
    const float idotn = dot(p.n, wo), odotn = dot(p.n, wi);
    if (idotn * odotn > 0.0f) {
        V3 h = normalize(wo + wi);
        float idoth = dot(wo, h);
        float D = ggx_d(to_local(p.sb, h), p.ax, p.ay);
        float other = 1.0f / (4.0f * idotn);
        S4 F, T;
        pbr_coefficients(p, fabsf(idoth), F, T);
        float G = smith_g1(to_local(p.sb, wo), p.ax, p.ay) * smith_g1(to_local(p.sb, wi), p.ax, p.ay);
        return scale(F, G * D * other);
    }
    if (idotn * odotn < 0.0f) {
        V3 h = -normalize(wo + wi * p.eta_refr);
        bool valid = (p.eta_refr >= 1.0f && dot(p.n, h) > 0.0f) || (p.eta_refr < 1.0f && dot(p.n, h) < 0.0f);
        if (!valid) return s4(0.0f);
        float idoth = dot(wo, h), odoth = dot(wi, h);
        if (idoth * odoth >= 0.0f) return s4(0.0f);
        float D = ggx_d(to_local(p.sb, h), p.ax, p.ay);
        float other = sqf(p.eta_refr) * fabsf(idoth) * fabsf(odoth) / (sqf(idoth + p.eta_refr * odoth) * idotn);
        S4 F, T;
        pbr_coefficients(p, fabsf(idoth), F, T);
        float G = smith_g1(to_local(p.sb, wo), p.ax, p.ay) * smith_g1(to_local(p.sb, wi), p.ax, p.ay);
        return scale(T, G * D * other);
    }
    return s4(0.0f);
}
__device__ __forceinline__ float spec_pdf(const Pbr &p, V3 wi) {
    const V3 wo = p.view;
    const float idotn = dot(p.n, wo), odotn = dot(p.n, wi);
    S4 F, T;
    float pr, pt;
    if (idotn * odotn > 0.0f) {
        V3 h = normalize(wo + wi);
        float idoth = dot(wo, h), odoth = dot(wi, h);
        if (idoth * odoth < 0.0f) return 0.0f;
        float D = ggx_d(to_local(p.sb, h), p.ax, p.ay);
        float other = idoth / (4.0f * odoth * idotn);
        pbr_probabilities(p, fabsf(idoth), F, T, pr, pt);
        return smith_g1(to_local(p.sb, wo), p.ax, p.ay) * (pr * D * other);
    }
    if (idotn * odotn < 0.0f) {
        V3 h = -normalize(wo + wi * p.eta_refr);
        bool valid = (p.eta_refr >= 1.0f && dot(p.n, h) > 0.0f) || (p.eta_refr < 1.0f && dot(p.n, h) < 0.0f);
        if (!valid) return 0.0f;
        float idoth = dot(wo, h), odoth = dot(wi, h);
        if (idoth * odoth >= 0.0f) return 0.0f;
        float D = ggx_d(to_local(p.sb, h), p.ax, p.ay);
        float other = sqf(p.eta_refr) * fabsf(idoth) * fabsf(odoth) / (sqf(idoth + p.eta_refr * odoth) * idotn);
        pbr_probabilities(p, fabsf(idoth), F, T, pr, pt);
        return smith_g1(to_local(p.sb, wo), p.ax, p.ay) * (pt * D * other);
    }
    return 0.0f;
}
__device__ __forceinline__ float d_factors(const Pbr &p, const EnergyTables &E, float odotn, bool with_ratio) {
    float ior = p.eta_refr, ratio = p.d_ratio;
    if (!(odotn > 0.0f)) {
        odotn = -odotn;
        ior = 1.0f / p.eta_refr;
        ratio = 1.0f - p.d_ratio;
    }
    float t = p.d_ec * diel_complement(E, ior, p.iso_alpha, odotn) /
              clamp_avg(lookup2(E, kLutDielectricAvg, ior_index(ior), alpha_index(p.iso_alpha)));
    if (with_ratio) t *= ratio;
    return fminf(t, safe_div(1.0f, odotn, 0.0f));
}
__device__ __forceinline__ S4 dmms_value(const Pbr &p, const EnergyTables &E, V3 wi) {
    if (!p.d_on) return s4(0.0f);
    float odotn = dot(wi, p.n);
    float u = d_factors(p, E, odotn, true) * kRcpPi * fabsf(odotn);
    return s4((odotn > 0.0f ? p.d_scale_refl : p.d_scale_trans) * u);
}
__device__ __forceinline__ float dmms_pdf(const Pbr &p, V3 wi) {
    if (!p.d_on) return 0.0f;
    return dot(wi, p.n) > 0.0f ? kRcpTwoPi * p.d_ratio : kRcpTwoPi * (1.0f - p.d_ratio);
}
__device__ __forceinline__ float m_factors(const Pbr &p, const EnergyTables &E, float odotn) {
    float t = p.m_ec * metal_complement(E, p.iso_alpha, odotn) / clamp_avg(metal_avg_complement(E, p.iso_alpha));
    return fminf(t, safe_div(1.0f, odotn, 0.0f));
}
__device__ __forceinline__ S4 mmms_value(const Pbr &p, const EnergyTables &E, V3 wi) {
    if (!p.m_on) return s4(0.0f);
    float odotn = dot(wi, p.n);
    if (!(odotn > 0.0f)) return s4(0.0f);
    return scale(p.m_scale, m_factors(p, E, odotn) * kRcpPi * odotn);
}
__device__ __forceinline__ float mmms_pdf(const Pbr &p, V3 wi) {
    if (!p.m_on) return 0.0f;
    return dot(wi, p.n) > 0.0f ? kRcpTwoPi : 0.0f;
}
// openpbr_f_EON at roughness 0 (approximate FON albedo): Lambert plus the 1e-7 multiple-scatter floor.
__device__ __forceinline__ S4 eon0(S4 rho) {
    S4 r;
    for (int k = 0; k < 4; ++k) {
        float fss = rho.v[k] * kRcpPi;
        float fms = (rho.v[k] * rho.v[k] * kRcpPi) * 1e-7f * 1e-7f / 1e-7f;
        r.v[k] = fss + fms;
    }
    return r;
}
__device__ __forceinline__ S4 diffuse_value(const Pbr &p, const EnergyTables &E, V3 wi) {
    float co = dot(wi, p.n);
    if (co <= 0.0f) return s4(0.0f);
    float comp = p.diffuse_comp * opaque_complement(E, p.eta_opaque, p.iso_alpha, co);
    return scale(scale(eon0(p.diffuse_albedo), comp), co);
}
__device__ __forceinline__ float diffuse_pdf(const Pbr &p, V3 wi) { return fmaxf(dot(wi, p.n), 0.0f) * kRcpPi; }

// The base layer (the aggregate lobe): f |cos| per lane.
__device__ __forceinline__ S4 base_eval(const Pbr &p, const EnergyTables &E, V3 wi) {
    return sadd(sadd(sadd(spec_value(p, E, wi), dmms_value(p, E, wi)), mmms_value(p, E, wi)), diffuse_value(p, E, wi));
}
__device__ __forceinline__ float base_pdf(const Pbr &p, V3 wi) {
    float s = p.w[0] * spec_pdf(p, wi) + p.w[1] * dmms_pdf(p, wi) + p.w[2] * mmms_pdf(p, wi) + p.w[3] * diffuse_pdf(p, wi);
    return p.total > 0.0f ? s / p.total : 0.0f;
}
// The aggregate lobe's sample: lobe by weight, the lobe's own weight and pdf, then every other lobe
// folded in.
__device__ __forceinline__ bool base_sample(const Pbr &p, const EnergyTables &E, float r0, float r1, float r2, V3 &wi,
                                            S4 &weight, float &pdf) {
    if (p.total <= 1.175494351e-38f) return false;
    float x = r0 * p.total, acc = 0.0f;
    int lobe = -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float wk = p.w[k];
        if (lobe < 0 && wk > 0.0f) {
            if (x < acc + wk) {
                lobe = k;
                x = fminf((x - acc) / wk, kLargestBelowOne);
            } else {
                acc += wk;
            }
        }
    }
    if (lobe < 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (p.w[k] > 0.0f) lobe = k;
        x = kLargestBelowOne;
    }
    const V3 wo = p.view;
    V3 d;
    S4 w;
    float pd;
    if (lobe == 0) {
        float idotn = dot(wo, p.n);
        V3 h = to_world(p.sb, sample_vndf(p.ax, p.ay, to_local(p.sb, wo), x, r1));
        float idoth = dot(wo, h);
        if (idoth < 0.0f) return false;
        S4 F, T;
        float pr, pt;
        pbr_probabilities(p, idoth, F, T, pr, pt);
        if (pr + pt == 0.0f) return false;
        if (r2 < pr) {
            d = normalize(-wo + h * (2.0f * idoth));
            float odotn = dot(p.n, d);
            if (idotn * odotn <= 0.0f) return false;
            float odoth = dot(d, h);
            w = scale(scale(F, 1.0f / pr), smith_g1(to_local(p.sb, d), p.ax, p.ay));
            float D = ggx_d(to_local(p.sb, h), p.ax, p.ay);
            if (D <= 0.0f) return false;
            pd = pr * D * smith_g1(to_local(p.sb, wo), p.ax, p.ay) * idoth / (4.0f * odoth * idotn);
        } else {
            if (!refract_dir(wo, h, idoth, p.eta_refr, d)) return false;
            float odotn = dot(d, p.n);
            if (idotn * odotn >= 0.0f) return false;
            float odoth = dot(d, h);
            w = scale(scale(T, 1.0f / pt), smith_g1(to_local(p.sb, d), p.ax, p.ay));
            float D = ggx_d(to_local(p.sb, h), p.ax, p.ay);
            if (D <= 0.0f) return false;
            pd = pt * D * smith_g1(to_local(p.sb, wo), p.ax, p.ay) * sqf(p.eta_refr) * fabsf(idoth) * fabsf(odoth) /
                 (sqf(idoth + p.eta_refr * odoth) * idotn);
        }
    } else if (lobe == 1) {
        float rx = x, z;
        if (rx < p.d_ratio) {
            rx = fminf(rx / p.d_ratio, kLargestBelowOne);
            z = rx;
        } else {
            rx = fminf((rx - p.d_ratio) / (1.0f - p.d_ratio), kLargestBelowOne);
            z = -rx;
        }
        float hz = sqrtf(1.0f - sqf(z)), sn, cs;
        sincosf(kTwoPi * r1, &sn, &cs);
        Basis B = make_basis(p.n);
        d = to_world(B, v3(cs * hz, sn * hz, z));
        float odotn = dot(d, p.n);
        float u = 2.0f * fabsf(odotn) * d_factors(p, E, odotn, false);
        if (odotn > 0.0f) {
            w = s4(p.d_scale_refl * u);
            pd = kRcpTwoPi * p.d_ratio;
        } else {
            w = s4(p.d_scale_trans * u);
            pd = kRcpTwoPi * (1.0f - p.d_ratio);
        }
    } else if (lobe == 2) {
        float z = x, hz = sqrtf(1.0f - sqf(z)), sn, cs;
        sincosf(kTwoPi * r1, &sn, &cs);
        Basis B = make_basis(p.n);
        d = to_world(B, v3(cs * hz, sn * hz, z));
        float odotn = dot(d, p.n);
        w = scale(p.m_scale, 2.0f * odotn * m_factors(p, E, odotn));
        pd = kRcpTwoPi;
    } else {
        float z = sqrtf(r1), r = sqrtf(1.0f - r1), sn, cs;
        if (z == 0.0f) return false;
        sincosf(kTwoPi * x, &sn, &cs);
        Basis B = make_basis(p.n);
        d = to_world(B, v3(cs * r, sn * r, z));
        float comp = p.diffuse_comp * opaque_complement(E, p.eta_opaque, p.iso_alpha, z);
        w = scale(scale(eon0(p.diffuse_albedo), comp), kPi);
        pd = z * kRcpPi;
    }
    S4 bc = scale(w, pd);
    const float wsel = lobe == 0 ? p.w[0] : lobe == 1 ? p.w[1] : lobe == 2 ? p.w[2] : p.w[3];
    float ptot = pd * wsel;
    if (lobe != 0) { bc = sadd(bc, spec_value(p, E, d)); ptot += p.w[0] * spec_pdf(p, d); }
    if (lobe != 1) { bc = sadd(bc, dmms_value(p, E, d)); ptot += p.w[1] * dmms_pdf(p, d); }
    if (lobe != 2) { bc = sadd(bc, mmms_value(p, E, d)); ptot += p.w[2] * mmms_pdf(p, d); }
    if (lobe != 3) { bc = sadd(bc, diffuse_value(p, E, d)); ptot += p.w[3] * diffuse_pdf(p, d); }
    ptot /= p.total;
    if (!(ptot > 0.0f)) return false;
    wi = d;
    pdf = ptot;
    weight = scale(bc, 1.0f / ptot);
    return true;
}

// ---------------------------------------------------------------- the coat
// The coat's reflection lobe: GGX with the coat's (grey) dielectric Fresnel, f |cos|.
__device__ __forceinline__ float coat_value(const Pbr &p, V3 wi) {
    const V3 wo = p.view;
    const float idotn = dot(p.n, wo), odotn = dot(p.n, wi);
    if (!(idotn * odotn > 0.0f)) return 0.0f;
    const V3 h = normalize(wo + wi);
    const float D = ggx_d(to_local(p.sb, h), p.c_alpha, p.c_alpha);
    const float F = fresnel(kCoatIor, fabsf(dot(wo, h)));
    const float G =
        smith_g1(to_local(p.sb, wo), p.c_alpha, p.c_alpha) * smith_g1(to_local(p.sb, wi), p.c_alpha, p.c_alpha);
    return G * (D * (1.0f / (4.0f * idotn)) * F);
}
__device__ __forceinline__ float coat_pdf(const Pbr &p, V3 wi) {
    const V3 wo = p.view;
    const float idotn = dot(p.n, wo), odotn = dot(p.n, wi);
    if (!(idotn * odotn > 0.0f)) return 0.0f;
    const V3 h = normalize(wo + wi);
    const float idoth = dot(wo, h), odoth = dot(wi, h);
    if (idoth * odoth < 0.0f) return 0.0f;
    const float D = ggx_d(to_local(p.sb, h), p.c_alpha, p.c_alpha);
    return smith_g1(to_local(p.sb, wo), p.c_alpha, p.c_alpha) * (D * (idoth / (4.0f * odoth * idotn)));
}
// The coating lobe: the coat (times its presence) over the base (times in_base_layer_scale; the light
// side's scale is 1 under OpenPBR's default view-side attenuation).
__device__ __forceinline__ S4 coated_eval(const Pbr &p, const EnergyTables &E, V3 wi) {
    const S4 b = base_eval(p, E, wi);
    if (!p.c_on) return b;
    const float c = coat_value(p, wi) * p.coat;
    S4 r;
    for (int k = 0; k < 4; ++k) r.v[k] = c + b.v[k] * p.c_scale.v[k];
    return r;
}
// The coating lobe's pdf: mix(base pdf, coat pdf, coat reflection probability).
__device__ __forceinline__ float coated_pdf(const Pbr &p, V3 wi) {
    const float b = base_pdf(p, wi);
    return p.c_on ? mixf(b, coat_pdf(p, wi), p.c_prob) : b;
}
// The coating lobe's sample (openpbr_coating_lobe.h:258-331): the first random number picks the coat
// or the base and is remapped; the other layer's value and pdf are folded in.
__device__ __forceinline__ bool coated_sample(const Pbr &p, const EnergyTables &E, float r0, float r1, float r2, V3 &wi,
                                              S4 &weight, float &pdf) {
    if (!p.c_on) return base_sample(p, E, r0, r1, r2, wi, weight, pdf);
    const float cp = p.c_prob;
    V3 d;
    S4 comb;
    float ptot;
    if (r0 < cp) {
        const V3 wo = p.view;
        const float idotn = dot(wo, p.n);
        const float x = fminf(r0 * (1.0f / cp), kLargestBelowOne);
        const V3 h = to_world(p.sb, sample_vndf(p.c_alpha, p.c_alpha, to_local(p.sb, wo), x, r1));
        const float idoth = dot(wo, h);
        if (idoth < 0.0f) return false;
        d = normalize(-wo + h * (2.0f * idoth));
        const float odotn = dot(p.n, d);
        if (idotn * odotn <= 0.0f) return false;
        const float D = ggx_d(to_local(p.sb, h), p.c_alpha, p.c_alpha);
        if (D <= 0.0f) return false;
        const float pd = D * smith_g1(to_local(p.sb, wo), p.c_alpha, p.c_alpha) * idoth / (4.0f * dot(d, h) * idotn);
        const float bc = fresnel(kCoatIor, idoth) * smith_g1(to_local(p.sb, d), p.c_alpha, p.c_alpha) * pd * p.coat;
        const S4 be = base_eval(p, E, d);
        for (int k = 0; k < 4; ++k) comb.v[k] = bc + be.v[k] * p.c_scale.v[k];
        ptot = mixf(base_pdf(p, d), pd, cp);
    } else {
        S4 w;
        float pd;
        if (!base_sample(p, E, fminf((r0 - cp) * (1.0f / (1.0f - cp)), kLargestBelowOne), r1, r2, d, w, pd)) return false;
        const float ce = coat_value(p, d) * p.coat;
        for (int k = 0; k < 4; ++k) comb.v[k] = ce + (w.v[k] * pd) * p.c_scale.v[k];
        ptot = mixf(pd, coat_pdf(p, d), cp);
    }
    wi = d;
    pdf = ptot;
    weight = scale(comb, 1.0f / ptot);
    return true;
}

// ---------------------------------------------------------------- the fuzz
// The sheen's pdf (the LTC density in the view-aligned frame) and its coefficients' reflectance R, in
// the normal's own frame; the coefficients are re-read per call (they depend only on the view).
__device__ __forceinline__ float sheen_pdf(const Pbr &p, const EnergyTables &E, V3 wi, float &refl) {
    const Basis B = make_basis(p.n);
    const V3 wo = to_local(B, p.view), wl = to_local(B, wi);
    float coef[3];
    ltc_fetch(E, wo.z, coef);
    refl = coef[2];
    if (wo.z <= 0.0f || wl.z <= 0.0f) return 0.0f;
    return ltc_eval(ltc_rotate(wl, wo.x, -wo.y), coef[0], coef[1]);
}
// openpbr_eval: the sheen over the coating lobe, which is scaled by 1 - view_reflected; f |cos|.
__device__ __forceinline__ S4 pbr_eval(const Pbr &p, const EnergyTables &E, V3 wi) {
    const S4 b = coated_eval(p, E, wi);
    if (!(p.fuzz > 0.0f)) return b;
    float r;
    const float sh = sheen_pdf(p, E, wi, r) * (p.fuzz * r * 1.0f);
    S4 out;
    for (int k = 0; k < 4; ++k) out.v[k] = b.v[k] * ((1.0f - p.f_refl) * 1.0f) + sh;
    return out;
}
// openpbr_pdf: mix(coating pdf, sheen pdf, sheen probability).
__device__ __forceinline__ float pbr_pdf(const Pbr &p, const EnergyTables &E, V3 wi) {
    const float b = coated_pdf(p, wi);
    if (!(p.fuzz > 0.0f)) return b;
    float r;
    return mixf(b, sheen_pdf(p, E, wi, r), p.f_prob);
}
// openpbr_sample through the fuzz lobe (openpbr_fuzz_lobe.h:468-545): the first random number picks the
// sheen or the coating lobe; the sheen samples its LTC with the other two.
__device__ __forceinline__ bool pbr_sample(const Pbr &p, const EnergyTables &E, float r0, float r1, float r2, V3 &wi,
                                           S4 &weight, float &pdf) {
    if (!(p.fuzz > 0.0f)) return coated_sample(p, E, r0, r1, r2, wi, weight, pdf);
    const float sp = p.f_prob, bp = 1.0f - sp;
    V3 d;
    if (r0 < sp) {
        const Basis B = make_basis(p.n);
        const V3 wo = to_local(B, p.view);
        if (wo.z <= 0.0f) return false;
        float coef[3];
        ltc_fetch(E, wo.z, coef);
        float sn, cs;
        sincosf(kTwoPi * r1, &sn, &cs);
        const float z = sqrtf(r2), rr = sqrtf(1.0f - r2), a = 1.0f / coef[0];
        const V3 o = v3(cs * rr, sn * rr, z);
        const V3 wl = ltc_rotate(normalize(v3(o.x * a - o.z * coef[1] * a, o.y * a, o.z)), wo.x, wo.y);
        if (!(wo.z * wl.z > 0.0f)) return false;
        d = normalize(to_world(B, wl));
        float r;
        float pd = sheen_pdf(p, E, d, r) * sp;
        pd += coated_pdf(p, d) * bp;
        weight = scale(pbr_eval(p, E, d), 1.0f / pd);
        pdf = pd;
    } else {
        S4 w;
        float pd;
        if (!coated_sample(p, E, fminf((r0 - sp) * (1.0f / bp), kLargestBelowOne), r1, r2, d, w, pd)) return false;
        float r;
        const float shp = sheen_pdf(p, E, d, r);
        const float sh = shp * (p.fuzz * r * 1.0f);
        S4 bc;
        for (int k = 0; k < 4; ++k) bc.v[k] = (w.v[k] * pd) * ((1.0f - p.f_refl) * 1.0f) + sh;
        pd *= bp;
        pd += shp * sp;
        weight = scale(bc, 1.0f / pd);
        pdf = pd;
    }
    wi = d;
    return true;
}
