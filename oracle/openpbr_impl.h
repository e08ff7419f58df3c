/*
 * openpbr_impl.h -- oracle (test infrastructure only): the subset of the reference's layered BSDF
 * that its GLOSSY / OPENPBR materials use on this path, restated per wavelength lane.  Included by
 * integrator.c (compiled with REAL = float and REAL = double).
 *
 * The reference evaluates the vendored Adobe OpenPBR (Vendor/OpenPBR, pinned c91aad1) through
 * prepare_openpbr_lanes / prepare_bsdf / prepared_eval / prepared_pdf / prepared_sample
 * (main.swift:2485-2645).  Restated here for the inputs those materials set: fuzz = subsurface = thin
 * film = dispersion = anisotropy = 0, base_diffuse_roughness = 0, and a coat of weight `coat` with
 * coat_roughness 0.12 (main.swift:2494-2495) and OpenPBR's other coat defaults (coat_ior 1.6, white
 * coat_color, coat_darkening 1: Vendor/OpenPBR/openpbr_resolved_inputs.h:123-127):
 *   openpbr_prepare_lobes          Vendor/OpenPBR/impl/openpbr_bsdf.h:289-560, 600-860 (the coat's
 *                                  IORs, roughening and darkening at :339-470, 498-552, 831-840)
 *   coating lobe                   impl/openpbr_coating_lobe.h (with the fuzz lobe at weight 0 passing
 *                                  through, impl/openpbr_fuzz_lobe.h:437-545) and its reflection lobe,
 *                                  impl/openpbr_generic_minimal_microfacet_lobe.h with
 *                                  impl/openpbr_ior_reflection_coefficient.h
 *   comprehensive microfacet lobe  impl/openpbr_comprehensive_microfacet_lobe.h (eval, pdf, sample)
 *   reflection/transmission coeff  impl/openpbr_reflection_transmission_coefficient.h
 *   dielectric / metal MMS lobes   impl/openpbr_microfacet_multiple_scattering_lobes.h
 *   energy-conserving diffuse lobe impl/openpbr_diffuse_lobe.h (EON at roughness 0)
 *   aggregate lobe                 impl/openpbr_aggregate_lobe.h (lobe weights, sample, pdf)
 *   GGX / Smith / VNDF / Fresnel / F82 / basis  impl/openpbr_lobe_utils.h, openpbr_basis.h
 *   energy-table lookups           impl/openpbr_microfacet_multiple_scattering_data.h, with the
 *                                  ideal-metal lookups replaced by Shaders/MetalEnergy.metal's
 *                                  129 x 129 table (scripts/prepare_shaders.py:28-39).
 * Colour quantities carry four lanes; lobe-selection and reflection/transmission probabilities use
 * lanes 0-2 (the reference's `a` BSDF), as prepared_sample does (main.swift:2633-2640).
 * The energy tables are baked by energy.c (see oracle.h) and read through vo_energy().
 */
#ifndef VO_OPENPBR_IMPL_H
#define VO_OPENPBR_IMPL_H

/* ------------------------------------------------------------------ helpers */
static inline real sq(real x) { return x * x; }
static inline real max3(s4 a) { return RFMAX(a.v[0], RFMAX(a.v[1], a.v[2])); }
static inline s4 sadd(s4 a, s4 b) { s4 r; for (int k = 0; k < 4; ++k) r.v[k] = a.v[k] + b.v[k]; return r; }
static inline s4 ssub1(s4 a) { s4 r; for (int k = 0; k < 4; ++k) r.v[k] = (real)1 - a.v[k]; return r; }
static inline s4 ssat(s4 a) {
    s4 r;
    for (int k = 0; k < 4; ++k) r.v[k] = RFMIN(RFMAX(a.v[k], (real)0), (real)1);
    return r;
}
static inline real safe_div(real a, real b, real fallback) { return b != 0 ? a / b : fallback; }
static const real LARGEST_BELOW_ONE = (real)0.999999940;

typedef struct { v3 t, b, n; } basis_t;
/* openpbr_make_basis(normal) (openpbr_basis.h): Duff et al.'s branchless frame. */
static basis_t make_basis(v3 n) {
    basis_t B;
    B.n = n;
    real s = n.z < 0 ? (real)-1 : (real)1;
    real a = (real)-1 / (s + n.z);
    real b = n.x * n.y * a;
    B.t = V(b, s + n.y * n.y * a, -n.y);
    B.b = V((real)1 + s * n.x * n.x * a, s * b, -s * n.x);
    return B;
}
/* openpbr_make_basis(normal, tangent, handedness 1): the tangent made perpendicular to the normal. */
static basis_t make_basis_t(v3 n, v3 t) {
    basis_t B;
    B.n = n;
    B.t = norm(sub(t, mulv(n, dot(t, n))));
    B.b = cross(n, B.t);
    return B;
}
static basis_t invert_basis(basis_t B) { B.t = neg(B.t); B.b = neg(B.b); B.n = neg(B.n); return B; }
static v3 to_local(const basis_t *B, v3 d) { return V(dot(d, B->t), dot(d, B->b), dot(d, B->n)); }
static v3 to_world(const basis_t *B, v3 d) { return add(add(mulv(B->t, d.x), mulv(B->b, d.y)), mulv(B->n, d.z)); }

/* openpbr_eval_aniso_ggx, openpbr_eval_aniso_smith_g1 (openpbr_lobe_utils.h:76-137): alphas along the
 * basis' tangent (ax) and bitangent (ay). */
static real ggx_d(v3 h, real ax, real ay) {
    return (real)1 / (PI_R * ax * ay * sq(sq(h.x / ax) + sq(h.y / ay) + sq(h.z)));
}
static real smith_g1(v3 v, real ax, real ay) {
    real z2 = sq(v.z);
    if (z2 == 0) return 0;
    return (real)2 / ((real)1 + RSQRT((real)1 + (sq(ax * v.x) + sq(ay * v.y)) / z2));
}
/* openpbr_sample_aniso_ggx_smith_vndf: spherical-cap VNDF sampling. */
static v3 sample_vndf(real ax, real ay, v3 wi, real r0, real r1) {
    v3 ih = norm(V(ax * wi.x, ay * wi.y, wi.z));
    real c = RFMIN(ih.z, (real)1);
    real ang = (real)6.28318530717958647692 * r0;
    real opc = (real)1 + c;
    real hz = ((real)1 - r1) * opc;
    real omz = r1 * opc;
    real omc = (real)1 - c;
    real opz = hz + omc;
    real tc = RSQRT(opz * omz);
    v3 m = V(tc * RCOS(ang) + ih.x, tc * RSIN(ang) + ih.y, hz);
    return norm(V(ax * m.x, ay * m.y, m.z));
}
static real fresnel(real eta, real ci) {
    if (eta == 1) return 0;
    real st2 = ((real)1 - sq(ci)) / sq(eta);
    if (st2 >= 1) return 1;
    real ct = RSQRT((real)1 - st2);
    real eci = eta * ci, ect = eta * ct;
    real rs = sq((ci - ect) / (ci + ect)), rp = sq((ct - eci) / (ct + eci));
    return (real)0.5 * (rs + rp);
}
static real average_fresnel(real eta) {
    if (eta > 1) return (eta - 1) / ((real)4.08567 + (real)1.00071 * eta);
    return (real)0.997118 + (real)0.1014 * eta - (real)0.965241 * (eta * eta) - (real)0.130607 * (eta * eta) * eta;
}
static real f0_from_ior(real eta) { return sq((eta - 1) / (eta + 1)); }
static real ior_from_f0(real f0) { real s = RSQRT(f0); return (1 + s) / (1 - s); }
static real apply_specular_weight_to_ior(real eta, real w) {
    real f0 = RFMIN(f0_from_ior(eta) * w, (real)0.9999);
    real ext = ior_from_f0(f0);
    return eta < 1 ? 1 / ext : ext;
}
/* F82-tint metal Fresnel ("Fresnel Equations Considered Harmful"), white tint here. */
static s4 metal_b_factor(s4 f0, s4 tint) {
    const real cmax = (real)1 / (real)7, om = 1 - cmax;
    real om5 = om * om * om * om * om, om6 = om5 * om;
    s4 r;
    for (int k = 0; k < 4; ++k) r.v[k] = (f0.v[k] + (1 - f0.v[k]) * om5) * (1 - tint.v[k]) * ((real)1 / (cmax * om6));
    return r;
}
static s4 metal_f82(s4 f0, s4 tint, real c) {
    s4 b = metal_b_factor(f0, tint), r;
    real omc = 1 - c, p5 = omc * omc * omc * omc * omc;
    for (int k = 0; k < 4; ++k) r.v[k] = f0.v[k] + ((1 - f0.v[k]) - b.v[k] * c * omc) * p5;
    return ssat(r);
}
static s4 metal_average_f82(s4 f0, s4 tint) {
    s4 b = metal_b_factor(f0, tint), r;
    for (int k = 0; k < 4; ++k) r.v[k] = f0.v[k] + (1 - f0.v[k]) * ((real)1 / (real)21) - b.v[k] * ((real)1 / (real)126);
    return ssat(r);
}
static int refract_dir(v3 wi, v3 n, real ci, real eta, v3 *wt) {
    real inv = 1 / eta;
    real s2i = RFMAX((real)0, 1 - sq(ci));
    real s2t = sq(inv) * s2i;
    if (s2t >= 1) { *wt = V(0, 0, 0); return 0; }
    real ct = RSQRT(1 - s2t);
    *wt = norm(add(mulv(wi, -inv), mulv(n, inv * ci - ct)));
    return 1;
}

/* ------------------------------------------------------------------ energy tables */
/* OpenPBR's 32-node axes (openpbr_microfacet_multiple_scattering_data.h). */
static real ior_index(real ior) {
    const real inv_range = (real)1 / ((real)2.5 - (real)1);
    if (ior < 1) return (real)15 - ((1 / ior - 1) * inv_range) * (real)15;
    return (real)16 + ((ior - 1) * inv_range) * (real)15;
}
static real clamp_index(real x) { return RFMIN(RFMAX(x, (real)0), (real)31); }
static real lut(int id, int i) { return (real)vo_energy_table(id)[i] * ((real)1 / (real)65535); }
static real lookup2(int id, real x, real y) {
    x = clamp_index(x); y = clamp_index(y);
    int x0 = (int)RFLOOR(x), y0 = (int)RFLOOR(y), x1 = x0 + 1 < 31 ? x0 + 1 : 31, y1 = y0 + 1 < 31 ? y0 + 1 : 31;
    real tx = x - (real)x0, ty = y - (real)y0;
    real v0 = lut(id, x0 * 32 + y0) * (1 - tx) + lut(id, x1 * 32 + y0) * tx;
    real v1 = lut(id, x0 * 32 + y1) * (1 - tx) + lut(id, x1 * 32 + y1) * tx;
    return v0 * (1 - ty) + v1 * ty;
}
static real lookup3(int id, real x, real y, real z) {
    x = clamp_index(x); y = clamp_index(y); z = clamp_index(z);
    int x0 = (int)RFLOOR(x), y0 = (int)RFLOOR(y), z0 = (int)RFLOOR(z);
    int x1 = x0 + 1 < 31 ? x0 + 1 : 31, y1 = y0 + 1 < 31 ? y0 + 1 : 31, z1 = z0 + 1 < 31 ? z0 + 1 : 31;
    real tx = x - (real)x0, ty = y - (real)y0, tz = z - (real)z0;
#define L3(a, b, c) lut(id, (a) * 1024 + (b) * 32 + (c))
    real v00 = L3(x0, y0, z0) * (1 - tx) + L3(x1, y0, z0) * tx;
    real v01 = L3(x0, y0, z1) * (1 - tx) + L3(x1, y0, z1) * tx;
    real v10 = L3(x0, y1, z0) * (1 - tx) + L3(x1, y1, z0) * tx;
    real v11 = L3(x0, y1, z1) * (1 - tx) + L3(x1, y1, z1) * tx;
#undef L3
    real v0 = v00 * (1 - ty) + v10 * ty, v1 = v01 * (1 - ty) + v11 * ty;
    return v0 * (1 - tz) + v1 * tz;
}
static real alpha_index(real alpha) { return RSQRT(alpha) * (real)31; }
static real diel_complement(real ior, real alpha, real c) {
    return lookup3(VO_LUT_DIELECTRIC, ior_index(ior), alpha_index(alpha), c * (real)31);
}
static real diel_avg_complement(real ior, real alpha) { return lookup2(VO_LUT_DIELECTRIC_AVG, ior_index(ior), alpha_index(alpha)); }
static real diel_ratio(real ior, real alpha) { return lookup2(VO_LUT_DIELECTRIC_RATIO, ior_index(ior), alpha_index(alpha)); }
static real extrapolate_ior(real v, real ior) {
    if (ior > (real)2.5 || ior < (real)0.4) {
        real f0max = f0_from_ior((real)2.5);
        return (1 - (f0_from_ior(ior) - f0max) / (1 - f0max)) * v;
    }
    return v;
}
static real opaque_complement(real ior, real alpha, real c) {
    return extrapolate_ior(lookup3(VO_LUT_OPAQUE, ior_index(ior), alpha_index(alpha), c * (real)31), ior);
}
static real opaque_avg_complement(real ior, real alpha) {
    return extrapolate_ior(lookup2(VO_LUT_OPAQUE_AVG, ior_index(ior), alpha_index(alpha)), ior);
}
/* vibe_metal_energy / vibe_metal_average_energy (Shaders/MetalEnergy.metal:2108-2119). */
static real metal_complement(real alpha, real c) {
    const float *t = vo_metal_energy_table();
    real x = RFMIN(RFMAX(RSQRT(alpha) * (real)128, (real)0), (real)128);
    real y = RSQRT(RFMIN(RFMAX(c, (real)0), (real)1)) * (real)128;
    int ix = (int)x < 127 ? (int)x : 127, iy = (int)y < 127 ? (int)y : 127;
    real fy = y - (real)iy, fx = x - (real)ix;
    real a = (real)t[ix * 129 + iy] + ((real)t[ix * 129 + iy + 1] - (real)t[ix * 129 + iy]) * fy;
    real b = (real)t[(ix + 1) * 129 + iy] + ((real)t[(ix + 1) * 129 + iy + 1] - (real)t[(ix + 1) * 129 + iy]) * fy;
    return a + (b - a) * fx;
}
static real metal_avg_complement(real alpha) {
    const float *t = vo_metal_energy_table() + 129 * 129;
    real x = RFMIN(RFMAX(RSQRT(alpha) * (real)128, (real)0), (real)128);
    int ix = (int)x < 127 ? (int)x : 127;
    return (real)t[ix] + ((real)t[ix + 1] - (real)t[ix]) * (x - (real)ix);
}
static real clamp_avg(real v) { return RFMAX(v, (real)1e-12); }
/* mix(x, y, a) = x + (y - x) a, as the shading language defines it. */
static real mixr(real x, real y, real a) { return x + (y - x) * a; }
#define COAT_IOR ((real)1.6)
#define COAT_ROUGHNESS ((real)0.12)
/* fuzz_color white and fuzz_roughness 0.5 (OpenPBR's defaults, openpbr_resolved_inputs.h:131-133) */
#define FUZZ_ROUGHNESS ((real)0.5)

/* ------------------------------------------------------------------ fuzz (Disney sheen LTC) */
/* openpbr_disney_sheen_fetch_coeffs, array mode (openpbr_fuzz_lobe.h:147-176): bilinear in the 32 x 32
 * table over (roughness, cosine). */
static void ltc_fetch(real cos_t, real out[3]) {
    const float *T = vo_ltc_table();
    const real row = RFMAX((real)0, RFMIN(FUZZ_ROUGHNESS, LARGEST_BELOW_ONE)) * (real)31;
    const real col = RFMAX((real)0, RFMIN(cos_t, LARGEST_BELOW_ONE)) * (real)31;
    const real r = RFLOOR(row), c = RFLOOR(col), rf = row - r, cf = col - c;
    const int ri = (int)r, ci = (int)c;
    for (int k = 0; k < 3; ++k) {
        const real a = mixr((real)T[(ri * 32 + ci) * 3 + k], (real)T[(ri * 32 + ci + 1) * 3 + k], cf);
        const real b = mixr((real)T[((ri + 1) * 32 + ci) * 3 + k], (real)T[((ri + 1) * 32 + ci + 1) * 3 + k], cf);
        out[k] = mixr(a, b, rf);
    }
}
/* openpbr_disney_sheen_rotate_vector: about z by the azimuth of (x, y). */
static v3 ltc_rotate(v3 v, real x, real y) {
    const real r2 = x * x + y * y;
    if (r2 == 0) return v;
    const real inv = (real)1 / RSQRT(r2), sn = y * inv, cs = x * inv;
    return V(cs * v.x + sn * -v.y, cs * v.y + sn * v.x, v.z);
}
/* openpbr_disney_sheen_eval_ltc: the LTC's density in the aligned frame. */
static real ltc_eval(v3 wi, const real coef[3]) {
    const real ai = coef[0], bi = coef[1];
    const v3 o = V(ai * wi.x + bi * wi.z, ai * wi.y, wi.z);
    const real z = RFMAX((real)0, wi.z), l2 = dot(o, o);
    if (l2 == 0 || ai == 0 || z == 0) return 0;
    const real k = ai / l2;
    return (real)0.318309886 * (z * k) * k;
}

/* ------------------------------------------------------------------ prepared BSDF */
typedef struct {
    v3 view;                 /* wo */
    basis_t spec_basis;      /* specular_anisotropy_basis_ff */
    v3 n;                    /* normal_ff */
    real ax, ay, iso_alpha;  /* the specular lobe's alphas (clamped) and specular_alpha */
    /* reflection / transmission coefficient */
    real eta_trans, eta_opaque, eta_refr;
    real scale_trans_refl, scale_opaque_refl, transmission, metal_amount;
    s4 f0_metal;
    /* dielectric MMS */
    real d_ratio, d_ec_idotn, d_scale_refl, d_scale_trans;
    int d_on;
    /* metal MMS */
    s4 m_scale;
    real m_ec_idotn;
    int m_on;
    /* diffuse */
    s4 diffuse_albedo;
    real diffuse_comp;       /* cached_specular_energy_compensation */
    real weights[4];         /* specular, dielectric MMS, metal MMS, diffuse */
    real total;
    /* coating lobe, used when a coated surface is entered from outside */
    int c_on;
    real coat, c_alpha, c_prob;  /* presence, coat alpha (clamped), coat reflection probability */
    s4 c_scale;              /* in_base_layer_scale: (1 - in_reflected) x darkening, per lane */
    real c_refl;             /* in_reflected */
    /* fuzz lobe (openpbr_fuzz_lobe.h), used when a fuzzy surface is hit from outside */
    int f_on;
    real fuzz, f_refl, f_prob;   /* presence, view_reflected, sheen probability */
    basis_t f_basis;
    v3 f_wo;                 /* the view direction in the fuzz basis */
    real f_coef[3];          /* LTC coefficients (a_inv, b_inv, R) at the view's cosine */
} pbr_t;

typedef struct {
    real type, roughness, ior, metalness, transmission, coat, anisotropy;
    v3 tangent;              /* the hit's tangent (dP/du); used by an anisotropic specular lobe */
    real fuzz;
} pbr_inputs;

/* all_coefficients at cosine c: reflection and transmission per lane. */
static void coefficients(const pbr_t *p, real c, s4 *refl, s4 *trans) {
    real ft = fresnel(p->eta_trans, c), fo = fresnel(p->eta_opaque, c);
    s4 fm = metal_f82(p->f0_metal, S4(1), c);
    for (int k = 0; k < 4; ++k) {
        refl->v[k] = p->scale_trans_refl * ft + p->scale_opaque_refl * fo + p->metal_amount * fm.v[k];
        trans->v[k] = p->transmission * (1 - ft);
    }
}
static void probabilities(const pbr_t *p, real c, s4 *refl, s4 *trans, real *pr, real *pt) {
    coefficients(p, c, refl, trans);
    real rm = max3(*refl), tm = max3(*trans);
    *pr = safe_div(rm, rm + tm, 0);
    *pt = safe_div(tm, rm + tm, 0);
}

/* prepare_openpbr_lanes for the subset (main.swift:2522-2560 -> openpbr_prepare_lobes).
 * n: the hit's shading normal (facing the arriving ray), wo: toward the viewer, inside: hit from
 * the back of a transmissive surface, rho: base colour per lane. */
static void pbr_prepare(pbr_t *p, const pbr_inputs *in, v3 n, v3 wo, int inside_flag, s4 rho) {
    const int glossy = in->type == (real)1;
    real rough = RFMAX((real)0.03, in->roughness);
    real metalness = glossy ? (real)1 : in->metalness;
    const real spec_w = 1;  /* specular_weight of non-MaterialX materials (main.swift:2525) */
    real ior = RFMAX((real)1.01, in->ior);
    real transw = in->transmission;
    v3 outward = transw > 0 && inside_flag ? neg(n) : n;
    /* geometry_basis (main.swift:2508-2509): the tangent frame where the tangent is usable.  Only an
     * anisotropic lobe depends on the frame's azimuth; isotropic ones keep the normal's own frame. */
    const real aniso = in->anisotropy;
    basis_t gb = aniso > 0 && dot(in->tangent, in->tangent) > (real)0.5 && RFABS(dot(in->tangent, outward)) < (real)0.99
                     ? make_basis_t(outward, in->tangent) : make_basis(outward);
    p->view = wo;
    real cos_theta = dot(wo, gb.n);
    int back = cos_theta < 0;
    int inside = back;  /* not thin-walled */
    p->spec_basis = back ? invert_basis(gb) : gb;
    p->n = p->spec_basis.n;
    /* IORs (openpbr_bsdf.h:339-418): the exterior is IOR 1; entering a coated surface, the base sits
     * under the coat, whose IOR is blended by the coat weight. */
    const real coat = in->coat;
    const int enter_coat = coat > 0 && !inside;
    const real wci = enter_coat ? mixr(1, COAT_IOR, coat) : (real)1;
    real wsi = wci * apply_specular_weight_to_ior(ior / wci, spec_w);
    real r_refr, r_trans, r_opaque, rel_coat = 1;
    if (inside) {
        r_refr = (real)1 / wsi;
        r_trans = r_refr;
        r_opaque = wsi / (real)1;
    } else {
        r_refr = wsi / (real)1;
        if (enter_coat) {
            /* no flip from external to internal reflection under the coat */
            r_trans = wci > wsi && wsi >= 1 ? wci / wsi : wsi / wci;
            rel_coat = COAT_IOR / (real)1;
        } else {
            r_trans = r_refr;
        }
        r_opaque = r_trans;
    }
    if (r_refr >= 1) r_refr = RFMAX(r_refr, (real)1.0001);
    else r_refr = RFMIN(r_refr, (real)1 / (real)1.0001);
    /* Roughnesses (openpbr_bsdf.h:424-460): the coat's adjusted for the fuzz over it, the base's for
     * the coat. */
    const real fuzz = in->fuzz;
    const real fuzz_adjust = ((real)3 * ((real)1 / (real)3)) * FUZZ_ROUGHNESS * (real)0.005;
    const real c_rough = mixr(COAT_ROUGHNESS,
                              RSQRT(RSQRT(RFMIN((real)1, sq(sq(COAT_ROUGHNESS)) + fuzz_adjust * sq(sq(FUZZ_ROUGHNESS))))),
                              fuzz);
    const real adjust = mixr(1, 0, rel_coat >= 1 ? 1 / rel_coat : rel_coat);
    const real adj_rough =
        mixr(rough, RSQRT(RSQRT(RFMIN((real)1, sq(sq(rough)) + adjust * sq(sq(c_rough))))), coat);
    real salpha = adj_rough * adj_rough;
    p->iso_alpha = salpha;
    /* openpbr_compute_anisotropic_alpha (openpbr_lobe_utils.h:41-72), floored at 1e-6 */
    p->ax = p->ay = salpha;
    if (aniso > 0) {
        p->ax = salpha * RSQRT((real)2 / ((real)1 + sq((real)1 - aniso)));
        p->ay = ((real)1 - aniso) * p->ax;
    }
    p->ax = RFMAX(p->ax, (real)1e-6);
    p->ay = RFMAX(p->ay, (real)1e-6);
    if (vo_get_fault() == 14) { real t = p->ax; p->ax = p->ay; p->ay = t; }  /* planted */
    real non_trans = 1 - transw, dielectric = 1 - metalness;
    real opaque = non_trans * dielectric;
    real trans = dielectric * transw;
    real dmetal = metalness * spec_w;
    s4 avg_f = metal_average_f82(rho, S4(1));
    real scale_trans_refl = trans;           /* specular_color (1) x trans; inside: trans */
    real scale_opaque_refl = opaque;         /* specular_color (1) x opaque_dielectric */
    p->eta_trans = r_trans;
    p->eta_opaque = r_opaque;
    p->eta_refr = r_refr;
    p->scale_trans_refl = scale_trans_refl;
    p->scale_opaque_refl = scale_opaque_refl;
    p->transmission = trans;                 /* transmission_color (1) x trans */
    p->metal_amount = dmetal;
    p->f0_metal = rho;                       /* base_color x base_weight (1) */
    /* dielectric MMS lobe */
    p->d_scale_refl = scale_trans_refl;
    p->d_scale_trans = trans;
    p->d_on = salpha >= (real)0.04 * (real)0.04;
    real idotn = dot(wo, p->n);
    if (p->d_on) {
        p->d_ratio = diel_ratio(r_refr, salpha);
        p->d_ec_idotn = diel_complement(r_refr, salpha, idotn);
    } else {
        p->d_ratio = 0;
        p->d_ec_idotn = 0;
    }
    /* metal MMS lobe */
    for (int k = 0; k < 4; ++k) p->m_scale.v[k] = sq(avg_f.v[k]) * dmetal;
    p->m_on = salpha >= (real)0.04 * (real)0.04;
    p->m_ec_idotn = p->m_on ? metal_complement(salpha, idotn) : 0;
    /* diffuse lobe */
    p->diffuse_albedo = sscale(rho, opaque);
    p->diffuse_comp = opaque_complement(r_opaque, salpha, idotn) / clamp_avg(opaque_avg_complement(r_opaque, salpha));
    /* aggregate lobe weights (openpbr_estimate_lobe_contribution of each lobe) */
    s4 rc, tc;
    coefficients(p, idotn, &rc, &tc);
    p->weights[0] = max3(sadd(rc, tc));
    p->weights[1] = p->d_on ? max3(S4(p->d_scale_refl * p->d_ratio + p->d_scale_trans * (1 - p->d_ratio))) * p->d_ec_idotn : 0;
    p->weights[2] = p->m_on ? max3(p->m_scale) * p->m_ec_idotn : 0;
    p->weights[3] = max3(sscale(p->diffuse_albedo, p->diffuse_comp));
    p->total = p->weights[0] + p->weights[1] + p->weights[2] + p->weights[3];
    /* The coating lobe (openpbr_bsdf.h:498-552, 831-840; openpbr_coating_lobe.h:167-185, 137-155).
     * Hit from inside, the coat only tints (white here: no change); at weight 0 it passes through. */
    p->c_on = enter_coat;
    p->coat = coat;
    p->c_alpha = RFMAX(sq(c_rough), (real)1e-6);
    p->c_prob = 0;
    p->c_scale = S4(1);
    if (enter_coat) {
        /* darkening by internal reflections under the coat, OpenPBR v1.1.1's interfaced-Lambertian
         * Delta = (1 - K) / (1 - E_b K), lerped in by the coat weight */
        const real ks = average_fresnel(rel_coat);
        const real kr = 1 - (1 - average_fresnel(rel_coat)) / sq(rel_coat);
        const real so = average_fresnel(r_opaque), st = average_fresnel(r_trans);
        const real spec_base = (opaque * so + ((real)0 + trans) * st) + metalness;
        const real kk = mixr(ks, kr, mixr(1, adj_rough, spec_base));
        real refl = 0;
        if (idotn > 0) {
            const real c_iso = sq(c_rough);
            refl = (c_iso == 0 ? fresnel(rel_coat, idotn) : 1 - opaque_complement(rel_coat, c_iso, idotn)) * coat;
        }
        for (int k = 0; k < 4; ++k) {
            const real eb = ((dmetal * avg_f.v[k] + opaque * mixr(rho.v[k], 1, so)) + (real)0) + trans;
            const real delta = (1 - kk) / (1 - eb * kk);
            const real dark = vo_get_fault() == 13 ? (real)1 : mixr(1, delta, coat * (real)1);
            p->c_scale.v[k] = (1 - refl) * dark;
        }
        p->c_prob = safe_div(refl, refl + p->total * max3(p->c_scale), (real)0.5);
        p->c_refl = refl;
    } else {
        p->c_refl = 0;
    }
    /* The fuzz lobe (openpbr_bsdf.h:842-858, openpbr_fuzz_lobe.h:395-430): presence `fuzz` outside,
     * 0 inside; its normal blends the base and coat normals, which coincide here. */
    p->f_on = fuzz > 0 && !inside;
    p->fuzz = p->f_on ? fuzz : 0;
    p->f_refl = 0;
    p->f_prob = 0;
    if (p->f_on) {
        p->f_basis = make_basis(p->n);
        p->f_wo = to_local(&p->f_basis, wo);
        ltc_fetch(p->f_wo.z, p->f_coef);
        p->f_refl = p->f_wo.z <= 0 ? 0 : p->fuzz * p->f_coef[2];
        /* the coating lobe's estimate: its reflectance plus the base's weight through it */
        const real below = p->c_on ? p->c_refl + p->total * max3(p->c_scale) : p->total;
        p->f_prob = safe_div(p->f_refl, p->f_refl + below * (1 - p->f_refl), (real)0.5);
    }
}

/* ---- lobe values (f |cos|) and pdfs */
static s4 spec_value(const pbr_t *p, v3 wi) {
    v3 wo = p->view;
    real idotn = dot(p->n, wo), odotn = dot(p->n, wi);
    if (idotn * odotn > 0) {
        v3 h = norm(add(wo, wi));
        real idoth = dot(wo, h);
        real D = ggx_d(to_local(&p->spec_basis, h), p->ax, p->ay);
        real other = (real)1 / ((real)4 * idotn);
        s4 F, T;
        coefficients(p, RFABS(idoth), &F, &T);
        real G = smith_g1(to_local(&p->spec_basis, wo), p->ax, p->ay) *
                 (vo_get_fault() == 5 ? (real)1 : smith_g1(to_local(&p->spec_basis, wi), p->ax, p->ay));
        return sscale(F, G * D * other);
    }
    if (idotn * odotn < 0) {
        v3 h = neg(norm(add(wo, mulv(wi, p->eta_refr))));
        int valid = (p->eta_refr >= 1 && dot(p->n, h) > 0) || (p->eta_refr < 1 && dot(p->n, h) < 0);
        if (!valid) return S4(0);
        real idoth = dot(wo, h), odoth = dot(wi, h);
        if (idoth * odoth >= 0) return S4(0);
        real D = ggx_d(to_local(&p->spec_basis, h), p->ax, p->ay);
        real other = sq(p->eta_refr) * RFABS(idoth) * RFABS(odoth) / (sq(idoth + p->eta_refr * odoth) * idotn);
        s4 F, T;
        coefficients(p, RFABS(idoth), &F, &T);
        real G = smith_g1(to_local(&p->spec_basis, wo), p->ax, p->ay) * smith_g1(to_local(&p->spec_basis, wi), p->ax, p->ay);
        return sscale(T, G * D * other);
    }
    return S4(0);
}
static real spec_pdf(const pbr_t *p, v3 wi) {
    v3 wo = p->view;
    real idotn = dot(p->n, wo), odotn = dot(p->n, wi);
    s4 F, T;
    real pr, pt;
    if (idotn * odotn > 0) {
        v3 h = norm(add(wo, wi));
        real idoth = dot(wo, h), odoth = dot(wi, h);
        if (idoth * odoth < 0) return 0;
        real D = ggx_d(to_local(&p->spec_basis, h), p->ax, p->ay);
        real other = idoth / ((real)4 * odoth * idotn);
        probabilities(p, RFABS(idoth), &F, &T, &pr, &pt);
        return smith_g1(to_local(&p->spec_basis, wo), p->ax, p->ay) * (pr * D * other);
    }
    if (idotn * odotn < 0) {
        v3 h = neg(norm(add(wo, mulv(wi, p->eta_refr))));
        int valid = (p->eta_refr >= 1 && dot(p->n, h) > 0) || (p->eta_refr < 1 && dot(p->n, h) < 0);
        if (!valid) return 0;
        real idoth = dot(wo, h), odoth = dot(wi, h);
        if (idoth * odoth >= 0) return 0;
        real D = ggx_d(to_local(&p->spec_basis, h), p->ax, p->ay);
        real other = sq(p->eta_refr) * RFABS(idoth) * RFABS(odoth) / (sq(idoth + p->eta_refr * odoth) * idotn);
        probabilities(p, RFABS(idoth), &F, &T, &pr, &pt);
        return smith_g1(to_local(&p->spec_basis, wo), p->ax, p->ay) * (pt * D * other);
    }
    return 0;
}
static real d_factors(const pbr_t *p, real odotn, int with_ratio) {
    real ior = p->eta_refr, ratio = p->d_ratio;
    if (!(odotn > 0)) {
        odotn = -odotn;
        ior = 1 / p->eta_refr;
        ratio = 1 - p->d_ratio;
    }
    real t = p->d_ec_idotn * diel_complement(ior, p->iso_alpha, odotn) / clamp_avg(diel_avg_complement(ior, p->iso_alpha));
    if (with_ratio) t *= ratio;
    return RFMIN(t, safe_div(1, odotn, 0));
}
static s4 dmms_value(const pbr_t *p, v3 wi) {
    if (!p->d_on) return S4(0);
    real odotn = dot(wi, p->n);
    real u = d_factors(p, odotn, 1) * (real)0.318309886 * RFABS(odotn);
    return S4((odotn > 0 ? p->d_scale_refl : p->d_scale_trans) * u);
}
static real dmms_pdf(const pbr_t *p, v3 wi) {
    if (!p->d_on) return 0;
    return dot(wi, p->n) > 0 ? (real)0.159154943 * p->d_ratio : (real)0.159154943 * (1 - p->d_ratio);
}
static real m_factors(const pbr_t *p, real odotn) {
    real t = p->m_ec_idotn * metal_complement(p->iso_alpha, odotn) / clamp_avg(metal_avg_complement(p->iso_alpha));
    return RFMIN(t, safe_div(1, odotn, 0));
}
static s4 mmms_value(const pbr_t *p, v3 wi) {
    if (!p->m_on) return S4(0);
    real odotn = dot(wi, p->n);
    if (!(odotn > 0)) return S4(0);
    return sscale(p->m_scale, m_factors(p, odotn) * (real)0.318309886 * odotn);
}
static real mmms_pdf(const pbr_t *p, v3 wi) {
    if (!p->m_on) return 0;
    return dot(wi, p->n) > 0 ? (real)0.159154943 : 0;
}
/* openpbr_f_EON at roughness 0 with the approximate FON albedo (exact = false). */
static s4 eon0(s4 rho) {
    s4 r;
    for (int k = 0; k < 4; ++k) {
        real fss = rho.v[k] * (real)0.318309886;
        real rms = rho.v[k] * rho.v[k];
        real fms = (rms * (real)0.318309886) * RFMAX((real)1e-7, 0) * RFMAX((real)1e-7, 0) / RFMAX((real)1e-7, 0);
        r.v[k] = fss + fms;
    }
    return r;
}
static s4 diffuse_value(const pbr_t *p, v3 wi) {
    real co = dot(wi, p->n);
    if (co <= 0) return S4(0);
    real comp = p->diffuse_comp * opaque_complement(p->eta_opaque, p->iso_alpha, co);
    return sscale(sscale(eon0(p->diffuse_albedo), comp), co);
}
static real diffuse_pdf(const pbr_t *p, v3 wi) { return RFMAX(dot(wi, p->n), (real)0) * (real)0.318309886; }

/* The base layer (the aggregate lobe): f |cos| per lane. */
static s4 base_eval(const pbr_t *p, v3 wi) {
    return sadd(sadd(sadd(spec_value(p, wi), dmms_value(p, wi)), mmms_value(p, wi)), diffuse_value(p, wi));
}
/* The aggregate lobe's pdf: weight-averaged lobe pdfs. */
static real base_pdf(const pbr_t *p, v3 wi) {
    real s = p->weights[0] * spec_pdf(p, wi) + p->weights[1] * dmms_pdf(p, wi) + p->weights[2] * mmms_pdf(p, wi) +
             p->weights[3] * diffuse_pdf(p, wi);
    return p->total > 0 ? s / p->total : 0;
}
/* openpbr_sample (aggregate lobe, openpbr_aggregate_lobe.h): choose a lobe by weight, sample it
 * with its own weight and pdf, then fold in every other lobe's value and pdf (no lobe of the subset
 * is a delta).  Returns 0 when the sample is invalid. */
static int base_sample(const pbr_t *p, real r0, real r1, real r2, v3 *wi, s4 *weight, real *pdf) {
    if (p->total <= (real)1.175494351e-38) return 0;
    real x = r0 * p->total, acc = 0;
    int lobe = -1;
    for (int k = 0; k < 4; ++k) {
        if (p->weights[k] <= 0) continue;
        if (x < acc + p->weights[k]) { lobe = k; x = RFMIN((x - acc) / p->weights[k], LARGEST_BELOW_ONE); break; }
        acc += p->weights[k];
    }
    if (lobe < 0) {
        for (int k = 3; k >= 0; --k)
            if (p->weights[k] > 0) { lobe = k; x = LARGEST_BELOW_ONE; break; }
    }
    v3 wo = p->view;
    v3 d;
    s4 w;
    real pd;
    if (lobe == 0) {
        /* openpbr_sample_lobe, comprehensive microfacet lobe */
        real idotn = dot(wo, p->n);
        v3 h = to_world(&p->spec_basis, sample_vndf(p->ax, p->ay, to_local(&p->spec_basis, wo), x, r1));
        real idoth = dot(wo, h);
        if (idoth < 0) return 0;
        s4 F, T;
        real pr, pt;
        probabilities(p, idoth, &F, &T, &pr, &pt);
        if (pr + pt == 0) return 0;
        real D, gm;
        if (r2 < pr) {
            d = norm(add(neg(wo), mulv(h, (real)2 * idoth)));
            real odotn = dot(p->n, d);
            if (idotn * odotn <= 0) return 0;
            real odoth = dot(d, h);
            real gs = smith_g1(to_local(&p->spec_basis, d), p->ax, p->ay);
            w = sscale(sscale(F, (real)1 / pr), gs);
            D = ggx_d(to_local(&p->spec_basis, h), p->ax, p->ay);
            if (D <= 0) return 0;
            gm = smith_g1(to_local(&p->spec_basis, wo), p->ax, p->ay);
            pd = pr * D * gm * idoth / ((real)4 * odoth * idotn);
        } else {
            if (!refract_dir(wo, h, idoth, p->eta_refr, &d)) return 0;
            real odotn = dot(d, p->n);
            if (idotn * odotn >= 0) return 0;
            real odoth = dot(d, h);
            real gs = smith_g1(to_local(&p->spec_basis, d), p->ax, p->ay);
            w = sscale(sscale(T, (real)1 / pt), gs);
            D = ggx_d(to_local(&p->spec_basis, h), p->ax, p->ay);
            if (D <= 0) return 0;
            gm = smith_g1(to_local(&p->spec_basis, wo), p->ax, p->ay);
            pd = pt * D * gm * sq(p->eta_refr) * RFABS(idoth) * RFABS(odoth) / (sq(idoth + p->eta_refr * odoth) * idotn);
        }
    } else if (lobe == 1) {
        real rx = x, z;
        if (rx < p->d_ratio) {
            rx = RFMIN(rx / p->d_ratio, LARGEST_BELOW_ONE);
            z = rx;
        } else {
            rx = RFMIN((rx - p->d_ratio) / (1 - p->d_ratio), LARGEST_BELOW_ONE);
            z = -rx;
        }
        real hz = RSQRT(1 - sq(z)), az = (real)6.28318530717958647692 * r1;
        basis_t B = make_basis(p->n);
        d = to_world(&B, V(RCOS(az) * hz, RSIN(az) * hz, z));
        real odotn = dot(d, p->n);
        real u = (real)2 * RFABS(odotn) * d_factors(p, odotn, 0);
        if (odotn > 0) {
            w = S4(p->d_scale_refl * u);
            pd = (real)0.159154943 * p->d_ratio;
        } else {
            w = S4(p->d_scale_trans * u);
            pd = (real)0.159154943 * (1 - p->d_ratio);
        }
    } else if (lobe == 2) {
        real z = x, hz = RSQRT(1 - sq(z)), az = (real)6.28318530717958647692 * r1;
        basis_t B = make_basis(p->n);
        d = to_world(&B, V(RCOS(az) * hz, RSIN(az) * hz, z));
        real odotn = dot(d, p->n);
        w = sscale(p->m_scale, (real)2 * odotn * m_factors(p, odotn));
        pd = (real)0.159154943;
    } else {
        real phi = (real)6.28318530717958647692 * x;
        real z = RSQRT(r1), r = RSQRT(1 - r1);
        if (z == 0) return 0;
        basis_t B = make_basis(p->n);
        d = to_world(&B, V(RCOS(phi) * r, RSIN(phi) * r, z));
        real comp = p->diffuse_comp * opaque_complement(p->eta_opaque, p->iso_alpha, z);
        w = sscale(sscale(eon0(p->diffuse_albedo), comp), PI_R);
        pd = z * (real)0.318309886;
    }
    /* Fold in the other lobes (all non-delta). */
    s4 bc = sscale(w, pd);
    real ptot = pd * p->weights[lobe];
    if (lobe != 0) { bc = sadd(bc, spec_value(p, d)); ptot += p->weights[0] * spec_pdf(p, d); }
    if (lobe != 1) { bc = sadd(bc, dmms_value(p, d)); ptot += p->weights[1] * dmms_pdf(p, d); }
    if (lobe != 2) { bc = sadd(bc, mmms_value(p, d)); ptot += p->weights[2] * mmms_pdf(p, d); }
    if (lobe != 3) { bc = sadd(bc, diffuse_value(p, d)); ptot += p->weights[3] * diffuse_pdf(p, d); }
    ptot /= p->total;
    if (!(ptot > 0)) return 0;
    *wi = d;
    *pdf = ptot;
    *weight = sscale(bc, (real)1 / ptot);
    return 1;
}

/* ---- the coat's reflection lobe: GGX with the coat's dielectric Fresnel (grey) */
static real coat_value(const pbr_t *p, v3 wi) {
    v3 wo = p->view;
    real idotn = dot(p->n, wo), odotn = dot(p->n, wi);
    if (!(idotn * odotn > 0)) return 0;
    v3 h = norm(add(wo, wi));
    real idoth = dot(wo, h);
    real D = ggx_d(to_local(&p->spec_basis, h), p->c_alpha, p->c_alpha);
    real other = (real)1 / ((real)4 * idotn);
    real F = fresnel(COAT_IOR, RFABS(idoth));
    real G = smith_g1(to_local(&p->spec_basis, wo), p->c_alpha, p->c_alpha) *
             smith_g1(to_local(&p->spec_basis, wi), p->c_alpha, p->c_alpha);
    return G * (D * other * F);
}
static real coat_pdf(const pbr_t *p, v3 wi) {
    v3 wo = p->view;
    real idotn = dot(p->n, wo), odotn = dot(p->n, wi);
    if (!(idotn * odotn > 0)) return 0;
    v3 h = norm(add(wo, wi));
    real idoth = dot(wo, h), odoth = dot(wi, h);
    if (idoth * odoth < 0) return 0;
    real D = ggx_d(to_local(&p->spec_basis, h), p->c_alpha, p->c_alpha);
    real other = idoth / ((real)4 * odoth * idotn);
    return smith_g1(to_local(&p->spec_basis, wo), p->c_alpha, p->c_alpha) * (D * other);
}
static int coat_sample(const pbr_t *p, real r0, real r1, v3 *wi, real *weight, real *pdf) {
    v3 wo = p->view;
    real idotn = dot(wo, p->n);
    v3 h = to_world(&p->spec_basis, sample_vndf(p->c_alpha, p->c_alpha, to_local(&p->spec_basis, wo), r0, r1));
    real idoth = dot(wo, h);
    if (idoth < 0) return 0;
    v3 d = norm(add(neg(wo), mulv(h, (real)2 * idoth)));
    real odotn = dot(p->n, d);
    if (idotn * odotn <= 0) return 0;
    real odoth = dot(d, h);
    real w = fresnel(COAT_IOR, idoth) * smith_g1(to_local(&p->spec_basis, d), p->c_alpha, p->c_alpha);
    real D = ggx_d(to_local(&p->spec_basis, h), p->c_alpha, p->c_alpha);
    if (D <= 0) return 0;
    *pdf = D * smith_g1(to_local(&p->spec_basis, wo), p->c_alpha, p->c_alpha) * idoth / ((real)4 * odoth * idotn);
    *weight = w;
    *wi = d;
    return 1;
}

/* The coating lobe: the coat (scaled by its presence) over the base (scaled by in_base_layer_scale;
 * the light side's scale is 1 with OpenPBR's default view-side attenuation), f |cos| per lane. */
static s4 coated_eval(const pbr_t *p, v3 wi) {
    s4 b = base_eval(p, wi);
    if (!p->c_on) return b;
    real c = coat_value(p, wi);
    s4 r;
    for (int k = 0; k < 4; ++k) r.v[k] = c * p->coat + b.v[k] * p->c_scale.v[k];
    return r;
}
/* The coating lobe's pdf: mix(base pdf, coat pdf, coat reflection probability). */
static real coated_pdf(const pbr_t *p, v3 wi) {
    if (!p->c_on) return base_pdf(p, wi);
    return mixr(base_pdf(p, wi), coat_pdf(p, wi), p->c_prob);
}
/* The coating lobe's sample (openpbr_coating_lobe.h:258-331): the first random number chooses the coat
 * or the base and is remapped; the other layer's value and pdf are folded in. */
static int coated_sample(const pbr_t *p, real r0, real r1, real r2, v3 *wi, s4 *weight, real *pdf) {
    if (!p->c_on) return base_sample(p, r0, r1, r2, wi, weight, pdf);
    const real cp = p->c_prob;
    v3 d;
    real ptot;
    s4 comb;
    if (r0 < cp) {
        real x = RFMIN(r0 * ((real)1 / cp), LARGEST_BELOW_ONE), w, pd;
        if (!coat_sample(p, x, r1, &d, &w, &pd)) return 0;
        real bc = w * pd;
        s4 be = base_eval(p, d);
        for (int k = 0; k < 4; ++k) comb.v[k] = bc * p->coat + be.v[k] * p->c_scale.v[k];
        ptot = mixr(base_pdf(p, d), pd, cp);
    } else {
        real x = RFMIN((r0 - cp) * ((real)1 / (1 - cp)), LARGEST_BELOW_ONE), pd;
        s4 w;
        if (!base_sample(p, x, r1, r2, &d, &w, &pd)) return 0;
        real ce = coat_value(p, d);
        for (int k = 0; k < 4; ++k) comb.v[k] = ce * p->coat + (w.v[k] * pd) * p->c_scale.v[k];
        ptot = mixr(pd, coat_pdf(p, d), cp);
    }
    *wi = d;
    *pdf = ptot;
    *weight = sscale(comb, (real)1 / ptot);
    return 1;
}

/* ---- the fuzz lobe (openpbr_fuzz_lobe.h:179-306, 437-545) */
static real sheen_pdf(const pbr_t *p, v3 wi) {
    const v3 wl = to_local(&p->f_basis, wi);
    if (p->f_wo.z <= 0 || wl.z <= 0) return 0;
    return ltc_eval(ltc_rotate(wl, p->f_wo.x, vo_get_fault() == 15 ? p->f_wo.y : -p->f_wo.y), p->f_coef);  /* 15: planted */
}
static real sheen_value(const pbr_t *p, v3 wi) { return sheen_pdf(p, wi) * (p->fuzz * p->f_coef[2] * (real)1); }
static int sheen_sample(const pbr_t *p, real u0, real u1, v3 *wi) {
    if (p->f_wo.z <= 0) return 0;
    const real phi = (real)6.28318530717958647692 * u0, z = RSQRT(u1), r = RSQRT(1 - u1);
    const real a = 1 / p->f_coef[0], bi = p->f_coef[1];
    const v3 o = V(RCOS(phi) * r, RSIN(phi) * r, z);
    const v3 std_dir = norm(V(o.x * a - o.z * bi * a, o.y * a, o.z));
    const v3 wl = ltc_rotate(std_dir, p->f_wo.x, p->f_wo.y);
    if (!(p->f_wo.z * wl.z > 0)) return 0;
    *wi = norm(to_world(&p->f_basis, wl));
    return 1;
}

/* openpbr_eval: the sheen over the coating lobe, which is scaled by 1 - view_reflected; f |cos|. */
static s4 pbr_eval(const pbr_t *p, v3 wi) {
    s4 b = coated_eval(p, wi);
    if (!p->f_on) return b;
    const real sh = sheen_value(p, wi);
    s4 r;
    for (int k = 0; k < 4; ++k) r.v[k] = b.v[k] * ((1 - p->f_refl) * (real)1) + sh;
    return r;
}
/* openpbr_pdf: mix(coating pdf, sheen pdf, sheen probability). */
static real pbr_pdf(const pbr_t *p, v3 wi) {
    if (!p->f_on) return coated_pdf(p, wi);
    return mixr(coated_pdf(p, wi), sheen_pdf(p, wi), p->f_prob);
}
/* openpbr_sample through the fuzz lobe: the first random number chooses the sheen or the coating lobe;
 * the sheen samples its LTC with the other two. */
static int pbr_sample(const pbr_t *p, real r0, real r1, real r2, v3 *wi, s4 *weight, real *pdf) {
    if (!p->f_on) return coated_sample(p, r0, r1, r2, wi, weight, pdf);
    const real sp = p->f_prob, bp = 1 - sp;
    v3 d;
    if (r0 < sp) {
        if (!sheen_sample(p, r1, r2, &d)) return 0;
        real pd = sheen_pdf(p, d) * sp;
        pd += coated_pdf(p, d) * bp;
        *weight = sscale(pbr_eval(p, d), (real)1 / pd);
        *pdf = pd;
    } else {
        real pd;
        s4 w;
        if (!coated_sample(p, RFMIN((r0 - sp) * ((real)1 / bp), LARGEST_BELOW_ONE), r1, r2, &d, &w, &pd)) return 0;
        const real sh = sheen_value(p, d);
        s4 bc;
        for (int k = 0; k < 4; ++k) bc.v[k] = (w.v[k] * pd) * ((1 - p->f_refl) * (real)1) + sh;
        pd *= bp;
        pd += sheen_pdf(p, d) * sp;
        *weight = sscale(bc, (real)1 / pd);
        *pdf = pd;
    }
    *wi = d;
    return 1;
}

#endif
