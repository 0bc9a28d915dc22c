// The seven BSDFs of the reference, with Fresnel and GGX helpers, as
// host/device functions over a flat "local BSDF" record (the material
// evaluated at one hit point).
//
// References: render/material/fresnel.h:7-85, render/material/ggx.h:6-64
// (GGX_Sample_Visible_Area defined), render/material/bsdf/*.h,
// render/material/optix_material.h:70-111 (Sample/Eval/GetAlbedo dispatch).
#pragma once

#include "pt_math.h"

namespace pupil {

// EMatType order of decl/material_decl.inl:3-9 (Unknown = 0).
enum MatType : uint32_t {
    kMatUnknown = 0,
    kMatDiffuse = 1,
    kMatDielectric = 2,
    kMatRoughDielectric = 3,
    kMatConductor = 4,
    kMatRoughConductor = 5,
    kMatPlastic = 6,
    kMatRoughPlastic = 7,
};

// EBsdfLobeType bits of render/material/bsdf/bsdf.h:7-24.
enum LobeType : uint32_t {
    kLobeUnknown = 0,
    kLobeDiffuseReflection = 1u << 1,
    kLobeGlossyReflection = 1u << 3,
    kLobeGlossyTransmission = 1u << 4,
    kLobeDeltaReflection = 1u << 5,
    kLobeDeltaTransmission = 1u << 6,
    kLobeDelta = (1u << 5) | (1u << 6),
};

// Material evaluated at a hit.  c0 is always the albedo slot
// (GetAlbedo, optix_material.h:97-111):
//   diffuse:          c0 = reflectance
//   dielectric:       c0 = specular_reflectance, c1 = specular_transmittance, eta
//   rough dielectric: + alpha
//   conductor:        c0 = specular_reflectance, c1 = eta, c2 = k
//   rough conductor:  + alpha
//   plastic:          c0 = diffuse_reflectance, c1 = specular_reflectance,
//                     eta, int_fdr, spec_weight, nonlinear
//   rough plastic:    + alpha
struct LocalBsdf {
    uint32_t type;
    uint32_t nonlinear;
    float alpha;
    float eta;
    float int_fdr;
    float spec_weight;
    f3 c0, c1, c2;
};

struct BsdfSample {
    f3 wi;
    f3 f;
    float pdf;
    uint32_t lobe;
};

// --- Fresnel (fresnel.h:7-85) ----------------------------------------------
PT_HD float fresnel_dielectric(float eta, float cos_theta_i, float &cos_theta_t) {
    float scale = cos_theta_i > 0.f ? 1.f / eta : eta;
    float cos_theta_t2 = 1.f - (1.f - cos_theta_i * cos_theta_i) * (scale * scale);
    if (cos_theta_t2 <= 0.0f) {
        cos_theta_t = 0.0f;
        return 1.0f;
    }
    float o_cos_theta_i = cos_theta_i;
    cos_theta_i = fabsf(cos_theta_i);
    cos_theta_t = sqrtf(fmaxf(0.f, cos_theta_t2));
    float rs = (cos_theta_i - eta * cos_theta_t) / (cos_theta_i + eta * cos_theta_t);
    float rp = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    cos_theta_t = o_cos_theta_i > 0.f ? -cos_theta_t : cos_theta_t;
    return 0.5f * (rs * rs + rp * rp);
}
PT_HD float fresnel_dielectric(float eta, float cos_theta_i) {
    float ct;
    return fresnel_dielectric(eta, cos_theta_i, ct);
}
PT_HD float fresnel_conductor(float eta, float k, float cos_theta_i) {
    float cos_theta_i2 = cos_theta_i * cos_theta_i;
    float sin_theta_i2 = 1.f - cos_theta_i2;
    float sin_theta_i4 = sin_theta_i2 * sin_theta_i2;
    float t1 = eta * eta - k * k - sin_theta_i2;
    float a2pb2 = sqrtf(fmaxf(0.f, t1 * t1 + 4.f * k * k * eta * eta));
    float a = sqrtf(fmaxf(0.f, 0.5f * (a2pb2 + t1)));
    float term1 = a2pb2 + cos_theta_i2;
    float term2 = 2.f * a * cos_theta_i;
    float rs2 = (term1 - term2) / (term1 + term2);
    float term3 = a2pb2 * cos_theta_i2 + sin_theta_i4;
    float term4 = term2 * sin_theta_i2;
    float rp2 = rs2 * (term3 - term4) / (term3 + term4);
    return 0.5f * (rp2 + rs2);
}
PT_HD f3 fresnel_conductor(f3 eta, f3 k, float cos_theta_i) {
    return f3{fresnel_conductor(eta.x, k.x, cos_theta_i), fresnel_conductor(eta.y, k.y, cos_theta_i),
              fresnel_conductor(eta.z, k.z, cos_theta_i)};
}
// Diffuse Fresnel reflectance fits (fresnel.h:58-85): Egan & Hilgeman 1973
// for eta < 1, d'Eon & Irving 2011 otherwise.  Used on the host only.
PT_HD float fresnel_diffuse_reflectance(float eta) {
    if (eta < 1) {
        return -1.4399f * (eta * eta) + 0.7099f * eta + 0.6681f + 0.0636f / eta;
    }
    float inv_eta = 1.0f / eta;
    float inv_eta2 = inv_eta * inv_eta;
    float inv_eta3 = inv_eta2 * inv_eta;
    float inv_eta4 = inv_eta3 * inv_eta;
    float inv_eta5 = inv_eta4 * inv_eta;
    return 0.919317f - 3.4793f * inv_eta + 6.75335f * inv_eta2 - 7.80989f * inv_eta3 +
           4.98554f * inv_eta4 - 1.36881f * inv_eta5;
}

// --- GGX, isotropic (ggx.h:10-64) -------------------------------------------
PT_HD float ggx_lambda(f3 w, float alpha) {
    float a2 = alpha * alpha;
    f3 v2 = w * w;
    return (-1.f + sqrtf(1.f + (v2.x + v2.y) * a2 / v2.z)) / 2.f;
}
PT_HD float ggx_g1(f3 w, float alpha) { return 1.f / (1.f + ggx_lambda(w, alpha)); }
PT_HD float ggx_g(f3 wi, f3 wo, float alpha) { return ggx_g1(wi, alpha) * ggx_g1(wo, alpha); }
PT_HD float ggx_d(f3 wh, float alpha) {
    float a2 = alpha * alpha;
    f3 v2 = wh * wh;
    float t = (v2.x + v2.y) / a2 + v2.z;
    return 1.f / (kPi * a2 * t * t);
}
// Visible-normal pdf (ggx.h:16-22).
PT_HD float ggx_pdf(f3 wo, f3 wh, float alpha) {
    return ggx_d(wh, alpha) * ggx_g1(wo, alpha) * dot(wo, wh) / fabsf(wo.z);
}
// Visible-normal sampling (Heitz 2018 form, ggx.h:24-42).
PT_HD f3 ggx_sample(f3 wo, float alpha, f2 xi) {
    f3 vh = normalize(f3{alpha * wo.x, alpha * wo.y, wo.z});
    f3 t1v = wo.z < 0.9999f ? normalize(cross(f3{0.f, 0.f, 1.f}, vh)) : f3{1.f, 0.f, 0.f};
    f3 t2v = cross(vh, t1v);
    float r = sqrtf(xi.x);
    float phi = 2.f * kPi * xi.y;
    float s_phi, c_phi;
    sincos_pt(phi, &s_phi, &c_phi);
    float t1 = r * c_phi;
    float t2 = r * s_phi;
    float s = 0.5f * (1.f + vh.z);
    t2 = (1.f - s) * sqrtf(1.f - t1 * t1) + s * t2;
    f3 nh = t1 * t1v + t2 * t2v + sqrtf(fmaxf(0.f, 1.f - t1 * t1 - t2 * t2)) * vh;
    f3 ne = f3{alpha * nh.x, alpha * nh.y, fmaxf(0.f, nh.z)};
    return normalize(ne);
}

// --- per-type Eval (GetBsdf then GetPdf) --------------------------------------
PT_HD f3 plastic_diffuse_term(const LocalBsdf &b) {
    f3 denom = 1.f - (b.nonlinear ? b.c0 * b.int_fdr : mk3(b.int_fdr));
    return b.c0 / denom;
}
PT_HD float plastic_specular_prob(const LocalBsdf &b, float fresnel_o) {
    return (fresnel_o * b.spec_weight) /
           (fresnel_o * b.spec_weight + (1 - fresnel_o) * (1.f - b.spec_weight));
}

// diffuse.h:14-27
PT_HD void eval_diffuse(const LocalBsdf &b, f3 wo, f3 wi, f3 &f, float &pdf) {
    f = mk3(0.f);
    if (wi.z > 0.f && wo.z > 0.f) f = b.c0 * kInvPi;
    pdf = 0.f;
    if (wi.z > 0.f && wo.z > 0.f) pdf = cosine_hemisphere_pdf(wi);
}
// rough_conductor.h:21-38 (the pdf normalizes the half vector twice).
PT_HD void eval_rough_conductor(const LocalBsdf &b, f3 wo, f3 wi, f3 &f, float &pdf) {
    f = mk3(0.f);
    pdf = 0.f;
    if (wi.z <= 0.f || wo.z <= 0.f) return;
    f3 wh = normalize(wi + wo);
    f3 fresnel_o = fresnel_conductor(b.c1, b.c2, dot(wo, wh));
    f = b.c0 * ggx_d(wh, b.alpha) * fresnel_o * ggx_g(wi, wo, b.alpha) / (4.f * wi.z * wo.z);
    f3 wh2 = normalize(normalize(wi + wo));
    pdf = ggx_pdf(wo, wh2, b.alpha) / (4.f * dot(wo, wh2));
}
// rough_dielectric.h:21-71
PT_HD void eval_rough_dielectric(const LocalBsdf &b, f3 wo, f3 wi, f3 &f, float &pdf) {
    const float eta = b.eta;
    const float alpha = b.alpha;
    // GetBsdf
    f = mk3(0.f);
    if (!is_zero(wo.z)) {
        f3 wh;
        bool reflect = wo.z * wi.z > 0.f;
        if (reflect) wh = normalize(wo + wi);
        else wh = normalize(wo + wi * (wo.z > 0.f ? eta : 1.f / eta));
        wh = wh * (wh.z > 0.f ? 1.f : -1.f);
        float F = fresnel_dielectric(eta, dot(wo, wh));
        float G = ggx_g(wi, wo, alpha);
        float D = ggx_d(wh, alpha);
        if (reflect) {
            f = b.c0 * F * G * D / (4.f * fabsf(wi.z) * fabsf(wo.z));
        } else {
            float e = wo.z > 0.f ? eta : 1.f / eta;
            float sqrt_denom = dot(wo, wh) + e * dot(wi, wh);
            f = b.c1 * fabsf((1.f - F) * D * G * dot(wi, wh) * dot(wo, wh) /
                             (sqrt_denom * sqrt_denom * wi.z * wo.z));
        }
    }
    // GetPdf
    {
        bool reflect = wo.z * wi.z > 0.f;
        f3 wh;
        float dwh_dwo;
        if (reflect) {
            wh = normalize(wo + wi);
            dwh_dwo = 1.f / (4.f * dot(wi, wh));
        } else {
            float e = wo.z > 0.f ? eta : 1.f / eta;
            wh = normalize(wo + wi * e);
            float sqrt_denom = dot(wo, wh) + e * dot(wi, wh);
            dwh_dwo = (e * e * dot(wi, wh)) / (sqrt_denom * sqrt_denom);
        }
        wh = wh * (wh.z > 0.f ? 1.f : -1.f);
        f3 wof = wo * (wo.z > 0.f ? 1.f : -1.f);
        float F = fresnel_dielectric(eta, dot(wo, wh));
        pdf = fabsf(ggx_pdf(wof, wh, alpha) * (reflect ? F : 1.f - F) * dwh_dwo);
    }
}
// plastic.h:32-53 (diffuse lobe only)
PT_HD void eval_plastic(const LocalBsdf &b, f3 wo, f3 wi, f3 &f, float &pdf) {
    f = mk3(0.f);
    pdf = 0.f;
    if (wi.z <= 0.f || wo.z <= 0.f) return;
    float fo = fresnel_dielectric(b.eta, wo.z);
    float fi = fresnel_dielectric(b.eta, wi.z);
    f3 diff = plastic_diffuse_term(b);
    f = diff * (1.f - fi) * (1.f - fo) * cosine_hemisphere_pdf(wi) / (b.eta * b.eta * wi.z);
    float sp = plastic_specular_prob(b, fo);
    pdf = cosine_hemisphere_pdf(wi) * (1.f - sp);
}
// rough_plastic.h:31-55
PT_HD void eval_rough_plastic(const LocalBsdf &b, f3 wo, f3 wi, f3 &f, float &pdf) {
    f = mk3(0.f);
    pdf = 0.f;
    if (wi.z <= 0.f || wo.z <= 0.f) return;
    float fo = fresnel_dielectric(b.eta, wo.z);
    f3 wh = normalize(wi + wo);
    f = b.c1 * fresnel_dielectric(b.eta, dot(wh, wo)) * ggx_d(wh, b.alpha) * ggx_g(wi, wo, b.alpha) /
        (4.f * wo.z * wi.z);
    float fi = fresnel_dielectric(b.eta, wi.z);
    f3 diff = plastic_diffuse_term(b);
    f = f + diff * (1.f - fi) * (1.f - fo) * kInvPi / (b.eta * b.eta);
    float sp = plastic_specular_prob(b, fo);
    float dp = 1.f - sp;
    f3 wh2 = normalize(wi + wo);
    pdf = sp * ggx_pdf(wo, wh2, b.alpha) / (4.f * dot(wi, wh2));
    pdf = pdf + dp * cosine_hemisphere_pdf(wi);
}

// LocalBsdf::Eval (optix_material.h:81-91).  Delta BSDFs evaluate to zero.
PT_HD void bsdf_eval(const LocalBsdf &b, f3 wo, f3 wi, f3 &f, float &pdf) {
    switch (b.type) {
        case kMatDiffuse: eval_diffuse(b, wo, wi, f, pdf); return;
        case kMatRoughDielectric: eval_rough_dielectric(b, wo, wi, f, pdf); return;
        case kMatRoughConductor: eval_rough_conductor(b, wo, wi, f, pdf); return;
        case kMatPlastic: eval_plastic(b, wo, wi, f, pdf); return;
        case kMatRoughPlastic: eval_rough_plastic(b, wo, wi, f, pdf); return;
        default: f = mk3(0.f); pdf = 0.f; return;
    }
}

// LocalBsdf::Sample (optix_material.h:70-79).  The record starts as in
// bsdf.h:26-36: f = 0, pdf = 0, lobe Unknown; wi is set on every path that
// returns a nonzero pdf.
PT_HD void bsdf_sample(const LocalBsdf &b, f3 wo, Rng &rng, BsdfSample &s) {
    s.f = mk3(0.f);
    s.pdf = 0.f;
    s.lobe = kLobeUnknown;
    s.wi = mk3(0.f);
    switch (b.type) {
        case kMatDiffuse: {  // diffuse.h:29-35
            float x = rng_next(rng);
            float y = rng_next(rng);
            s.wi = cosine_sample_hemisphere(x, y);
            eval_diffuse(b, wo, s.wi, s.f, s.pdf);
            s.lobe = kLobeDiffuseReflection;
        } break;
        case kMatConductor: {  // conductor.h:27-34
            s.wi = reflect_z(wo);
            s.pdf = 1.f;
            f3 fr = fresnel_conductor(b.c1, b.c2, wo.z);
            s.f = b.c0 * fr / fabsf(s.wi.z);
            s.lobe = kLobeDeltaReflection;
        } break;
        case kMatRoughConductor: {  // rough_conductor.h:40-46
            float x = rng_next(rng);
            float y = rng_next(rng);
            s.wi = reflect_n(wo, ggx_sample(wo, b.alpha, mk2(x, y)));
            eval_rough_conductor(b, wo, s.wi, s.f, s.pdf);
            s.lobe = kLobeDiffuseReflection;
        } break;
        case kMatDielectric: {  // dielectric.h:28-44
            float cos_theta_t;
            float F = fresnel_dielectric(b.eta, wo.z, cos_theta_t);
            if (rng_next(rng) < F) {
                s.wi = reflect_z(wo);
                s.pdf = F;
                s.f = b.c0 * F / fabsf(s.wi.z);
                s.lobe = kLobeDeltaReflection;
            } else {
                s.wi = refract_z(wo, cos_theta_t, b.eta);
                s.pdf = 1.f - F;
                float factor = cos_theta_t < 0.f ? 1.f / b.eta : b.eta;
                s.f = b.c1 * (1.f - F) * factor * factor / fabsf(s.wi.z);
                s.lobe = kLobeDeltaTransmission;
            }
        } break;
        case kMatRoughDielectric: {  // rough_dielectric.h:73-96
            float x = rng_next(rng);
            float y = rng_next(rng);
            f3 wof = wo * (wo.z > 0.f ? 1.f : -1.f);
            f3 wh = ggx_sample(wof, b.alpha, mk2(x, y));
            float cos_theta_t = 0.f;
            float F = fresnel_dielectric(b.eta, dot(wo, wh), cos_theta_t);
            if (rng_next(rng) < F) {
                s.wi = reflect_n(wo, wh);
                s.lobe = kLobeGlossyReflection;
            } else {
                if (is_zero(cos_theta_t)) return;
                s.wi = refract_n(wo, wh, cos_theta_t, b.eta);
                s.lobe = kLobeGlossyTransmission;
                if (s.wi.z * wo.z >= 0.f) return;
            }
            eval_rough_dielectric(b, wo, s.wi, s.f, s.pdf);
        } break;
        case kMatPlastic: {  // plastic.h:55-80
            if (wo.z <= 0.f) return;
            float fo = fresnel_dielectric(b.eta, wo.z);
            float x = rng_next(rng);
            float y = rng_next(rng);
            float sp = plastic_specular_prob(b, fo);
            if (x < sp) {
                s.lobe = kLobeDeltaReflection;
                s.wi = reflect_z(wo);
                s.f = b.c1 * fo / s.wi.z;
                s.pdf = sp;
            } else {
                s.lobe = kLobeDiffuseReflection;
                s.wi = cosine_sample_hemisphere((x - sp) / (1.f - sp), y);
                float fi = fresnel_dielectric(b.eta, s.wi.z);
                f3 diff = plastic_diffuse_term(b);
                s.f = diff * (1.f - fi) * (1.f - fo) * cosine_hemisphere_pdf(s.wi) /
                      (b.eta * b.eta * s.wi.z);
                s.pdf = cosine_hemisphere_pdf(s.wi) * (1.f - sp);
            }
        } break;
        case kMatRoughPlastic: {  // rough_plastic.h:57-80
            if (wo.z <= 0.f) return;
            float fo = fresnel_dielectric(b.eta, wo.z);
            float sp = plastic_specular_prob(b, fo);
            float x = rng_next(rng);
            float y = rng_next(rng);
            if (y < sp) {
                y /= sp;
                f3 wh = ggx_sample(wo, b.alpha, mk2(x, y));
                s.wi = reflect_n(wo, wh);
                s.lobe = kLobeGlossyReflection;
            } else {
                y = (y - sp) / (1.f - sp);
                s.wi = cosine_sample_hemisphere(x, y);
                s.lobe = kLobeDiffuseReflection;
            }
            eval_rough_plastic(b, wo, s.wi, s.f, s.pdf);
        } break;
        default:
            break;
    }
}

}  // namespace pupil
