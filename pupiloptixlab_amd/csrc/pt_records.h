// Device-resident scene records (textures, materials, shapes, emitters,
// camera) and the per-hit functions that read them: texture lookup, material
// to local BSDF, area/constant-environment emitter sampling and evaluation.
//
// References: cuda/texture.h:33-57, render/material/optix_material.h:119-131,
// render/material/bsdf/*.h GetLocal, render/emitter.h:110-136,
// render/emitter/area.h:17-45, render/emitter/sphere.h:14-43,
// render/emitter/env.h:67-86,
// render/camera.h:7-10, example/path_tracer/main.cu:58-78.
#pragma once

#include "pt_bsdf.h"

namespace pupil {

enum TexType : uint32_t { kTexRgb = 0, kTexChecker = 1, kTexBitmap = 2 };

// 80 bytes.  uv transform rows r0, r1 are applied to (u, v, 0, 1).
struct TextureRec {
    uint32_t type;
    uint32_t bitmap;
    uint32_t pad0, pad1;
    f4 c0;  // rgb / checkerboard patch1
    f4 c1;  // checkerboard patch2
    f4 r0;
    f4 r1;
};

enum MaterialFlags : uint32_t { kMatTwosided = 1u, kMatNonlinear = 2u };

struct MaterialRec {
    uint32_t type;
    uint32_t flags;
    float eta;          // int_ior / ext_ior
    float int_fdr;      // plastic: DiffuseReflectance(1 / eta)
    float spec_weight;  // plastic: specular sampling weight
    float pad0, pad1, pad2;
    TextureRec tex[4];  // slot meaning per material type, see LocalBsdf
};

enum ShapeFlags : uint32_t {
    kShapeFlipNormals = 1u,
    kShapeFlipTexcoords = 2u,
    kShapeHasNormals = 4u,
    kShapeHasTexcoords = 8u,
    kShapeNegFaceNormal = 16u,  // instance transform mirrors: flip cross(e1, e2)
};

struct ShapeRec {
    uint32_t material;
    uint32_t flags;
    int32_t emitter_offset;  // emitter index of this shape's first triangle, or -1
    uint32_t first_tri;
};

// EEmitterType (render/emitter/types.h:7-15).
enum EmitterType : uint32_t { kEmitterNone = 0, kEmitterTriArea = 1, kEmitterSphere = 2, kEmitterConstEnv = 3 };

// One emitting triangle (TriAreaEmitter, area.h:4-15) or sphere
// (SphereEmitter, sphere.h:4-12: p0 = center xyz + radius) in world space.
struct EmitterRec {
    uint32_t type;
    float select_prob;
    float area;
    float weight;
    f4 p0, p1, p2;   // positions
    f4 n0, n1, n2;   // unit normals
    f4 uv01;         // uv0.xy, uv1.xy
    f4 uv2;          // uv2.xy
    TextureRec radiance;
};

// Shading record of an analytic sphere (Geometry::Sphere, render/geometry.h:
// 23-27, as instanced by world/render_object.cpp:30-35): the world -> object
// transform of its unit sphere and the owning shape.
struct SphereRec {
    f4 w2o[3];       // rows of the 3x4 world -> object transform
    uint32_t shape;
    uint32_t pad0, pad1, pad2;
};

struct EnvRec {
    uint32_t type;   // kEmitterNone or kEmitterConstEnv
    float select_prob;
    float pad0, pad1;
    f4 color;
};

struct CameraRec {
    f4 s2c[4];  // sample_to_camera rows
    f4 c2w[4];  // camera_to_world rows
};

struct LocalGeometry {
    f3 position;
    f3 normal;
    f2 texcoord;
};

// Texture::Sample (cuda/texture.h:33-57).  Bitmaps are not supported yet and
// read as black.
PT_HD f3 texture_sample(const TextureRec &t, f2 uv) {
    if (t.type == kTexRgb) return mk3(t.c0.x, t.c0.y, t.c0.z);
    const f4 tc = f4{uv.x, uv.y, 0.f, 1.f};
    float tx = dot(t.r0, tc);
    float ty = dot(t.r1, tc);
    if (t.type == kTexChecker) {
        tx = tx - (tx > 0.f ? floorf(tx) : ceilf(tx));
        ty = ty - (ty > 0.f ? floorf(ty) : ceilf(ty));
        if (tx < 0.f) tx += 1.f;
        if (ty < 0.f) ty += 1.f;
        f3 p1 = mk3(t.c0.x, t.c0.y, t.c0.z), p2 = mk3(t.c1.x, t.c1.y, t.c1.z);
        if (tx > 0.5f) return ty > 0.5f ? p1 : p2;
        return ty > 0.5f ? p2 : p1;
    }
    return mk3(0.f);
}

PT_HD bool material_uses_texcoords(const MaterialRec &m) {
    return (m.tex[0].type | m.tex[1].type | m.tex[2].type | m.tex[3].type) != kTexRgb;
}

// Material::GetLocalBsdf (optix_material.h:119-131) with the per-type GetLocal.
PT_HD LocalBsdf material_local_bsdf(const MaterialRec &m, f2 uv) {
    LocalBsdf b;
    b.type = m.type;
    b.nonlinear = (m.flags & kMatNonlinear) ? 1u : 0u;
    b.alpha = 0.f;
    b.eta = m.eta;
    b.int_fdr = m.int_fdr;
    b.spec_weight = m.spec_weight;
    b.c0 = texture_sample(m.tex[0], uv);
    b.c1 = mk3(0.f);
    b.c2 = mk3(0.f);
    switch (m.type) {
        case kMatDielectric:
        case kMatPlastic:
            b.c1 = texture_sample(m.tex[1], uv);
            break;
        case kMatRoughDielectric:
        case kMatRoughPlastic:
            b.c1 = texture_sample(m.tex[1], uv);
            b.alpha = texture_sample(m.tex[2], uv).x;
            break;
        case kMatConductor:
            b.c1 = texture_sample(m.tex[1], uv);
            b.c2 = texture_sample(m.tex[2], uv);
            break;
        case kMatRoughConductor:
            b.c1 = texture_sample(m.tex[1], uv);
            b.c2 = texture_sample(m.tex[2], uv);
            b.alpha = texture_sample(m.tex[3], uv).x;
            break;
        default:
            break;
    }
    return b;
}

// EmitterSampleRecord (emitter/types.h:17-26); is_delta and distance start at
// false / 0 (the reference leaves them unset on the zero-pdf path).
struct EmitterSample {
    f3 radiance;
    f3 wi;
    float distance;
    float pdf;
};

// TriAreaEmitter::SampleDirect (area.h:17-35).
PT_HD void area_sample_direct(const EmitterRec &e, const LocalGeometry &hit, f2 xi, EmitterSample &ret) {
    f3 t = uniform_sample_triangle(xi.x, xi.y);
    f3 p0 = mk3(e.p0.x, e.p0.y, e.p0.z), p1 = mk3(e.p1.x, e.p1.y, e.p1.z), p2 = mk3(e.p2.x, e.p2.y, e.p2.z);
    f3 n0 = mk3(e.n0.x, e.n0.y, e.n0.z), n1 = mk3(e.n1.x, e.n1.y, e.n1.z), n2 = mk3(e.n2.x, e.n2.y, e.n2.z);
    f3 position = p0 * t.x + p1 * t.y + p2 * t.z;
    f3 normal = normalize(n0 * t.x + n1 * t.y + n2 * t.z);
    f2 uv0 = mk2(e.uv01.x, e.uv01.y), uv1 = mk2(e.uv01.z, e.uv01.w), uv2 = mk2(e.uv2.x, e.uv2.y);
    f2 tex = uv0 * t.x + uv1 * t.y + uv2 * t.z;
    ret.radiance = texture_sample(e.radiance, tex);
    ret.wi = normalize(position - hit.position);
    float nol = dot(hit.normal, ret.wi);
    float lnol = dot(normal, -ret.wi);
    ret.pdf = 0.f;
    ret.distance = 0.f;
    if (nol > 0.f && lnol > 0.f) {
        float distance = length(position - hit.position);
        ret.pdf = distance * distance / (lnol * e.area);
        ret.distance = distance;
    }
}

// SphereEmitter::SampleDirect (sphere.h:14-33): a uniform point on the whole
// sphere, texcoord from the world-frame direction.
PT_HD void sphere_sample_direct(const EmitterRec &e, const LocalGeometry &hit, f2 xi, EmitterSample &ret) {
    const f3 t = uniform_sample_sphere(xi.x, xi.y);
    const f3 position = t * e.p0.w + mk3(e.p0.x, e.p0.y, e.p0.z);
    const f3 normal = normalize(t);
    ret.radiance = texture_sample(e.radiance, sphere_texcoord(t));
    ret.wi = normalize(position - hit.position);
    const float nol = dot(hit.normal, ret.wi);
    const float lnol = dot(normal, -ret.wi);
    ret.pdf = 0.f;
    ret.distance = 0.f;
    if (nol > 0.f && lnol > 0.f) {
        const float distance = length(position - hit.position);
        ret.pdf = distance * distance / (lnol * e.area);
        ret.distance = distance;
    }
}

// Emitter::SampleDirect (render/emitter.h:74-90) for the area kinds.
PT_HD void emitter_sample_direct(const EmitterRec &e, const LocalGeometry &hit, f2 xi, EmitterSample &ret) {
    if (e.type == kEmitterSphere) sphere_sample_direct(e, hit, xi, ret);
    else area_sample_direct(e, hit, xi, ret);
}

// Solid-angle density with which emitter_sample_direct produces the emitter
// point `p` (barycentrics b1, b2 on a triangle) seen from `o`: the emitter's
// own outward normal, as in SampleDirect, not the shading normal of the hit.
PT_HD float emitter_nee_pdf(const EmitterRec &e, f3 p, float b1, float b2, f3 o) {
    f3 en;
    if (e.type == kEmitterSphere) {
        en = normalize(p - mk3(e.p0.x, e.p0.y, e.p0.z));
    } else {
        const float wv = 1.f - b1 - b2;
        en = normalize(mk3(e.n0.x, e.n0.y, e.n0.z) * wv + mk3(e.n1.x, e.n1.y, e.n1.z) * b1 +
                       mk3(e.n2.x, e.n2.y, e.n2.z) * b2);
    }
    const f3 dir = normalize(o - p);
    const float nl = dot(en, dir);
    const float dist = length(o - p);
    return nl > 0.f ? dist * dist / (nl * e.area) : 0.f;
}

// TriAreaEmitter::Eval (area.h:37-45) and SphereEmitter::Eval (sphere.h:
// 35-43), which are the same function of the hit geometry; pdf stays 0 when
// the hit faces away.
PT_HD void area_eval(const EmitterRec &e, const LocalGeometry &emit_geo, f3 scatter_pos, f3 &radiance, float &pdf) {
    f3 dir = normalize(scatter_pos - emit_geo.position);
    float lnol = dot(emit_geo.normal, dir);
    pdf = 0.f;
    radiance = mk3(0.f);
    if (lnol > 0.f) {
        float distance = length(scatter_pos - emit_geo.position);
        pdf = distance * distance / (lnol * e.area);
        radiance = texture_sample(e.radiance, emit_geo.texcoord);
    }
}

// ConstEnvEmitter::SampleDirect (env.h:70-80).
PT_HD void const_env_sample_direct(const EnvRec &env, const LocalGeometry &hit, f2 xi, EmitterSample &ret) {
    f3 local_wi = uniform_sample_hemisphere(xi.x, xi.y);
    ret.wi = to_world(local_wi, hit.normal);
    ret.pdf = uniform_hemisphere_pdf(local_wi);
    ret.distance = kMaxDistance;
    ret.radiance = mk3(env.color.x, env.color.y, env.color.z);
}

// Camera ray of main.cu:58-78: d = S2C * (u, v, 0, 1); d /= d.w; d.w = 0;
// d = normalize(d); dir = normalize(C2W * d); origin = C2W column 3.
PT_HD void camera_ray(const CameraRec &cam, uint32_t x, uint32_t y, uint32_t w, uint32_t h, float jx, float jy,
                      f3 &origin, f3 &dir) {
    f4 pf = f4{((float)x + jx) / (float)w, ((float)y + jy) / (float)h, 0.f, 1.f};
    f4 d = f4{dot(cam.s2c[0], pf), dot(cam.s2c[1], pf), dot(cam.s2c[2], pf), dot(cam.s2c[3], pf)};
    float inv_w = 1.0f / d.w;
    d = f4{d.x * inv_w, d.y * inv_w, d.z * inv_w, d.w * inv_w};
    d.w = 0.f;
    d = normalize(d);
    f4 wd = f4{dot(cam.c2w[0], d), dot(cam.c2w[1], d), dot(cam.c2w[2], d), dot(cam.c2w[3], d)};
    dir = normalize(mk3(wd.x, wd.y, wd.z));
    origin = mk3(cam.c2w[0].w, cam.c2w[1].w, cam.c2w[2].w);
}

}  // namespace pupil
