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

// Bitmap sampling modes (util/texture.h:10-20).
enum BitmapFilter : uint32_t { kFilterPoint = 0, kFilterLinear = 1 };
enum BitmapWrap : uint32_t { kWrapRepeat = 0, kWrapClamp = 1, kWrapMirror = 2, kWrapBorder = 3 };

// One bitmap (util::BitmapTexture with its CUDA texture-object settings,
// cuda/texture.cpp:60-100): w x h RGBA texels from texels[offset], row 0 at
// v = 0.
struct BitmapRec {
    uint64_t offset;
    uint32_t w, h;
    uint32_t filter, wrap;
    uint32_t pad0, pad1;
};

// Device arrays that texture lookups read.
struct TexArrays {
    const BitmapRec *bitmaps;
    const f4 *texels;
};

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
enum EmitterType : uint32_t {
    kEmitterNone = 0,
    kEmitterTriArea = 1,
    kEmitterSphere = 2,
    kEmitterConstEnv = 3,
    kEmitterEnvMap = 4
};

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

// The environment: ConstEnvEmitter (env.h:67-86) or EnvMapEmitter (env.h:
// 6-65) with its sampling tables (world/emitter.cpp:107-149).
struct EnvRec {
    uint32_t type;   // kEmitterNone, kEmitterConstEnv or kEmitterEnvMap
    float select_prob;
    uint32_t bitmap; // env map: radiance bitmap (bilinear, repeat)
    float scale;     // env map: radiance scale
    f4 color;        // constant environment radiance
    f4 to_world[3];  // env map frame (rows; xyz used)
    f4 to_local[3];
    uint32_t w, h;   // env map size
    float normalization;
    uint32_t pad;
    const float *row_cdf;     // h + 1 entries
    const float *col_cdf;     // (w + 1) per row
    const float *row_weight;  // sin of each row's center polar angle
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

// Texel index under an address mode; `border` is set for texels outside a
// bordered image.
PT_HD int bitmap_wrap_index(int i, int n, uint32_t mode, bool &border) {
    if (mode == kWrapClamp) return i < 0 ? 0 : (i >= n ? n - 1 : i);
    if (mode == kWrapBorder) {
        if (i < 0 || i >= n) {
            border = true;
            return 0;
        }
        return i;
    }
    if (mode == kWrapMirror) {
        int m = i % (2 * n);
        if (m < 0) m += 2 * n;
        return m < n ? m : 2 * n - 1 - m;
    }
    int m = i % n;
    return m < 0 ? m + n : m;
}

PT_HD f4 bitmap_texel(const TexArrays &tx, const BitmapRec &b, int x, int y) {
    bool border = false;
    const int ix = bitmap_wrap_index(x, (int)b.w, b.wrap, border);
    const int iy = bitmap_wrap_index(y, (int)b.h, b.wrap, border);
    if (border) return f4{1.f, 0.f, 0.f, 0.f};  // borderColor (1, 0, 0, 0), cuda/texture.cpp:91
    return tx.texels[b.offset + (size_t)iy * b.w + (size_t)ix];
}

// The coordinate brought into a range where texel indices fit an int, with
// the same texels selected: repeat by the fractional part, mirror modulo 2,
// clamp / border by clamping to [-1, 2].  Non-finite coordinates read 0.
PT_HD float bitmap_reduce(float u, uint32_t mode) {
    if (!(u >= -3.0e38f && u <= 3.0e38f)) return 0.f;
    if (mode == kWrapRepeat) return u - floorf(u);
    if (mode == kWrapMirror) return u - 2.f * floorf(u * 0.5f);
    return fminf(fmaxf(u, -1.f), 2.f);
}

// tex2D<float4> on a normalized-coordinate texture (cuda/texture.h:52-54):
// point sampling takes texel floor(u w); linear filtering blends the four
// texels around (u w - 0.5, v h - 0.5) with float weights (the hardware's
// 8-bit weight fraction is not reproduced).
PT_HD f3 bitmap_sample(const TexArrays &tx, uint32_t index, float u, float v) {
    const BitmapRec b = tx.bitmaps[index];
    const float fu = bitmap_reduce(u, b.wrap) * (float)b.w;
    const float fv = bitmap_reduce(v, b.wrap) * (float)b.h;
    if (b.filter == kFilterPoint) {
        const f4 t = bitmap_texel(tx, b, (int)floorf(fu), (int)floorf(fv));
        return mk3(t.x, t.y, t.z);
    }
    const float x = fu - 0.5f, y = fv - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float a = x - x0, c = y - y0;
    const int ix = (int)x0, iy = (int)y0;
    const f4 t00 = bitmap_texel(tx, b, ix, iy), t10 = bitmap_texel(tx, b, ix + 1, iy);
    const f4 t01 = bitmap_texel(tx, b, ix, iy + 1), t11 = bitmap_texel(tx, b, ix + 1, iy + 1);
    const float w00 = (1.f - a) * (1.f - c), w10 = a * (1.f - c), w01 = (1.f - a) * c, w11 = a * c;
    return mk3(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
               w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
               w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z);
}

// Scene features a shading kernel is compiled for.  A kernel compiled
// without a feature drops its code (a scene without bitmap textures, an
// environment map or spheres shades with fewer registers); the host picks a
// kernel whose features cover the scene's, so the dropped branches are never
// taken and the arithmetic of the others is unchanged.
constexpr uint32_t kFeatBitmap = 1u;   // bitmap textures (materials, emitters, environment map)
constexpr uint32_t kFeatChecker = 2u;  // checkerboard textures
constexpr uint32_t kFeatEnvMap = 4u;   // environment map emitter
constexpr uint32_t kFeatSphere = 8u;   // analytic spheres (shapes and emitters)
constexpr uint32_t kFeatAll = 15u;

// Texture::Sample (cuda/texture.h:33-57).
PT_HD f3 texture_sample(const TextureRec &t, f2 uv, const TexArrays &tex, uint32_t feat = kFeatAll) {
    if (t.type == kTexRgb || !(feat & (kFeatBitmap | kFeatChecker))) return mk3(t.c0.x, t.c0.y, t.c0.z);
    const f4 tc = f4{uv.x, uv.y, 0.f, 1.f};
    float tx = dot(t.r0, tc);
    float ty = dot(t.r1, tc);
    if ((feat & kFeatChecker) && t.type == kTexChecker) {
        tx = tx - (tx > 0.f ? floorf(tx) : ceilf(tx));
        ty = ty - (ty > 0.f ? floorf(ty) : ceilf(ty));
        if (tx < 0.f) tx += 1.f;
        if (ty < 0.f) ty += 1.f;
        f3 p1 = mk3(t.c0.x, t.c0.y, t.c0.z), p2 = mk3(t.c1.x, t.c1.y, t.c1.z);
        if (tx > 0.5f) return ty > 0.5f ? p1 : p2;
        return ty > 0.5f ? p2 : p1;
    }
    if (!(feat & kFeatBitmap)) return mk3(t.c0.x, t.c0.y, t.c0.z);
    return bitmap_sample(tex, t.bitmap, tx, ty);
}

PT_HD bool material_uses_texcoords(const MaterialRec &m) {
    return (m.tex[0].type | m.tex[1].type | m.tex[2].type | m.tex[3].type) != kTexRgb;
}

// Material::GetLocalBsdf (optix_material.h:119-131) with the per-type GetLocal.
PT_HD LocalBsdf material_local_bsdf(const MaterialRec &m, f2 uv, const TexArrays &tx, uint32_t feat = kFeatAll) {
    LocalBsdf b;
    b.type = m.type;
    b.nonlinear = (m.flags & kMatNonlinear) ? 1u : 0u;
    b.alpha = 0.f;
    b.eta = m.eta;
    b.int_fdr = m.int_fdr;
    b.spec_weight = m.spec_weight;
    b.c0 = texture_sample(m.tex[0], uv, tx, feat);
    b.c1 = mk3(0.f);
    b.c2 = mk3(0.f);
    switch (m.type) {
        case kMatDielectric:
        case kMatPlastic:
            b.c1 = texture_sample(m.tex[1], uv, tx, feat);
            break;
        case kMatRoughDielectric:
        case kMatRoughPlastic:
            b.c1 = texture_sample(m.tex[1], uv, tx, feat);
            b.alpha = texture_sample(m.tex[2], uv, tx, feat).x;
            break;
        case kMatConductor:
            b.c1 = texture_sample(m.tex[1], uv, tx, feat);
            b.c2 = texture_sample(m.tex[2], uv, tx, feat);
            break;
        case kMatRoughConductor:
            b.c1 = texture_sample(m.tex[1], uv, tx, feat);
            b.c2 = texture_sample(m.tex[2], uv, tx, feat);
            b.alpha = texture_sample(m.tex[3], uv, tx, feat).x;
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
PT_HD void area_sample_direct(const EmitterRec &e, const LocalGeometry &hit, f2 xi, EmitterSample &ret,
                              const TexArrays &tx, uint32_t feat = kFeatAll) {
    f3 t = uniform_sample_triangle(xi.x, xi.y);
    f3 p0 = mk3(e.p0.x, e.p0.y, e.p0.z), p1 = mk3(e.p1.x, e.p1.y, e.p1.z), p2 = mk3(e.p2.x, e.p2.y, e.p2.z);
    f3 n0 = mk3(e.n0.x, e.n0.y, e.n0.z), n1 = mk3(e.n1.x, e.n1.y, e.n1.z), n2 = mk3(e.n2.x, e.n2.y, e.n2.z);
    f3 position = p0 * t.x + p1 * t.y + p2 * t.z;
    f3 normal = normalize(n0 * t.x + n1 * t.y + n2 * t.z);
    f2 uv0 = mk2(e.uv01.x, e.uv01.y), uv1 = mk2(e.uv01.z, e.uv01.w), uv2 = mk2(e.uv2.x, e.uv2.y);
    f2 tex = uv0 * t.x + uv1 * t.y + uv2 * t.z;
    ret.radiance = texture_sample(e.radiance, tex, tx, feat);
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
PT_HD void sphere_sample_direct(const EmitterRec &e, const LocalGeometry &hit, f2 xi, EmitterSample &ret,
                                const TexArrays &tx, uint32_t feat = kFeatAll) {
    const f3 t = uniform_sample_sphere(xi.x, xi.y);
    const f3 position = t * e.p0.w + mk3(e.p0.x, e.p0.y, e.p0.z);
    const f3 normal = normalize(t);
    ret.radiance = texture_sample(e.radiance, sphere_texcoord(t), tx, feat);
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
PT_HD void emitter_sample_direct(const EmitterRec &e, const LocalGeometry &hit, f2 xi, EmitterSample &ret,
                                 const TexArrays &tx, uint32_t feat = kFeatAll) {
    if ((feat & kFeatSphere) && e.type == kEmitterSphere) sphere_sample_direct(e, hit, xi, ret, tx, feat);
    else area_sample_direct(e, hit, xi, ret, tx, feat);
}

// Solid-angle density with which emitter_sample_direct produces the emitter
// point `p` (barycentrics b1, b2 on a triangle) seen from `o`: the emitter's
// own outward normal, as in SampleDirect, not the shading normal of the hit.
PT_HD float emitter_nee_pdf(const EmitterRec &e, f3 p, float b1, float b2, f3 o, uint32_t feat = kFeatAll) {
    f3 en;
    if ((feat & kFeatSphere) && e.type == kEmitterSphere) {
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
PT_HD void area_eval(const EmitterRec &e, const LocalGeometry &emit_geo, f3 scatter_pos, f3 &radiance, float &pdf,
                     const TexArrays &tx, uint32_t feat = kFeatAll) {
    f3 dir = normalize(scatter_pos - emit_geo.position);
    float lnol = dot(emit_geo.normal, dir);
    pdf = 0.f;
    radiance = mk3(0.f);
    if (lnol > 0.f) {
        float distance = length(scatter_pos - emit_geo.position);
        pdf = distance * distance / (lnol * e.area);
        radiance = texture_sample(e.radiance, emit_geo.texcoord, tx, feat);
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

// --- environment map (EnvMapEmitter, env.h:6-65) ----------------------------------
// Map coordinates: local direction (sin t sin(pi - p), cos t, sin t cos(pi - p))
// for azimuth p in [0, 2pi) and polar angle t in [0, pi]; texcoord (p / 2pi,
// t / pi).
PT_HD f3 env_local_dir(float phi, float theta) {
    float st, ct, sp, cp;
    sincos_pt(theta, &st, &ct);
    sincos_pt(kPi - phi, &sp, &cp);
    return mk3(st * sp, ct, st * cp);
}

PT_HD f3 env_rows(const f4 *r, f3 v) {
    return mk3(r[0].x * v.x + r[0].y * v.y + r[0].z * v.z, r[1].x * v.x + r[1].y * v.y + r[1].z * v.z,
               r[2].x * v.x + r[2].y * v.y + r[2].z * v.z);
}

// Texcoord of a world direction (EnvMapEmitter::Eval, env.h:52-56); also
// returns the polar angle.
PT_HD f2 env_dir_texcoord(const EnvRec &e, f3 dir, float &theta) {
    const f3 d = env_rows(e.to_local, dir);
    const float phi = kPi - atan2_pt(d.x, d.z);
    theta = acos_pt(d.y);
    return f2{phi * 0.5f * kInvPi, theta * kInvPi};
}

PT_HD f3 env_map_radiance(const EnvRec &e, f2 tex, const TexArrays &tx) {
    return bitmap_sample(tx, e.bitmap, tex.x, tex.y) * e.scale;
}

// First index i in [lo, hi) with x <= cdf[i], or hi: the linear scans of
// env.h:24-31 on a non-decreasing table.
PT_HD uint32_t cdf_search(const float *cdf, uint32_t lo, uint32_t hi, float x) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (x <= cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Exact density of env_map_sample_exact over solid angle: the texel's share
// of sum(luminance x row weight), spread uniformly over the texel's (phi,
// theta) rectangle.
PT_HD float env_map_pdf_exact(const EnvRec &e, f2 tex, float theta, const TexArrays &tx) {
    uint32_t x = (uint32_t)fminf(fmaxf(tex.x * (float)e.w, 0.f), (float)(e.w - 1u));
    uint32_t y = (uint32_t)fminf(fmaxf(tex.y * (float)e.h, 0.f), (float)(e.h - 1u));
    const BitmapRec b = tx.bitmaps[e.bitmap];
    const f4 t = tx.texels[b.offset + (size_t)y * b.w + x];
    const float st = sin_pt(theta);
    if (!(st > 0.f)) return 0.f;
    return luminance(mk3(t.x, t.y, t.z)) * e.row_weight[y] * e.normalization / st;
}

// EnvMapEmitter::Eval (env.h:51-63): radiance and pdf of a direction.  The
// reference pdf interpolates the row weight and uses the filtered radiance
// (scaled), which is what env.h computes; the exact pdf is the density of
// env_map_sample_exact.
PT_HD void env_map_eval(const EnvRec &e, f3 dir, bool exact, const TexArrays &tx, f3 &radiance, float &pdf) {
    float theta;
    const f2 tex = env_dir_texcoord(e, dir, theta);
    radiance = env_map_radiance(e, tex, tx);
    if (exact) {
        pdf = env_map_pdf_exact(e, tex, theta, tx);
        return;
    }
    uint32_t row = (uint32_t)(tex.y * (float)e.h);
    if (row > e.h - 2u) row = e.h - 2u;
    const float a = e.row_weight[row], b = e.row_weight[row + 1u];
    const float t = tex.y * (float)e.h - 1.f * (float)row;
    pdf = luminance(radiance) * (a + t * (b - a)) * e.normalization / fmaxf(1e-4f, fabsf(sin_pt(theta)));
}

// EnvMapEmitter::SampleDirect (env.h:23-49): the row / column scans stop at
// the first table entry not below the draw, and the direction is that
// entry's texel corner.  The scan of env.h can run one row past the table
// (a draw above the second-to-last row entry); that row is clamped to the
// last one here.
PT_HD void env_map_sample_reference(const EnvRec &e, f2 xi, EmitterSample &ret, const TexArrays &tx) {
    uint32_t row = cdf_search(e.row_cdf, 0u, e.h, xi.x);
    if (row > e.h - 1u) row = e.h - 1u;
    const uint32_t col = cdf_search(e.col_cdf + (size_t)row * (e.w + 1u), 0u, e.w - 1u, xi.y);
    const float phi = (float)col * kPi * 2.f / (float)e.w;
    const float theta = (float)row * kPi / (float)e.h;
    ret.wi = env_rows(e.to_world, env_local_dir(phi, theta));
    ret.distance = kMaxDistance;
    ret.radiance = env_map_radiance(e, f2{phi * 0.5f * kInvPi, theta * kInvPi}, tx);
    ret.pdf = luminance(ret.radiance) * e.row_weight[row] * e.normalization / fmaxf(1e-4f, fabsf(sin_pt(theta)));
    if (ret.pdf < 0.f) ret.pdf = 0.f;
}

// Continuous sampling of the same tables: row and column by their cdf
// intervals, the position inside the texel from the remainder of each draw;
// pdf = env_map_pdf_exact of the sampled direction.
PT_HD void env_map_sample_exact(const EnvRec &e, f2 xi, EmitterSample &ret, const TexArrays &tx) {
    uint32_t y = cdf_search(e.row_cdf, 1u, e.h + 1u, xi.x) - 1u;
    if (y > e.h - 1u) y = e.h - 1u;
    const float r0 = e.row_cdf[y], r1 = e.row_cdf[y + 1u];
    const float fy = r1 > r0 ? fminf(fmaxf((xi.x - r0) / (r1 - r0), 0.f), 1.f) : 0.5f;
    const float *cc = e.col_cdf + (size_t)y * (e.w + 1u);
    uint32_t x = cdf_search(cc, 1u, e.w + 1u, xi.y) - 1u;
    if (x > e.w - 1u) x = e.w - 1u;
    const float c0 = cc[x], c1 = cc[x + 1u];
    const float fx = c1 > c0 ? fminf(fmaxf((xi.y - c0) / (c1 - c0), 0.f), 1.f) : 0.5f;
    const float phi = ((float)x + fx) * kPi * 2.f / (float)e.w;
    const float theta = ((float)y + fy) * kPi / (float)e.h;
    ret.wi = env_rows(e.to_world, env_local_dir(phi, theta));
    ret.distance = kMaxDistance;
    float pdf;
    env_map_eval(e, ret.wi, true, tx, ret.radiance, pdf);
    ret.pdf = pdf;
}

// Next-event sample of the environment (ConstEnv or EnvMap).
PT_HD void env_sample_direct(const EnvRec &env, const LocalGeometry &hit, f2 xi, bool exact, EmitterSample &ret,
                             const TexArrays &tx, uint32_t feat = kFeatAll) {
    if ((feat & kFeatEnvMap) && env.type == kEmitterEnvMap) {
        if (exact) env_map_sample_exact(env, xi, ret, tx);
        else env_map_sample_reference(env, xi, ret, tx);
    } else {
        const_env_sample_direct(env, hit, xi, ret);
    }
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
