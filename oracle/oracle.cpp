// CPU oracle of the MIS path tracer -- TEST INFRASTRUCTURE ONLY.
//
// Scalar restatement of the reference path tracer, one pixel sample at a time,
// following the control flow of example/path_tracer/main.cu:38-197.  Used only
// by tests/, __graft_entry__.smoke() and bench.py's CPU baseline as the
// checker; the product never links it.  Built with -ffp-contract=off so each
// expression rounds as written (IEEE division and square root, no fused
// multiply-add).
//
// Parity status of each part is recorded next to it ("pinned by ...").
#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Vector math with the semantics of cuda/vec_math.h (vector / scalar is a
// multiply by the reciprocal; dot sums left to right).
struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 v3s(float s) { return V3{s, s, s}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 mul(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 mul(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 divv(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 divs(V3 a, float s) { float inv = 1.0f / s; return mul(a, inv); }
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float dot4(V4 a, V4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
inline V3 cross3(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len3(V3 v) { return sqrtf(dot3(v, v)); }
inline V3 norm3(V3 v) { float inv = 1.0f / sqrtf(dot3(v, v)); return mul(v, inv); }
inline V2 add2(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
inline V2 mul2(V2 a, float s) { return V2{a.x * s, a.y * s}; }
inline bool iszero(float v) { return fabsf(v) < 0.000001f; }
inline bool iszero3(V3 v) { return iszero(v.x) && iszero(v.y) && iszero(v.z); }
inline float misw(float x, float y) { return x / (x + y); }
inline float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

const float PI_F = 3.14159265358979323846f;
const float INV_PI_F = 0.318309886183790671538f;

// sin/cos by Cody-Waite reduction modulo pi/2 (three-part constant, leading
// part 1.5703125) and Taylor polynomials on [-pi/4, pi/4].
// Pinned by: tests compare against numpy float64 sin/cos (<= 2 ulp).
void sincos_or(float x, float &s, float &c) {
    float q = rintf(x * 0.636619772367581343076f);
    float r = ((x - q * 1.5703125f) - q * 4.837512969970703125e-4f) - q * 7.54978995489188216e-8f;
    float r2 = r * r;
    float ps = -1.984126984126984e-4f + r2 * 2.755731922398589e-6f;
    ps = 8.333333333333333e-3f + r2 * ps;
    ps = -1.666666666666667e-1f + r2 * ps;
    float sr = r + r * (r2 * ps);
    float pc = 2.480158730158730e-5f + r2 * -2.755731922398589e-7f;
    pc = -1.388888888888889e-3f + r2 * pc;
    pc = 4.166666666666667e-2f + r2 * pc;
    pc = -0.5f + r2 * pc;
    float cr = 1.0f + r2 * pc;
    switch (((int)q) & 3) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}

// cuda/random.h:14-36.  Pinned by: golden vectors (tests/golden/rng.json) and
// the LCG / TEA definitions restated in numpy in the tests.
struct Random {
    uint32_t seed;
    void init(uint32_t v0, uint32_t v1) {
        uint32_t s0 = 0;
        for (uint32_t n = 0; n < 4; n++) {
            s0 += 0x9e3779b9u;
            v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
            v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
        }
        seed = v0;
    }
    float next() {
        seed = 1664525u * seed + 1013904223u;
        return (float)(seed & 0x00FFFFFFu) / 16777216.0f;
    }
};

// optix/util.h:74-156.  Pinned by: pdf normalization / orthonormality tests.
V3 sample_tri(float u1, float u2) { float su = sqrtf(u1); return v3(1.f - su, su * (1.f - u2), u2 * su); }
V3 cos_hemi(float u1, float u2) {
    float st = sqrtf(u1), phi = 2.0f * PI_F * u2, s, c;
    sincos_or(phi, s, c);
    return v3(st * c, st * s, sqrtf(fmaxf(0.f, 1.f - st * st)));
}
float cos_hemi_pdf(V3 v) { return v.z > 0.f ? INV_PI_F * v.z : 0.f; }
V3 uni_hemi(float u1, float u2) {
    float z = 1.f - 2.f * u1, st = sqrtf(fmaxf(0.f, 1.f - z * z)), phi = 2.0f * PI_F * u2, s, c;
    sincos_or(phi, s, c);
    return v3(st * c, st * s, fabsf(z));
}
float uni_hemi_pdf(V3 v) { return v.z > 0.f ? INV_PI_F * 0.5f : 0.f; }
V3 uni_sphere(float u1, float u2) {  // UniformSampleSphere (optix/util.h:38-43)
    float z = 1.f - 2.f * u1, st = sqrtf(fmaxf(0.f, 1.f - z * z)), phi = 2.0f * PI_F * u2, s, c;
    sincos_or(phi, s, c);
    return v3(st * c, st * s, z);
}

// atan on [0, 1] from the Taylor series of atan (through u^19) after the
// reduction atan(a) = pi/4 + atan((a-1)/(a+1)) for a > tan(pi/8); atan2 by
// octant symmetry; acos(z) = atan2(sqrt((1-z)(1+z)), z).
// Pinned by: tests against numpy float64 arctan2/arccos (a few ulp).
float atan01(float a) {
    float base = 0.f, u = a;
    if (a > 0.41421356f) { u = (a - 1.f) / (a + 1.f); base = 0.785398163397448f; }
    const float u2 = u * u;
    const float coef[8] = {1.f / 17.f, -1.f / 15.f, 1.f / 13.f, -1.f / 11.f, 1.f / 9.f, -1.f / 7.f, 1.f / 5.f, -1.f / 3.f};
    float p = -1.f / 19.f;
    for (int i = 0; i < 8; ++i) p = coef[i] + u2 * p;
    return base + (u + u * (u2 * p));
}
float atan2_or(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    float hi = fmaxf(ax, ay), lo = fminf(ax, ay);
    if (hi == 0.f) return 0.f;
    float r = atan01(lo / hi);
    if (ay > ax) r = 1.57079632679490f - r;
    if (x < 0.f) r = 3.14159265358979f - r;
    return y < 0.f ? -r : r;
}
float acos_or(float z) {
    z = fminf(fmaxf(z, -1.f), 1.f);
    return atan2_or(sqrtf((1.f - z) * (1.f + z)), z);
}
V2 sphere_uv(V3 p) {  // GetSphereTexcoord (optix/util.h:117-128)
    float phi = atan2_or(p.y, p.x);
    if (phi < 0.f) phi = phi + PI_F * 2.f;
    float theta = acos_or(p.z);
    return V2{phi * INV_PI_F * 0.5f, theta * INV_PI_F};
}
V3 reflect_local(V3 v) { return v3(-v.x, -v.y, v.z); }
V3 reflect_about(V3 v, V3 n) { return add(neg(v), mul(2.f * dot3(v, n), n)); }
V3 refract_local(V3 v, float ct, float eta) {
    float sc = -(ct < 0.f ? 1.f / eta : eta);
    return norm3(v3(sc * v.x, sc * v.y, ct));
}
V3 refract_about(V3 v, V3 n, float ct, float eta) {
    if (ct < 0) eta = 1 / eta;
    return sub(mul(n, dot3(v, n) * eta + ct), mul(v, eta));
}
void onb(V3 n, V3 &b1, V3 &b2) {
    float sign = copysignf(1.f, n.z);
    float a = -1.f / (sign + n.z);
    float b = n.x * n.y * a;
    b1 = v3(1.f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b2 = v3(b, sign + n.y * n.y * a, -n.y);
}
V3 to_local(V3 v, V3 n) { V3 b1, b2; onb(n, b1, b2); return v3(dot3(v, b1), dot3(v, b2), dot3(v, n)); }
V3 to_world(V3 v, V3 n) { V3 b1, b2; onb(n, b1, b2); return add(add(mul(b1, v.x), mul(b2, v.y)), mul(n, v.z)); }

// fresnel.h:7-85.  Pinned by: normal incidence closed form, TIR, and the
// textbook complex-index formula for conductors (tests).
float fr_diel(float eta, float ci, float &ct) {
    float scale = ci > 0.f ? 1.f / eta : eta;
    float ct2 = 1.f - (1.f - ci * ci) * (scale * scale);
    if (ct2 <= 0.0f) { ct = 0.0f; return 1.0f; }
    float oci = ci;
    ci = fabsf(ci);
    ct = sqrtf(fmaxf(0.f, ct2));
    float rs = (ci - eta * ct) / (ci + eta * ct);
    float rp = (eta * ci - ct) / (eta * ci + ct);
    ct = oci > 0.f ? -ct : ct;
    return 0.5f * (rs * rs + rp * rp);
}
float fr_diel(float eta, float ci) { float ct; return fr_diel(eta, ci, ct); }
float fr_cond(float eta, float k, float ci) {
    float ci2 = ci * ci, si2 = 1.f - ci2, si4 = si2 * si2;
    float t1 = eta * eta - k * k - si2;
    float a2pb2 = sqrtf(fmaxf(0.f, t1 * t1 + 4.f * k * k * eta * eta));
    float a = sqrtf(fmaxf(0.f, 0.5f * (a2pb2 + t1)));
    float term1 = a2pb2 + ci2, term2 = 2.f * a * ci;
    float rs2 = (term1 - term2) / (term1 + term2);
    float term3 = a2pb2 * ci2 + si4, term4 = term2 * si2;
    float rp2 = rs2 * (term3 - term4) / (term3 + term4);
    return 0.5f * (rp2 + rs2);
}
V3 fr_cond3(V3 eta, V3 k, float ci) { return v3(fr_cond(eta.x, k.x, ci), fr_cond(eta.y, k.y, ci), fr_cond(eta.z, k.z, ci)); }
float fr_diffuse(float eta) {
    if (eta < 1) return -1.4399f * (eta * eta) + 0.7099f * eta + 0.6681f + 0.0636f / eta;
    float i1 = 1.0f / eta, i2 = i1 * i1, i3 = i2 * i1, i4 = i3 * i1, i5 = i4 * i1;
    return 0.919317f - 3.4793f * i1 + 6.75335f * i2 - 7.80989f * i3 + 4.98554f * i4 - 1.36881f * i5;
}

// ggx.h:10-64 (visible-normal variant).  Pinned by: integral of D cos = 1 and
// VNDF pdf normalization (tests, Monte Carlo).
float ggx_lambda(V3 w, float a) { float a2 = a * a; V3 v2 = mul(w, w); return (-1.f + sqrtf(1.f + (v2.x + v2.y) * a2 / v2.z)) / 2.f; }
float ggx_g1(V3 w, float a) { return 1.f / (1.f + ggx_lambda(w, a)); }
float ggx_g(V3 wi, V3 wo, float a) { return ggx_g1(wi, a) * ggx_g1(wo, a); }
float ggx_d(V3 wh, float a) {
    float a2 = a * a; V3 v2 = mul(wh, wh);
    float t = (v2.x + v2.y) / a2 + v2.z;
    return 1.f / (PI_F * a2 * t * t);
}
float ggx_pdf(V3 wo, V3 wh, float a) { return ggx_d(wh, a) * ggx_g1(wo, a) * dot3(wo, wh) / fabsf(wo.z); }
V3 ggx_sample(V3 wo, float a, V2 xi) {
    V3 vh = norm3(v3(a * wo.x, a * wo.y, wo.z));
    V3 T1 = wo.z < 0.9999f ? norm3(cross3(v3(0.f, 0.f, 1.f), vh)) : v3(1.f, 0.f, 0.f);
    V3 T2 = cross3(vh, T1);
    float r = sqrtf(xi.x), phi = 2.f * PI_F * xi.y, sp, cp;
    sincos_or(phi, sp, cp);
    float t1 = r * cp, t2 = r * sp;
    float s = 0.5f * (1.f + vh.z);
    t2 = (1.f - s) * sqrtf(1.f - t1 * t1) + s * t2;
    V3 nh = add(add(mul(t1, T1), mul(t2, T2)), mul(sqrtf(fmaxf(0.f, 1.f - t1 * t1 - t2 * t2)), vh));
    return norm3(v3(a * nh.x, a * nh.y, fmaxf(0.f, nh.z)));
}

// ---------------------------------------------------------------------------
// Materials: optix_material.{h,cpp} + bsdf/*.h.  Pinned by: sample/eval
// consistency, pdf normalization, reciprocity and white-furnace bounds
// (tests), plus golden vectors.
enum { DIFFUSE = 1, DIELECTRIC, ROUGH_DIELECTRIC, CONDUCTOR, ROUGH_CONDUCTOR, PLASTIC, ROUGH_PLASTIC };
enum { LOBE_DIFFUSE_R = 1 << 1, LOBE_GLOSSY_R = 1 << 3, LOBE_GLOSSY_T = 1 << 4, LOBE_DELTA_R = 1 << 5, LOBE_DELTA_T = 1 << 6 };

// Bitmap lookup with the CUDA texture-object semantics of cuda/texture.cpp:
// 60-100 (normalized coordinates; point = texel floor(u w); bilinear = the
// four texels around (u w - 1/2, v h - 1/2) with float weights; repeat /
// clamp / mirror / border addressing, border color (1, 0, 0)).  The
// coordinate is first reduced (fractional part, modulo 2, or clamped to
// [-1, 2]) so texel indices fit an int.  Pinned by: texel-center, edge and
// wrap closed forms (tests).
int wrap_idx(int i, int n, uint32_t mode, bool &border) {
    switch (mode) {
        case 1: return std::min(std::max(i, 0), n - 1);
        case 3: if (i < 0 || i >= n) { border = true; return 0; } return i;
        case 2: { int m = ((i % (2 * n)) + 2 * n) % (2 * n); return m < n ? m : 2 * n - 1 - m; }
        default: return ((i % n) + n) % n;
    }
}
V4 texel(const or_bitmap &b, int x, int y) {
    bool border = false;
    int ix = wrap_idx(x, (int)b.width, b.wrap, border), iy = wrap_idx(y, (int)b.height, b.wrap, border);
    if (border) return V4{1.f, 0.f, 0.f, 0.f};
    const float *p = b.rgba + 4 * ((size_t)iy * b.width + (size_t)ix);
    return V4{p[0], p[1], p[2], p[3]};
}
float reduce_coord(float u, uint32_t mode) {
    if (!(u >= -3.0e38f && u <= 3.0e38f)) return 0.f;
    if (mode == 0) return u - floorf(u);
    if (mode == 2) return u - 2.f * floorf(u * 0.5f);
    return fminf(fmaxf(u, -1.f), 2.f);
}
V3 bitmap_lookup(const or_bitmap &b, float u, float v) {
    const float fu = reduce_coord(u, b.wrap) * (float)b.width, fv = reduce_coord(v, b.wrap) * (float)b.height;
    if (b.filter == 0) {
        V4 t = texel(b, (int)floorf(fu), (int)floorf(fv));
        return v3(t.x, t.y, t.z);
    }
    const float x = fu - 0.5f, y = fv - 0.5f, x0 = floorf(x), y0 = floorf(y), a = x - x0, c = y - y0;
    const int i = (int)x0, j = (int)y0;
    V4 t00 = texel(b, i, j), t10 = texel(b, i + 1, j), t01 = texel(b, i, j + 1), t11 = texel(b, i + 1, j + 1);
    float w00 = (1.f - a) * (1.f - c), w10 = a * (1.f - c), w01 = (1.f - a) * c, w11 = a * c;
    return v3(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x, w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
              w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z);
}

V3 tex_sample(const or_texture &t, V2 uv, const or_bitmap *bm) {
    if (t.type == 0) return v3(t.color0[0], t.color0[1], t.color0[2]);
    V4 tc = V4{uv.x, uv.y, 0.f, 1.f};
    float tx = dot4(V4{t.uv_row0[0], t.uv_row0[1], t.uv_row0[2], t.uv_row0[3]}, tc);
    float ty = dot4(V4{t.uv_row1[0], t.uv_row1[1], t.uv_row1[2], t.uv_row1[3]}, tc);
    if (t.type == 1) {  // cuda/texture.h:42-51
        tx = tx - (tx > 0.f ? floorf(tx) : ceilf(tx));
        ty = ty - (ty > 0.f ? floorf(ty) : ceilf(ty));
        if (tx < 0.f) tx += 1.f;
        if (ty < 0.f) ty += 1.f;
        V3 p1 = v3(t.color0[0], t.color0[1], t.color0[2]), p2 = v3(t.color1[0], t.color1[1], t.color1[2]);
        if (tx > 0.5f) return ty > 0.5f ? p1 : p2;
        return ty > 0.5f ? p2 : p1;
    }
    return bitmap_lookup(bm[t.bitmap], tx, ty);
}
V3 tex_average(const or_texture &t, const or_bitmap *bm) {  // optix_material.cpp:9-36
    if (t.type == 1) return mul(v3(t.color0[0] + t.color1[0], t.color0[1] + t.color1[1], t.color0[2] + t.color1[2]), 0.5f);
    if (t.type == 2) {
        const or_bitmap &b = bm[t.bitmap];
        float r = 0.f, g = 0.f, bl = 0.f;
        for (size_t i = 0; i < (size_t)b.width * b.height; ++i) {
            r += b.rgba[4 * i];
            g += b.rgba[4 * i + 1];
            bl += b.rgba[4 * i + 2];
        }
        return divs(v3(r, g, bl), 1.f * (float)b.height * (float)b.width);
    }
    return v3(t.color0[0], t.color0[1], t.color0[2]);
}

struct Bsdf {
    int type;
    bool nonlinear;
    float alpha, eta, int_fdr, spec_w;
    V3 diffuse_refl, spec_refl, spec_trans, cond_eta, cond_k;
};

// Material::GetLocalBsdf with host precompute of optix_material.cpp:88-119.
Bsdf make_bsdf(const or_material &m, V2 uv, const or_bitmap *bm) {
    Bsdf b;
    std::memset(&b, 0, sizeof(b));
    b.type = (int)m.type;
    b.nonlinear = m.nonlinear != 0;
    b.eta = m.int_ior / m.ext_ior;
    switch (m.type) {
        case DIFFUSE: b.diffuse_refl = tex_sample(m.tex[0], uv, bm); break;
        case DIELECTRIC:
            b.spec_refl = tex_sample(m.tex[0], uv, bm);
            b.spec_trans = tex_sample(m.tex[1], uv, bm);
            break;
        case ROUGH_DIELECTRIC:
            b.spec_refl = tex_sample(m.tex[0], uv, bm);
            b.spec_trans = tex_sample(m.tex[1], uv, bm);
            b.alpha = tex_sample(m.tex[2], uv, bm).x;
            break;
        case CONDUCTOR:
            b.spec_refl = tex_sample(m.tex[0], uv, bm);
            b.cond_eta = tex_sample(m.tex[1], uv, bm);
            b.cond_k = tex_sample(m.tex[2], uv, bm);
            break;
        case ROUGH_CONDUCTOR:
            b.spec_refl = tex_sample(m.tex[0], uv, bm);
            b.cond_eta = tex_sample(m.tex[1], uv, bm);
            b.cond_k = tex_sample(m.tex[2], uv, bm);
            b.alpha = tex_sample(m.tex[3], uv, bm).x;
            break;
        case PLASTIC:
        case ROUGH_PLASTIC: {
            b.diffuse_refl = tex_sample(m.tex[0], uv, bm);
            b.spec_refl = tex_sample(m.tex[1], uv, bm);
            if (m.type == ROUGH_PLASTIC) b.alpha = tex_sample(m.tex[2], uv, bm).x;
            float dl = lum(tex_average(m.tex[0], bm)), sl = lum(tex_average(m.tex[1], bm));
            b.spec_w = sl / (sl + dl);
            b.int_fdr = fr_diffuse(1.f / b.eta);
        } break;
        default: break;
    }
    return b;
}
V3 albedo_of(const Bsdf &b) {
    switch (b.type) {
        case DIFFUSE: case PLASTIC: case ROUGH_PLASTIC: return b.diffuse_refl;
        case DIELECTRIC: case ROUGH_DIELECTRIC: case CONDUCTOR: case ROUGH_CONDUCTOR: return b.spec_refl;
        default: return v3s(0.f);
    }
}

struct BRec { V3 wi, wo, f; float pdf; int lobe; };

V3 plastic_diff(const Bsdf &b) {
    V3 den = sub(v3s(1.f), b.nonlinear ? mul(b.diffuse_refl, b.int_fdr) : v3s(b.int_fdr));
    return divv(b.diffuse_refl, den);
}
float plastic_sp(const Bsdf &b, float fo) {
    return (fo * b.spec_w) / (fo * b.spec_w + (1 - fo) * (1.f - b.spec_w));
}

void bsdf_f(const Bsdf &b, BRec &r) {  // GetBsdf
    r.f = v3s(0.f);
    V3 wi = r.wi, wo = r.wo;
    switch (b.type) {
        case DIFFUSE:
            if (wi.z > 0.f && wo.z > 0.f) r.f = mul(b.diffuse_refl, INV_PI_F);
            break;
        case ROUGH_CONDUCTOR: {
            if (wi.z <= 0.f || wo.z <= 0.f) return;
            V3 wh = norm3(add(wi, wo));
            V3 fo = fr_cond3(b.cond_eta, b.cond_k, dot3(wo, wh));
            r.f = divs(mul(mul(mul(b.spec_refl, ggx_d(wh, b.alpha)), fo), ggx_g(wi, wo, b.alpha)), 4.f * wi.z * wo.z);
        } break;
        case ROUGH_DIELECTRIC: {
            if (iszero(wo.z)) return;
            bool refl = wo.z * wi.z > 0.f;
            V3 wh = refl ? norm3(add(wo, wi)) : norm3(add(wo, mul(wi, wo.z > 0.f ? b.eta : 1.f / b.eta)));
            wh = mul(wh, wh.z > 0.f ? 1.f : -1.f);
            float F = fr_diel(b.eta, dot3(wo, wh)), G = ggx_g(wi, wo, b.alpha), D = ggx_d(wh, b.alpha);
            if (refl) {
                r.f = divs(mul(mul(mul(b.spec_refl, F), G), D), 4.f * fabsf(wi.z) * fabsf(wo.z));
            } else {
                float e = wo.z > 0.f ? b.eta : 1.f / b.eta;
                float sd = dot3(wo, wh) + e * dot3(wi, wh);
                r.f = mul(b.spec_trans, fabsf((1.f - F) * D * G * dot3(wi, wh) * dot3(wo, wh) / (sd * sd * wi.z * wo.z)));
            }
        } break;
        case PLASTIC: {
            if (wi.z <= 0.f || wo.z <= 0.f) return;
            float fo = fr_diel(b.eta, wo.z), fi = fr_diel(b.eta, wi.z);
            r.f = divs(mul(mul(mul(plastic_diff(b), 1.f - fi), 1.f - fo), cos_hemi_pdf(wi)), b.eta * b.eta * wi.z);
        } break;
        case ROUGH_PLASTIC: {
            if (wi.z <= 0.f || wo.z <= 0.f) return;
            float fo = fr_diel(b.eta, wo.z);
            V3 wh = norm3(add(wi, wo));
            r.f = divs(mul(mul(mul(b.spec_refl, fr_diel(b.eta, dot3(wh, wo))), ggx_d(wh, b.alpha)), ggx_g(wi, wo, b.alpha)),
                       4.f * wo.z * wi.z);
            float fi = fr_diel(b.eta, wi.z);
            r.f = add(r.f, divs(mul(mul(mul(plastic_diff(b), 1.f - fi), 1.f - fo), INV_PI_F), b.eta * b.eta));
        } break;
        default: break;
    }
}

void bsdf_pdf(const Bsdf &b, BRec &r) {  // GetPdf
    r.pdf = 0.f;
    V3 wi = r.wi, wo = r.wo;
    switch (b.type) {
        case DIFFUSE:
            if (wi.z > 0.f && wo.z > 0.f) r.pdf = cos_hemi_pdf(wi);
            break;
        case ROUGH_CONDUCTOR: {
            if (wi.z <= 0.f || wo.z <= 0.f) return;
            V3 wh = norm3(norm3(add(wi, wo)));
            r.pdf = ggx_pdf(wo, wh, b.alpha) / (4.f * dot3(wo, wh));
        } break;
        case ROUGH_DIELECTRIC: {
            bool refl = wo.z * wi.z > 0.f;
            V3 wh;
            float dwh;
            if (refl) {
                wh = norm3(add(wo, wi));
                dwh = 1.f / (4.f * dot3(wi, wh));
            } else {
                float e = wo.z > 0.f ? b.eta : 1.f / b.eta;
                wh = norm3(add(wo, mul(wi, e)));
                float sd = dot3(wo, wh) + e * dot3(wi, wh);
                dwh = (e * e * dot3(wi, wh)) / (sd * sd);
            }
            wh = mul(wh, wh.z > 0.f ? 1.f : -1.f);
            V3 wof = mul(wo, wo.z > 0.f ? 1.f : -1.f);
            float F = fr_diel(b.eta, dot3(wo, wh));
            r.pdf = fabsf(ggx_pdf(wof, wh, b.alpha) * (refl ? F : 1.f - F) * dwh);
        } break;
        case PLASTIC: {
            if (wi.z <= 0.f || wo.z <= 0.f) return;
            float sp = plastic_sp(b, fr_diel(b.eta, wo.z));
            r.pdf = cos_hemi_pdf(wi) * (1.f - sp);
        } break;
        case ROUGH_PLASTIC: {
            if (wi.z <= 0.f || wo.z <= 0.f) return;
            float sp = plastic_sp(b, fr_diel(b.eta, wo.z));
            float dp = 1.f - sp;
            V3 wh = norm3(add(wi, wo));
            r.pdf = sp * ggx_pdf(wo, wh, b.alpha) / (4.f * dot3(wi, wh));
            r.pdf += dp * cos_hemi_pdf(wi);
        } break;
        default: break;
    }
}

void bsdf_eval(const Bsdf &b, BRec &r) { bsdf_f(b, r); bsdf_pdf(b, r); }

void bsdf_sample(const Bsdf &b, BRec &r, Random &rng) {
    r.f = v3s(0.f);
    r.pdf = 0.f;
    r.lobe = 0;
    r.wi = v3s(0.f);
    V3 wo = r.wo;
    switch (b.type) {
        case DIFFUSE: {
            float x = rng.next(), y = rng.next();
            r.wi = cos_hemi(x, y);
            bsdf_pdf(b, r);
            bsdf_f(b, r);
            r.lobe = LOBE_DIFFUSE_R;
        } break;
        case CONDUCTOR:
            r.wi = reflect_local(wo);
            r.pdf = 1.f;
            r.f = divs(mul(b.spec_refl, fr_cond3(b.cond_eta, b.cond_k, wo.z)), fabsf(r.wi.z));
            r.lobe = LOBE_DELTA_R;
            break;
        case ROUGH_CONDUCTOR: {
            float x = rng.next(), y = rng.next();
            r.wi = reflect_about(wo, ggx_sample(wo, b.alpha, V2{x, y}));
            bsdf_pdf(b, r);
            bsdf_f(b, r);
            r.lobe = LOBE_DIFFUSE_R;  // rough_conductor.h:45
        } break;
        case DIELECTRIC: {
            float ct;
            float F = fr_diel(b.eta, wo.z, ct);
            if (rng.next() < F) {
                r.wi = reflect_local(wo);
                r.pdf = F;
                r.f = divs(mul(b.spec_refl, F), fabsf(r.wi.z));
                r.lobe = LOBE_DELTA_R;
            } else {
                r.wi = refract_local(wo, ct, b.eta);
                r.pdf = 1.f - F;
                float fac = ct < 0.f ? 1.f / b.eta : b.eta;
                r.f = divs(mul(mul(mul(b.spec_trans, 1.f - F), fac), fac), fabsf(r.wi.z));
                r.lobe = LOBE_DELTA_T;
            }
        } break;
        case ROUGH_DIELECTRIC: {
            float x = rng.next(), y = rng.next();
            V3 wof = mul(wo, wo.z > 0.f ? 1.f : -1.f);
            V3 wh = ggx_sample(wof, b.alpha, V2{x, y});
            float ct = 0.f;
            float F = fr_diel(b.eta, dot3(wo, wh), ct);
            if (rng.next() < F) {
                r.wi = reflect_about(wo, wh);
                r.lobe = LOBE_GLOSSY_R;
            } else {
                if (iszero(ct)) return;
                r.wi = refract_about(wo, wh, ct, b.eta);
                r.lobe = LOBE_GLOSSY_T;
                if (r.wi.z * wo.z >= 0.f) return;
            }
            bsdf_pdf(b, r);
            bsdf_f(b, r);
        } break;
        case PLASTIC: {
            if (wo.z <= 0.f) return;
            float fo = fr_diel(b.eta, wo.z);
            float x = rng.next(), y = rng.next();
            float sp = plastic_sp(b, fo);
            if (x < sp) {
                r.lobe = LOBE_DELTA_R;
                r.wi = reflect_local(wo);
                r.f = divs(mul(b.spec_refl, fo), r.wi.z);
                r.pdf = sp;
            } else {
                r.lobe = LOBE_DIFFUSE_R;
                r.wi = cos_hemi((x - sp) / (1.f - sp), y);
                float fi = fr_diel(b.eta, r.wi.z);
                r.f = divs(mul(mul(mul(plastic_diff(b), 1.f - fi), 1.f - fo), cos_hemi_pdf(r.wi)), b.eta * b.eta * r.wi.z);
                r.pdf = cos_hemi_pdf(r.wi) * (1.f - sp);
            }
        } break;
        case ROUGH_PLASTIC: {
            if (wo.z <= 0.f) return;
            float fo = fr_diel(b.eta, wo.z);
            float sp = plastic_sp(b, fo);
            float x = rng.next(), y = rng.next();
            if (y < sp) {
                y /= sp;
                r.wi = reflect_about(wo, ggx_sample(wo, b.alpha, V2{x, y}));
                r.lobe = LOBE_GLOSSY_R;
            } else {
                y = (y - sp) / (1.f - sp);
                r.wi = cos_hemi(x, y);
                r.lobe = LOBE_DIFFUSE_R;
            }
            bsdf_pdf(b, r);
            bsdf_f(b, r);
        } break;
        default: break;
    }
}

// ---------------------------------------------------------------------------
// Geometry: watertight ray/triangle test (Woop, Benthin, Wald 2013).
// Pinned by: brute-force agreement and shared-edge/vertex watertightness tests.
inline float cmp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
struct RayK { int kx, ky, kz; float sx, sy, sz; };
RayK ray_k(V3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dz = cmp(d, kz);
    if (dz < 0.f) std::swap(kx, ky);
    return RayK{kx, ky, kz, cmp(d, kx) / dz, cmp(d, ky) / dz, 1.f / dz};
}
bool tri_hit(V3 o, const RayK &k, V3 v0, V3 v1, V3 v2, float &t, float &b1, float &b2) {
    V3 A = sub(v0, o), B = sub(v1, o), C = sub(v2, o);
    float Az0 = cmp(A, k.kz), Bz0 = cmp(B, k.kz), Cz0 = cmp(C, k.kz);
    float Ax = cmp(A, k.kx) - k.sx * Az0, Ay = cmp(A, k.ky) - k.sy * Az0;
    float Bx = cmp(B, k.kx) - k.sx * Bz0, By = cmp(B, k.ky) - k.sy * Bz0;
    float Cx = cmp(C, k.kx) - k.sx * Cz0, Cy = cmp(C, k.ky) - k.sy * Cz0;
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        double cb = (double)Cx * (double)By, cy = (double)Cy * (double)Bx;
        U = (float)(cb - cy);
        double ac = (double)Ax * (double)Cy, ay = (double)Ay * (double)Cx;
        V = (float)(ac - ay);
        double ba = (double)Bx * (double)Ay, bb = (double)By * (double)Ax;
        W = (float)(ba - bb);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float T = U * (k.sz * Az0) + V * (k.sz * Bz0) + W * (k.sz * Cz0);
    t = T / det;
    b1 = V / det;
    b2 = W / det;
    return true;
}

// Ray / sphere (Haines et al., Ray Tracing Gems 2019, ch. 7): discriminant
// from the squared distance of the center to the ray line, roots c/q and q/a.
// Nearest root in [tmin, tmax].  Pinned by: float64 closed-form roots and
// brute-force agreement (tests).
bool sph_hit(V3 o, V3 d, V3 c, float r, float tmin, float tmax, float &t) {
    V3 f = sub(o, c);
    float a = dot3(d, d), b = dot3(f, d);
    V3 l = sub(f, mul(b / a, d));
    float r2 = r * r;
    float disc = a * (r2 - dot3(l, l));
    if (!(disc >= 0.f)) return false;
    float sq = sqrtf(disc);
    float q = b >= 0.f ? -(b + sq) : -(b - sq);
    if (q == 0.f) return false;
    float t0 = (dot3(f, f) - r2) / q, t1 = q / a;
    if (t0 > t1) std::swap(t0, t1);
    if (t0 >= tmin && t0 <= tmax) { t = t0; return true; }
    if (t1 >= tmin && t1 <= tmax) { t = t1; return true; }
    return false;
}

const uint32_t SPHERE_BIT = 1u << 30;

// Sphere placement as in SetSphereAreaEmitter (world/emitter.cpp:224-241):
// center = M (0,0,0), radius = |M (1,0,0) - center|.
void sphere_cr(const float *m, V3 &c, float &r) {
    c = v3(m[3], m[7], m[11]);
    V3 px = v3(m[0] + m[3], m[4] + m[7], m[8] + m[11]);
    r = len3(sub(c, px));
}

// Binary BVH over float boxes (median split on the widest centroid axis).
struct BNode { float lo[3], hi[3]; uint32_t left_or_first, count; };  // count > 0: leaf

struct Accel {
    std::vector<BNode> nodes;
    std::vector<uint32_t> prims;  // primitive ids (spheres carry SPHERE_BIT)
    std::vector<V3> p0, p1, p2;  // per triangle vertices
    std::vector<V3> sc;          // per sphere center
    std::vector<float> sr;       // per sphere radius
};

// Closest-hit candidate test shared by the BVH walk and the brute force.
inline bool prim_hit(const Accel &acc, uint32_t p, V3 o, V3 d, const RayK &k, float tmin, float tmax, float &t,
                     float &b1, float &b2) {
    b1 = b2 = 0.f;
    if (p & SPHERE_BIT) {
        const uint32_t i = p & (SPHERE_BIT - 1u);
        return sph_hit(o, d, acc.sc[i], acc.sr[i], tmin, tmax, t);
    }
    return tri_hit(o, k, acc.p0[p], acc.p1[p], acc.p2[p], t, b1, b2);
}

void build_accel(const or_scene &s, Accel &acc) {
    const uint32_t nt = s.num_triangles, ns = s.num_spheres, np = nt + ns;
    acc.p0.resize(nt); acc.p1.resize(nt); acc.p2.resize(nt);
    acc.sc.resize(ns); acc.sr.resize(ns);
    std::vector<float> cen(3 * (size_t)np), blo(3 * (size_t)np), bhi(3 * (size_t)np);
    for (uint32_t i = 0; i < ns; ++i) {
        sphere_cr(s.spheres + 12 * (size_t)i, acc.sc[i], acc.sr[i]);
        const float cc[3] = {acc.sc[i].x, acc.sc[i].y, acc.sc[i].z};
        const size_t j = nt + (size_t)i;
        for (int k = 0; k < 3; ++k) {
            blo[3 * j + k] = cc[k] - acc.sr[i];
            bhi[3 * j + k] = cc[k] + acc.sr[i];
            cen[3 * j + k] = cc[k];
        }
    }
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t *ix = s.indices + 3 * (size_t)t;
        const float *a = s.positions + 3 * (size_t)ix[0], *b = s.positions + 3 * (size_t)ix[1], *c = s.positions + 3 * (size_t)ix[2];
        acc.p0[t] = v3(a[0], a[1], a[2]);
        acc.p1[t] = v3(b[0], b[1], b[2]);
        acc.p2[t] = v3(c[0], c[1], c[2]);
        for (int k = 0; k < 3; ++k) {
            blo[3 * t + k] = std::min(a[k], std::min(b[k], c[k]));
            bhi[3 * t + k] = std::max(a[k], std::max(b[k], c[k]));
            cen[3 * t + k] = 0.5f * (blo[3 * t + k] + bhi[3 * t + k]);
        }
    }
    // Work on slots 0..np-1 (triangles, then spheres); prims[] maps them to ids.
    std::vector<uint32_t> slot(np);
    for (uint32_t i = 0; i < np; ++i) slot[i] = i;
    acc.nodes.clear();
    acc.nodes.reserve(2 * (size_t)np + 1);
    struct Item { uint32_t node, first, count; };
    std::vector<Item> work;
    acc.nodes.push_back(BNode{});
    work.push_back(Item{0, 0, np});
    while (!work.empty()) {
        Item it = work.back();
        work.pop_back();
        BNode n;
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; ++k) { n.lo[k] = INFINITY; n.hi[k] = -INFINITY; }
        for (uint32_t i = it.first; i < it.first + it.count; ++i) {
            uint32_t p = slot[i];
            for (int k = 0; k < 3; ++k) {
                n.lo[k] = std::min(n.lo[k], blo[3 * p + k]);
                n.hi[k] = std::max(n.hi[k], bhi[3 * p + k]);
                clo[k] = std::min(clo[k], cen[3 * p + k]);
                chi[k] = std::max(chi[k], cen[3 * p + k]);
            }
        }
        if (it.count <= 4) {
            n.left_or_first = it.first;
            n.count = it.count;
            acc.nodes[it.node] = n;
            continue;
        }
        int axis = 0;
        for (int k = 1; k < 3; ++k) if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
        uint32_t mid = it.first + it.count / 2;
        std::nth_element(slot.begin() + it.first, slot.begin() + mid, slot.begin() + it.first + it.count,
                         [&](uint32_t a, uint32_t b) { return cen[3 * a + axis] < cen[3 * b + axis]; });
        uint32_t l = (uint32_t)acc.nodes.size();
        acc.nodes.push_back(BNode{});
        acc.nodes.push_back(BNode{});
        n.left_or_first = l;
        n.count = 0;
        acc.nodes[it.node] = n;
        work.push_back(Item{l, it.first, mid - it.first});
        work.push_back(Item{l + 1, mid, it.first + it.count - mid});
    }
    acc.prims.resize(np);
    for (uint32_t i = 0; i < np; ++i) acc.prims[i] = slot[i] < nt ? slot[i] : (SPHERE_BIT | (slot[i] - nt));
}

struct HitRec { float t; uint32_t prim; float b1, b2; };

// Closest hit = lexicographic minimum of (t, prim) over triangles with
// t in [tmin, tmax]; node intervals are widened by a relative 1e-4.
bool accel_trace(const Accel &acc, V3 o, V3 d, float tmin, float tmax, bool any, HitRec &h, uint64_t *nodes_v,
                 uint64_t *tris_v) {
    h.t = tmax;
    h.prim = 0xFFFFFFFFu;
    h.b1 = h.b2 = 0.f;
    if (acc.nodes.empty() || acc.prims.empty()) return false;
    RayK k = ray_k(d);
    float inv[3];
    const float dd[3] = {d.x, d.y, d.z}, oo[3] = {o.x, o.y, o.z};
    for (int a = 0; a < 3; ++a) inv[a] = 1.f / (fabsf(dd[a]) > 1e-12f ? dd[a] : copysignf(1e-12f, dd[a]));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    float tbest = tmax;
    while (sp) {
        const BNode &n = acc.nodes[stack[--sp]];
        if (nodes_v) (*nodes_v)++;
        float tn = tmin, tf = tbest;
        float t0[3], t1[3];
        for (int a = 0; a < 3; ++a) {
            t0[a] = (n.lo[a] - oo[a]) * inv[a];
            t1[a] = (n.hi[a] - oo[a]) * inv[a];
        }
        tn = fmaxf(fmaxf(fminf(t0[0], t1[0]), fminf(t0[1], t1[1])), fmaxf(fminf(t0[2], t1[2]), tmin));
        tf = fminf(fminf(fmaxf(t0[0], t1[0]), fmaxf(t0[1], t1[1])), fminf(fmaxf(t0[2], t1[2]), tbest));
        if (!(tn * (1.f - 1e-4f) <= tf * (1.f + 1e-4f))) continue;
        if (n.count == 0) {
            stack[sp++] = n.left_or_first;
            stack[sp++] = n.left_or_first + 1;
            continue;
        }
        for (uint32_t i = n.left_or_first; i < n.left_or_first + n.count; ++i) {
            uint32_t p = acc.prims[i];
            if (tris_v) (*tris_v)++;
            float t, b1, b2;
            if (!prim_hit(acc, p, o, d, k, tmin, tmax, t, b1, b2)) continue;
            if (t >= tmin && (t < tbest || (t == tbest && p < h.prim))) {
                h.t = t; h.prim = p; h.b1 = b1; h.b2 = b2;
                tbest = t;
                if (any) return true;
            }
        }
    }
    return h.prim != 0xFFFFFFFFu;
}

bool brute_trace(const Accel &acc, V3 o, V3 d, float tmin, float tmax, bool any, HitRec &h) {
    h.t = tmax;
    h.prim = 0xFFFFFFFFu;
    h.b1 = h.b2 = 0.f;
    RayK k = ray_k(d);
    const uint32_t nt = (uint32_t)acc.p0.size(), np = nt + (uint32_t)acc.sc.size();
    for (uint32_t i = 0; i < np; ++i) {
        const uint32_t p = i < nt ? i : (SPHERE_BIT | (i - nt));
        float t, b1, b2;
        if (!prim_hit(acc, p, o, d, k, tmin, tmax, t, b1, b2)) continue;
        if (t >= tmin && (t < h.t || (t == h.t && p < h.prim))) {
            h.t = t; h.prim = p; h.b1 = b1; h.b2 = b2;
            if (any) return true;
        }
    }
    return h.prim != 0xFFFFFFFFu;
}

// ---------------------------------------------------------------------------
// World: per-triangle area emitters and selection probabilities
// (world/emitter.cpp:169-222, 321-337), camera, environment.
struct AreaEmitter {
    bool sphere;     // SphereEmitter (sphere.h): center p[0], radius rad
    V3 p[3], n[3];
    V2 uv[3];
    float rad;
    float area, weight, sel;
    or_texture radiance;
};

struct World {
    const or_scene *s;
    Accel acc;
    std::vector<int32_t> tri_shape;
    std::vector<int32_t> sph_shape;
    std::vector<std::array<float, 12>> sph_w2o;  // world -> object rows
    std::vector<int32_t> shape_emitter_offset;
    std::vector<AreaEmitter> emitters;
    std::vector<float> cdf;  // running sums of selection probabilities
    bool has_env;
    float env_sel;
    V3 env_color;
    // EnvMapEmitter (env.h:6-65) when s->env_type == 2
    bool env_map;
    const or_bitmap *env_bm;
    float env_scale, env_norm;
    uint32_t env_w, env_h;
    std::vector<float> env_row_cdf, env_col_cdf, env_row_w;
    float env_to_world[9], env_to_local[9];
};

void build_env_tables(const or_scene &s, World &w);

float max3f(float r, float g, float b) { return r > g ? (r > b ? r : b) : (g > b ? g : b); }
float tex_weight(const or_texture &t, const or_bitmap *bm) {  // GetWeight (world/emitter.cpp:73-101)
    if (t.type == 1) return (max3f(t.color0[0], t.color0[1], t.color0[2]) + max3f(t.color1[0], t.color1[1], t.color1[2])) * 0.5f;
    if (t.type == 2) {
        const or_bitmap &b = bm[t.bitmap];
        float acc = 0.f;
        for (size_t i = 0; i < (size_t)b.width * b.height; ++i) acc += max3f(b.rgba[4 * i], b.rgba[4 * i + 1], b.rgba[4 * i + 2]);
        return acc / (1.f * (float)b.width * (float)b.height);
    }
    return max3f(t.color0[0], t.color0[1], t.color0[2]);
}

void build_world(const or_scene &s, World &w) {
    w.s = &s;
    build_accel(s, w.acc);
    w.tri_shape.assign(s.num_triangles, 0);
    w.sph_shape.assign(s.num_spheres, 0);
    w.sph_w2o.resize(s.num_spheres);
    for (uint32_t k = 0; k < s.num_spheres; ++k) {
        // Inverse of the affine [A | t] in double precision: [A^-1 | -A^-1 t].
        const float *m = s.spheres + 12 * (size_t)k;
        double A[3][3], inv[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = m[4 * i + j];
        double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                     A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                // cofactor of A[j][i] / det
                int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
                inv[i][j] = (A[r0][c0] * A[r1][c1] - A[r0][c1] * A[r1][c0]) / det;
            }
        for (int i = 0; i < 3; ++i) {
            double tt = -(inv[i][0] * m[3] + inv[i][1] * m[7] + inv[i][2] * m[11]);
            for (int j = 0; j < 3; ++j) w.sph_w2o[k][4 * i + j] = (float)inv[i][j];
            w.sph_w2o[k][4 * i + 3] = (float)tt;
        }
    }
    w.shape_emitter_offset.assign(s.num_shapes, -1);
    w.emitters.clear();
    for (uint32_t si = 0; si < s.num_shapes; ++si) {
        const or_shape &sh = s.shapes[si];
        if (sh.is_emitter) w.shape_emitter_offset[si] = (int32_t)w.emitters.size();
        if (sh.flags & 32u) {
            for (uint32_t k = sh.first_tri; k < sh.first_tri + sh.num_tris; ++k) {
                w.sph_shape[k] = (int32_t)si;
                if (!sh.is_emitter) continue;
                AreaEmitter e{};
                e.sphere = true;
                e.p[0] = w.acc.sc[k];
                e.rad = w.acc.sr[k];
                e.area = 4 * 3.14159265358979323846f * e.rad * e.rad;  // emitter.cpp:238
                e.radiance = sh.radiance;
                e.weight = tex_weight(sh.radiance, s.bitmaps) * e.area;
                w.emitters.push_back(e);
            }
            continue;
        }
        for (uint32_t t = sh.first_tri; t < sh.first_tri + sh.num_tris; ++t) {
            w.tri_shape[t] = (int32_t)si;
            if (!sh.is_emitter) continue;
            AreaEmitter e{};
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = s.indices[3 * (size_t)t + k];
                e.p[k] = v3(s.positions[3 * vi], s.positions[3 * vi + 1], s.positions[3 * vi + 2]);
                e.uv[k] = s.texcoords ? V2{s.texcoords[2 * vi], s.texcoords[2 * vi + 1]} : V2{0.f, 0.f};
            }
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = s.indices[3 * (size_t)t + k];
                V3 nn;
                if (s.normals && (sh.flags & 4u)) {
                    nn = v3(s.normals[3 * vi], s.normals[3 * vi + 1], s.normals[3 * vi + 2]);
                } else {
                    nn = cross3(sub(e.p[1], e.p[0]), sub(e.p[2], e.p[0]));
                    if (sh.flags & 16u) nn = neg(nn);
                }
                float l = sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
                e.n[k] = v3(nn.x / l, nn.y / l, nn.z / l);  // util/transform.cpp:115-123
            }
            e.area = len3(cross3(sub(e.p[1], e.p[0]), sub(e.p[2], e.p[0]))) * 0.5f;
            e.radiance = sh.radiance;
            e.weight = tex_weight(sh.radiance, s.bitmaps) * e.area;
            w.emitters.push_back(e);
        }
    }
    w.has_env = s.env_type != 0;
    w.env_color = v3(s.env_radiance[0], s.env_radiance[1], s.env_radiance[2]);
    w.env_map = s.env_type == 2;
    if (w.env_map) build_env_tables(s, w);
    float sum = 0.f;
    for (auto &e : w.emitters) sum += e.weight;
    size_t na = w.emitters.size(), total = na + (w.has_env ? 1 : 0);
    for (auto &e : w.emitters) {
        e.sel = e.weight / sum * (float)na;
        e.sel = e.sel / (float)total;
    }
    w.env_sel = w.has_env ? 1.f / (float)total : 0.f;
    w.cdf.resize(na);
    float run = 0.f;
    for (size_t i = 0; i < na; ++i) { run = run + w.emitters[i].sel; w.cdf[i] = run; }
}

// BuildEnvMapCdfTable (world/emitter.cpp:107-149) and the frame of
// world/emitter.cpp:293-312.  Float sums in the reference order; a black row
// keeps its 0/0 column entries; an all-black map gets normalization 0 (the
// reference divides by zero).  Pinned by: uniform-map and single-texel
// closed forms of the rendered irradiance (tests).
void build_env_tables(const or_scene &s, World &w) {
    const or_bitmap &b = s.bitmaps[s.env_bitmap];
    w.env_bm = &b;
    w.env_scale = s.env_scale;
    const uint32_t W = b.width, H = b.height;
    w.env_w = W;
    w.env_h = H;
    w.env_row_cdf.assign(H + 1, 0.f);
    w.env_col_cdf.assign((size_t)(W + 1) * H, 0.f);
    w.env_row_w.assign(H, 0.f);
    float total = 0.f;
    for (uint32_t y = 0; y < H; ++y) {
        float *cc = &w.env_col_cdf[(size_t)y * (W + 1)];
        float run = 0.f;
        for (uint32_t x = 0; x < W; ++x) {
            const float *p = b.rgba + 4 * ((size_t)y * W + x);
            run += lum(v3(p[0], p[1], p[2]));
            cc[x + 1] = run;
        }
        for (uint32_t x = 1; x < W; ++x) cc[x] /= run;
        cc[W] = 1.f;
        w.env_row_w[y] = std::sin(((float)y + 0.5f) * PI_F / (float)H);
        total += run * w.env_row_w[y];
        w.env_row_cdf[y + 1] = total;
    }
    for (uint32_t y = 1; y < H; ++y) w.env_row_cdf[y] /= total;
    w.env_row_cdf[H] = 1.f;
    w.env_norm = total == 0.f ? 0.f : 1.f / (total * (2.f * PI_F / (float)W) * (PI_F / (float)H));
    const float *m = s.env_to_world;
    double det = (double)m[0] * ((double)m[4] * m[8] - (double)m[5] * m[7]) - (double)m[1] * ((double)m[3] * m[8] - (double)m[5] * m[6]) +
                 (double)m[2] * ((double)m[3] * m[7] - (double)m[4] * m[6]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
            w.env_to_local[3 * i + j] = (float)(((double)m[3 * r0 + c0] * m[3 * r1 + c1] - (double)m[3 * r0 + c1] * m[3 * r1 + c0]) / det);
            w.env_to_world[3 * i + j] = m[3 * i + j];
        }
}

V3 rows3(const float *r, V3 v) {
    return v3(r[0] * v.x + r[1] * v.y + r[2] * v.z, r[3] * v.x + r[4] * v.y + r[5] * v.z, r[6] * v.x + r[7] * v.y + r[8] * v.z);
}
// Local direction of map angles (env.h:36): (sin t sin(pi - p), cos t, sin t cos(pi - p)).
V3 env_dir(float phi, float theta) {
    float st, ct, sp, cp;
    sincos_or(theta, st, ct);
    sincos_or(PI_F - phi, sp, cp);
    return v3(st * sp, ct, st * cp);
}
V3 env_lookup(const World &w, V2 tex) { return mul(bitmap_lookup(*w.env_bm, tex.x, tex.y), w.env_scale); }
V2 env_uv(const World &w, V3 dir, float &theta) {  // env.h:52-56
    V3 d = rows3(w.env_to_local, dir);
    float phi = PI_F - atan2_or(d.x, d.z);
    theta = acos_or(d.y);
    return V2{phi * 0.5f * INV_PI_F, theta * INV_PI_F};
}
// Density of the continuous sampler: texel share over its (phi, theta) cell.
float env_pdf_cell(const World &w, V2 tex, float theta) {
    uint32_t x = (uint32_t)fminf(fmaxf(tex.x * (float)w.env_w, 0.f), (float)(w.env_w - 1));
    uint32_t y = (uint32_t)fminf(fmaxf(tex.y * (float)w.env_h, 0.f), (float)(w.env_h - 1));
    const float *p = w.env_bm->rgba + 4 * ((size_t)y * w.env_bm->width + x);
    float st, ct;
    sincos_or(theta, st, ct);
    if (!(st > 0.f)) return 0.f;
    return lum(v3(p[0], p[1], p[2])) * w.env_row_w[y] * w.env_norm / st;
}
// Eval (env.h:51-63) or, with exact, the continuous sampler's density.
void env_eval(const World &w, V3 dir, bool exact, V3 &rad, float &pdf) {
    float theta;
    V2 tex = env_uv(w, dir, theta);
    rad = env_lookup(w, tex);
    if (exact) { pdf = env_pdf_cell(w, tex, theta); return; }
    uint32_t row = (uint32_t)(tex.y * (float)w.env_h);
    row = std::min(row, w.env_h - 2);
    float t = tex.y * (float)w.env_h - 1.f * (float)row;
    float wr = w.env_row_w[row] + t * (w.env_row_w[row + 1] - w.env_row_w[row]);
    float st, ct;
    sincos_or(theta, st, ct);
    pdf = lum(rad) * wr * w.env_norm / fmaxf(1e-4f, fabsf(st));
}
// SampleDirect (env.h:23-49): scans for the first entry not below the draw;
// direction = that entry's texel corner.  The row is clamped to the last one
// when the draw exceeds every row entry (the reference reads past its table).
void env_sample_ref(const World &w, V2 xi, V3 &wi, V3 &rad, float &pdf) {
    uint32_t row = 0;
    while (row < w.env_h && !(xi.x <= w.env_row_cdf[row])) ++row;
    row = std::min(row, w.env_h - 1);
    const float *cc = &w.env_col_cdf[(size_t)row * (w.env_w + 1)];
    uint32_t col = 0;
    while (col < w.env_w - 1 && !(xi.y <= cc[col])) ++col;
    float phi = (float)col * PI_F * 2.f / (float)w.env_w, theta = (float)row * PI_F / (float)w.env_h;
    wi = rows3(w.env_to_world, env_dir(phi, theta));
    rad = env_lookup(w, V2{phi * 0.5f * INV_PI_F, theta * INV_PI_F});
    float st, ct;
    sincos_or(theta, st, ct);
    pdf = lum(rad) * w.env_row_w[row] * w.env_norm / fmaxf(1e-4f, fabsf(st));
    if (pdf < 0.f) pdf = 0.f;
}
// Continuous sampling of the same tables: the cdf interval of the row and of
// the column, the in-texel position from the rest of each draw.
void env_sample_exact(const World &w, V2 xi, V3 &wi, V3 &rad, float &pdf) {
    uint32_t y = 0;
    while (y + 1 < w.env_h && !(xi.x <= w.env_row_cdf[y + 1])) ++y;
    float r0 = w.env_row_cdf[y], r1 = w.env_row_cdf[y + 1];
    float fy = r1 > r0 ? fminf(fmaxf((xi.x - r0) / (r1 - r0), 0.f), 1.f) : 0.5f;
    const float *cc = &w.env_col_cdf[(size_t)y * (w.env_w + 1)];
    uint32_t x = 0;
    while (x + 1 < w.env_w && !(xi.y <= cc[x + 1])) ++x;
    float c0 = cc[x], c1 = cc[x + 1];
    float fx = c1 > c0 ? fminf(fmaxf((xi.y - c0) / (c1 - c0), 0.f), 1.f) : 0.5f;
    float phi = ((float)x + fx) * PI_F * 2.f / (float)w.env_w, theta = ((float)y + fy) * PI_F / (float)w.env_h;
    wi = rows3(w.env_to_world, env_dir(phi, theta));
    env_eval(w, wi, true, rad, pdf);
}

// SelectOneEmiiter (render/emitter.h:110-136): linear scan, `p <= sum`.
int select_emitter(const World &w, float p) {
    float sum_p = 0.f;
    for (size_t i = 0; i < w.emitters.size(); ++i) {
        if (p <= sum_p + w.emitters[i].sel) return (int)i;
        sum_p += w.emitters[i].sel;
    }
    if (w.has_env) return -1;
    if (!w.emitters.empty()) return (int)w.emitters.size() - 1;
    return -2;
}

struct Geo { V3 pos, n; V2 uv; };

// Geometry::GetHitLocalGeometry (render/geometry.h:60-100, 176-180).
void hit_geo(const World &w, uint32_t prim, float t_hit, float b1, float b2, V3 ray_o, V3 ray_dir, Geo &g, int &mat,
             int &emitter) {
    const or_scene &s = *w.s;
    if (prim & SPHERE_BIT) {
        // Sphere: position o + t d; object-space point -> texcoord; normal by
        // the inverse transpose of object -> world (transposed w2o).
        const uint32_t k = prim & (SPHERE_BIT - 1u);
        const int si = w.sph_shape[k];
        const or_shape &sh = s.shapes[si];
        const float *M = w.sph_w2o[k].data();
        g.pos = add(ray_o, mul(t_hit, ray_dir));
        V4 hp{g.pos.x, g.pos.y, g.pos.z, 1.f};
        V3 lp = v3(dot4(V4{M[0], M[1], M[2], M[3]}, hp), dot4(V4{M[4], M[5], M[6], M[7]}, hp),
                   dot4(V4{M[8], M[9], M[10], M[11]}, hp));
        g.uv = sphere_uv(norm3(lp));
        V3 n = v3(M[0] * lp.x + M[4] * lp.y + M[8] * lp.z, M[1] * lp.x + M[5] * lp.y + M[9] * lp.z,
                  M[2] * lp.x + M[6] * lp.y + M[10] * lp.z);
        n = norm3(n);
        if (sh.flags & 1u) n = mul(n, -1.f);
        mat = (int)sh.material;
        if (dot3(neg(ray_dir), n) < 0.f && s.materials[sh.material].twosided) n = neg(n);
        g.n = n;
        emitter = w.shape_emitter_offset[si] >= 0 ? w.shape_emitter_offset[si] + (int)(k - sh.first_tri) : -1;
        return;
    }
    const uint32_t *ix = s.indices + 3 * (size_t)prim;
    const int si = w.tri_shape[prim];
    const or_shape &sh = s.shapes[si];
    V3 p0 = w.acc.p0[prim], p1 = w.acc.p1[prim], p2 = w.acc.p2[prim];
    float wgt = 1.f - b1 - b2;
    g.pos = add(add(mul(wgt, p0), mul(b1, p1)), mul(b2, p2));
    V3 n;
    if (s.normals && (sh.flags & 4u)) {
        const float *n0 = s.normals + 3 * (size_t)ix[0], *n1 = s.normals + 3 * (size_t)ix[1], *n2 = s.normals + 3 * (size_t)ix[2];
        n = add(add(mul(wgt, v3(n0[0], n0[1], n0[2])), mul(b1, v3(n1[0], n1[1], n1[2]))), mul(b2, v3(n2[0], n2[1], n2[2])));
    } else {
        n = cross3(sub(p1, p0), sub(p2, p0));
        if (sh.flags & 16u) n = neg(n);
    }
    n = norm3(n);
    if (sh.flags & 1u) n = mul(n, -1.f);
    V2 uv{0.f, 0.f};
    if (s.texcoords && (sh.flags & 8u)) {
        const float *t0 = s.texcoords + 2 * (size_t)ix[0], *t1 = s.texcoords + 2 * (size_t)ix[1], *t2 = s.texcoords + 2 * (size_t)ix[2];
        uv = add2(add2(mul2(V2{t0[0], t0[1]}, wgt), mul2(V2{t1[0], t1[1]}, b1)), mul2(V2{t2[0], t2[1]}, b2));
        if (sh.flags & 2u) uv.y = 1.f - uv.y;
    }
    mat = (int)sh.material;
    if (dot3(neg(ray_dir), n) < 0.f && s.materials[sh.material].twosided) n = neg(n);
    g.n = n;
    g.uv = uv;
    emitter = w.shape_emitter_offset[si] >= 0 ? w.shape_emitter_offset[si] + (int)(prim - sh.first_tri) : -1;
}

struct ESample { V3 rad, wi; float dist, pdf; };

void area_sample(const AreaEmitter &e, const Geo &hit, V2 xi, ESample &r, const or_bitmap *bm) {  // area.h:17-35
    if (e.sphere) {  // sphere.h:14-33
        V3 t = uni_sphere(xi.x, xi.y);
        V3 pos = add(mul(t, e.rad), e.p[0]);
        V3 nrm = norm3(t);
        r.rad = tex_sample(e.radiance, sphere_uv(t), bm);
        r.wi = norm3(sub(pos, hit.pos));
        float nol = dot3(hit.n, r.wi), lnol = dot3(nrm, neg(r.wi));
        r.pdf = 0.f;
        r.dist = 0.f;
        if (nol > 0.f && lnol > 0.f) {
            float dist = len3(sub(pos, hit.pos));
            r.pdf = dist * dist / (lnol * e.area);
            r.dist = dist;
        }
        return;
    }
    V3 t = sample_tri(xi.x, xi.y);
    V3 pos = add(add(mul(e.p[0], t.x), mul(e.p[1], t.y)), mul(e.p[2], t.z));
    V3 nrm = norm3(add(add(mul(e.n[0], t.x), mul(e.n[1], t.y)), mul(e.n[2], t.z)));
    V2 tex = add2(add2(mul2(e.uv[0], t.x), mul2(e.uv[1], t.y)), mul2(e.uv[2], t.z));
    r.rad = tex_sample(e.radiance, tex, bm);
    r.wi = norm3(sub(pos, hit.pos));
    float nol = dot3(hit.n, r.wi), lnol = dot3(nrm, neg(r.wi));
    r.pdf = 0.f;
    r.dist = 0.f;
    if (nol > 0.f && lnol > 0.f) {
        float dist = len3(sub(pos, hit.pos));
        r.pdf = dist * dist / (lnol * e.area);
        r.dist = dist;
    }
}

// Camera ray (main.cu:58-78).
void camera_ray(const or_scene &s, uint32_t x, uint32_t y, float jx, float jy, V3 &o, V3 &d) {
    const float *S = s.sample_to_camera, *C = s.camera_to_world;
    V4 pf{((float)x + jx) / (float)s.width, ((float)y + jy) / (float)s.height, 0.f, 1.f};
    V4 v{dot4(V4{S[0], S[1], S[2], S[3]}, pf), dot4(V4{S[4], S[5], S[6], S[7]}, pf), dot4(V4{S[8], S[9], S[10], S[11]}, pf),
         dot4(V4{S[12], S[13], S[14], S[15]}, pf)};
    float iw = 1.0f / v.w;
    v = V4{v.x * iw, v.y * iw, v.z * iw, v.w * iw};
    v.w = 0.f;
    float il = 1.0f / sqrtf(dot4(v, v));
    v = V4{v.x * il, v.y * il, v.z * il, v.w * il};
    V3 wd = v3(dot4(V4{C[0], C[1], C[2], C[3]}, v), dot4(V4{C[4], C[5], C[6], C[7]}, v), dot4(V4{C[8], C[9], C[10], C[11]}, v));
    d = norm3(wd);
    o = v3(C[3], C[7], C[11]);
}

struct PixelOut {
    V3 L, albedo, normal;
    float test;
    uint32_t prim;
};

// One camera sample of one pixel: main.cu:38-197.
void trace_path(const World &w, const or_params &p, uint32_t max_depth, uint32_t px, uint32_t py, uint32_t frame,
                PixelOut &out, uint64_t *counts) {
    const or_scene &s = *w.s;
    const uint32_t pixel = py * s.width + px;
    const bool consistent = p.mis_mode != 0;
    Random rng;
    rng.init(pixel, frame);
    V3 radiance = v3s(0.f), env_radiance = v3s(0.f), throughput = v3s(1.f);
    float env_pdf = 0.f;
    float jx = rng.next(), jy = rng.next();
    V3 ray_o, ray_d;
    camera_ray(s, px, py, jx, jy, ray_o, ray_d);

    HitRec h;
    counts[0]++;
    bool hit = accel_trace(w.acc, ray_o, ray_d, 0.001f, 1e16f, false, h, counts + 2, counts + 3);
    bool done = !hit;
    Geo g;
    int mat = 0, emitter = -1;
    Bsdf bsdf;
    std::memset(&bsdf, 0, sizeof(bsdf));
    out.prim = h.prim;
    if (done) {
        if (w.env_map) {
            if (consistent) env_eval(w, ray_d, true, env_radiance, env_pdf);
            else env_eval(w, norm3(sub(add(ray_o, ray_d), ray_o)), false, env_radiance, env_pdf);
        } else if (w.has_env) {
            env_radiance = w.env_color;
            env_pdf = 0.25f * INV_PI_F;
        }
        out.albedo = v3s(0.f);
        out.normal = v3s(0.f);
    } else {
        hit_geo(w, h.prim, h.t, h.b1, h.b2, ray_o, ray_d, g, mat, emitter);
        bsdf = make_bsdf(s.materials[mat], g.uv, s.bitmaps);
        if (emitter >= 0) radiance = add(radiance, tex_sample(w.emitters[emitter].radiance, g.uv, s.bitmaps));
        out.albedo = albedo_of(bsdf);
        out.normal = g.n;
    }
    out.test = rng.next();

    uint32_t depth = 0;
    bool prev_upper = false;
    while (!done) {
        ++depth;
        if (depth >= max_depth) break;
        float rr = depth > p.rr_depth ? 0.95f : 1.0f;
        if (rng.next() > rr) break;
        throughput = divs(throughput, rr);
        // direct light sampling
        {
            float ps = rng.next();
            float x = rng.next(), y = rng.next();
            int sel = select_emitter(w, ps);
            ESample es;
            es.pdf = 0.f;
            es.dist = 0.f;
            float sel_prob = 0.f;
            if (sel >= 0) {
                area_sample(w.emitters[sel], g, V2{x, y}, es, s.bitmaps);
                sel_prob = w.emitters[sel].sel;
            } else if (sel == -1 && w.env_map) {  // EnvMapEmitter::SampleDirect
                if (consistent) env_sample_exact(w, V2{x, y}, es.wi, es.rad, es.pdf);
                else env_sample_ref(w, V2{x, y}, es.wi, es.rad, es.pdf);
                es.dist = 1e16f;
                sel_prob = w.env_sel;
            } else if (sel == -1) {  // ConstEnvEmitter::SampleDirect (env.h:70-80)
                V3 lw = uni_hemi(x, y);
                es.wi = to_world(lw, g.n);
                es.pdf = uni_hemi_pdf(lw);
                es.dist = 1e16f;
                es.rad = w.env_color;
                sel_prob = w.env_sel;
            }
            if (es.pdf != 0.f) {
                HitRec sh;
                counts[1]++;
                bool occluded = accel_trace(w.acc, g.pos, es.wi, 0.0001f, es.dist - 0.0001f, true, sh, counts + 2, counts + 3);
                if (!occluded) {
                    BRec er;
                    er.wi = to_local(es.wi, g.n);
                    er.wo = to_local(neg(ray_d), g.n);
                    bsdf_eval(bsdf, er);
                    if (!iszero3(mul(er.f, es.pdf))) {
                        float nol = dot3(g.n, es.wi);
                        if (nol > 0.f) {
                            float mis;
                            if (consistent) {
                                es.pdf *= sel_prob;
                                mis = misw(es.pdf, er.pdf);
                            } else {
                                mis = misw(es.pdf, er.pdf);
                                es.pdf *= sel_prob;
                            }
                            radiance = add(radiance, divs(mul(mul(mul(mul(throughput, es.rad), er.f), nol), mis), es.pdf));
                        }
                    }
                }
            }
        }
        // bsdf sampling
        BRec br;
        br.wo = to_local(neg(ray_d), g.n);
        bsdf_sample(bsdf, br, rng);
        // The sample's value and pdf must be finite (a GGX reflection exactly in
        // the tangent plane evaluates 0 / 0 in rough_dielectric.h:39-48); the
        // reference would carry the NaN into the film, the path ends here.
        if (iszero3(mul(br.f, fabsf(br.wi.z))) || iszero(br.pdf) || !std::isfinite(br.f.x + br.f.y + br.f.z) ||
            !std::isfinite(br.pdf))
            break;
        throughput = mul(throughput, divs(mul(br.f, fabsf(br.wi.z)), br.pdf));
        ray_o = g.pos;
        ray_d = to_world(br.wi, g.n);
        prev_upper = dot3(g.n, ray_d) > 0.f;
        const bool prev_delta = (br.lobe & (LOBE_DELTA_R | LOBE_DELTA_T)) != 0;
        counts[0]++;
        bool h2 = accel_trace(w.acc, ray_o, ray_d, 0.001f, 1e16f, false, h, counts + 2, counts + 3);
        if (!h2) {
            float lp = 0.f;  // selection x next-event density of this direction
            if (w.env_map) {
                if (consistent) {
                    float p;
                    env_eval(w, ray_d, true, env_radiance, p);
                    lp = w.env_sel * p;
                } else {  // __miss__default evaluates the direction (o + d) - o
                    env_eval(w, norm3(sub(add(ray_o, ray_d), ray_o)), false, env_radiance, env_pdf);
                }
            } else if (w.has_env) {
                env_radiance = w.env_color;
                env_pdf = 0.25f * INV_PI_F;
                lp = prev_upper ? w.env_sel * (INV_PI_F * 0.5f) : 0.f;
            }
            float mis;
            if (consistent) {
                mis = prev_delta ? 1.f : misw(br.pdf, lp);
            } else {
                mis = misw(br.pdf, env_pdf);
            }
            env_radiance = mul(env_radiance, mul(throughput, mis));
            break;
        }
        hit_geo(w, h.prim, h.t, h.b1, h.b2, ray_o, ray_d, g, mat, emitter);
        bsdf = make_bsdf(s.materials[mat], g.uv, s.bitmaps);
        if (emitter >= 0) {
            const AreaEmitter &e = w.emitters[emitter];
            // TriAreaEmitter::Eval (area.h:37-45)
            V3 dir = norm3(sub(ray_o, g.pos));
            float lnol = dot3(g.n, dir);
            float epdf = 0.f;
            V3 erad = v3s(0.f);
            if (lnol > 0.f) {
                float dist = len3(sub(ray_o, g.pos));
                epdf = dist * dist / (lnol * e.area);
                erad = tex_sample(e.radiance, g.uv, s.bitmaps);
            }
            if (!iszero(epdf)) {
                float lpdf = epdf;
                if (consistent) {
                    // The pdf next-event estimation would assign to this point:
                    // it uses the emitter's own normals (area.h:19-33,
                    // sphere.h:14-33).
                    float wv = 1.f - h.b1 - h.b2;
                    V3 en = e.sphere ? norm3(sub(g.pos, e.p[0]))
                                     : norm3(add(add(mul(e.n[0], wv), mul(e.n[1], h.b1)), mul(e.n[2], h.b2)));
                    float nl = dot3(en, dir);
                    float dist = len3(sub(ray_o, g.pos));
                    lpdf = nl > 0.f ? dist * dist / (nl * e.area) : 0.f;
                }
                float mis = prev_delta ? 1.f : misw(br.pdf, lpdf * e.sel);
                radiance = add(radiance, mul(mul(throughput, erad), mis));
            }
        }
    }
    radiance = add(radiance, env_radiance);
    out.L = radiance;
}

}  // namespace

extern "C" {

int oracle_render(const or_scene *scene, const or_params *params, float *accum, float *albedo, float *normal,
                  float *test, uint32_t *prim, double *sum64, or_counts *counts) {
    if (!scene || !params) return -1;
    const or_scene &s = *scene;
    const auto t_build = std::chrono::steady_clock::now();
    World w;
    build_world(s, w);
    const auto t_render = std::chrono::steady_clock::now();
    const uint32_t max_depth = params->max_depth ? params->max_depth : (s.max_depth ? s.max_depth : 1);
    const uint32_t y0 = params->row_begin, y1 = params->row_end ? std::min(params->row_end, s.height) : s.height;
    unsigned nthreads = params->threads ? params->threads : std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next_row{y0};
    std::vector<std::thread> pool;
    std::vector<uint64_t> cnt(4 * (size_t)nthreads, 0);
    for (unsigned ti = 0; ti < nthreads; ++ti) {
        pool.emplace_back([&, ti] {
            uint64_t *c = &cnt[4 * (size_t)ti];
            for (;;) {
                uint32_t y = next_row.fetch_add(1);
                if (y >= y1) break;
                for (uint32_t x = 0; x < s.width; ++x) {
                    const size_t pix = (size_t)y * s.width + x;
                    V3 acc = v3s(0.f);
                    double sx = 0, sy = 0, sz = 0;
                    uint32_t n = 0;
                    PixelOut po{};
                    for (uint32_t f = 0; f < params->spp; ++f) {
                        trace_path(w, *params, max_depth, x, y, f, po, c);
                        V3 L = po.L;
                        if (params->accumulate && n > 0) {
                            const float t = 1.f / ((float)n + 1.f);
                            L = add(acc, mul(t, sub(L, acc)));  // lerp (main.cu:190-196)
                        }
                        acc = L;
                        if (params->accumulate) n++;
                        sx += (double)po.L.x;
                        sy += (double)po.L.y;
                        sz += (double)po.L.z;
                    }
                    if (accum) { accum[4 * pix] = acc.x; accum[4 * pix + 1] = acc.y; accum[4 * pix + 2] = acc.z; accum[4 * pix + 3] = 1.f; }
                    if (albedo) { albedo[3 * pix] = po.albedo.x; albedo[3 * pix + 1] = po.albedo.y; albedo[3 * pix + 2] = po.albedo.z; }
                    if (normal) { normal[3 * pix] = po.normal.x; normal[3 * pix + 1] = po.normal.y; normal[3 * pix + 2] = po.normal.z; }
                    if (test) test[pix] = po.test;
                    if (prim) prim[pix] = po.prim;
                    if (sum64) { sum64[3 * pix] = sx; sum64[3 * pix + 1] = sy; sum64[3 * pix + 2] = sz; }
                }
            }
        });
    }
    for (auto &t : pool) t.join();
    const auto t_end = std::chrono::steady_clock::now();
    if (counts) {
        std::memset(counts, 0, sizeof(*counts));
        counts->build_ms = std::chrono::duration<double, std::milli>(t_render - t_build).count();
        counts->render_ms = std::chrono::duration<double, std::milli>(t_end - t_render).count();
        counts->threads = nthreads;
        for (unsigned ti = 0; ti < nthreads; ++ti) {
            counts->rays_extension += cnt[4 * ti];
            counts->rays_shadow += cnt[4 * ti + 1];
            counts->node_visits += cnt[4 * ti + 2];
            counts->tri_tests += cnt[4 * ti + 3];
        }
    }
    return 0;
}

int oracle_trace(const or_scene *scene, const float *rays, uint32_t count, int any_hit, int brute, uint32_t *hits) {
    if (!scene) return -1;
    Accel acc;
    build_accel(*scene, acc);
    for (uint32_t i = 0; i < count; ++i) {
        const float *r = rays + 8 * (size_t)i;
        HitRec h;
        V3 o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
        if (brute) brute_trace(acc, o, d, r[3], r[7], any_hit != 0, h);
        else accel_trace(acc, o, d, r[3], r[7], any_hit != 0, h, nullptr, nullptr);
        hits[4 * i] = fbits(h.t);
        hits[4 * i + 1] = h.prim;
        hits[4 * i + 2] = fbits(h.b1);
        hits[4 * i + 3] = fbits(h.b2);
    }
    return 0;
}

void oracle_rng(uint32_t pixel, uint32_t sample, uint32_t n, float *out, uint32_t *seeds) {
    Random r;
    r.init(pixel, sample);
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = r.next();
        if (seeds) seeds[i] = r.seed;
    }
}

void oracle_sincos(float x, float *s, float *c) { sincos_or(x, *s, *c); }

void oracle_camera_ray(const or_scene *scene, uint32_t x, uint32_t y, float jx, float jy, float *out6) {
    V3 o, d;
    camera_ray(*scene, x, y, jx, jy, o, d);
    out6[0] = o.x; out6[1] = o.y; out6[2] = o.z; out6[3] = d.x; out6[4] = d.y; out6[5] = d.z;
}

void oracle_bsdf_eval(const or_material *m, const float *wo, const float *wi, float *out) {
    Bsdf b = make_bsdf(*m, V2{0.f, 0.f}, nullptr);
    BRec r;
    r.wo = v3(wo[0], wo[1], wo[2]);
    r.wi = v3(wi[0], wi[1], wi[2]);
    bsdf_eval(b, r);
    out[0] = r.f.x; out[1] = r.f.y; out[2] = r.f.z; out[3] = r.pdf;
}

void oracle_bsdf_sample(const or_material *m, const float *wo, uint32_t *seed, float *out) {
    Bsdf b = make_bsdf(*m, V2{0.f, 0.f}, nullptr);
    Random rng;
    rng.seed = *seed;
    BRec r;
    r.wo = v3(wo[0], wo[1], wo[2]);
    bsdf_sample(b, r, rng);
    *seed = rng.seed;
    out[0] = r.wi.x; out[1] = r.wi.y; out[2] = r.wi.z;
    out[3] = r.f.x; out[4] = r.f.y; out[5] = r.f.z;
    out[6] = r.pdf;
    out[7] = (float)r.lobe;
}

float oracle_fresnel_dielectric(float eta, float ci, float *ct) { float c; float f = fr_diel(eta, ci, c); if (ct) *ct = c; return f; }
float oracle_fresnel_conductor(float eta, float k, float ci) { return fr_cond(eta, k, ci); }
float oracle_ggx_d(const float *wh, float a) { return ggx_d(v3(wh[0], wh[1], wh[2]), a); }
float oracle_ggx_pdf(const float *wo, const float *wh, float a) { return ggx_pdf(v3(wo[0], wo[1], wo[2]), v3(wh[0], wh[1], wh[2]), a); }
void oracle_ggx_sample(const float *wo, float a, float u1, float u2, float *wh) {
    V3 h = ggx_sample(v3(wo[0], wo[1], wo[2]), a, V2{u1, u2});
    wh[0] = h.x; wh[1] = h.y; wh[2] = h.z;
}

int oracle_intersect(const float *o, const float *d, const float *v0, const float *v1, const float *v2, float *out) {
    V3 dd = v3(d[0], d[1], d[2]);
    RayK k = ray_k(dd);
    float t, b1, b2;
    if (!tri_hit(v3(o[0], o[1], o[2]), k, v3(v0[0], v0[1], v0[2]), v3(v1[0], v1[1], v1[2]), v3(v2[0], v2[1], v2[2]), t, b1, b2))
        return 0;
    out[0] = t; out[1] = b1; out[2] = b2;
    return 1;
}

void oracle_bitmap_sample(const or_bitmap *b, float u, float v, float *out3) {
    V3 c = bitmap_lookup(*b, u, v);
    out3[0] = c.x;
    out3[1] = c.y;
    out3[2] = c.z;
}
float oracle_atan2(float y, float x) { return atan2_or(y, x); }
float oracle_acos(float z) { return acos_or(z); }
void oracle_sphere_texcoord(const float *dir, float *uv) {
    V2 t = sphere_uv(v3(dir[0], dir[1], dir[2]));
    uv[0] = t.x;
    uv[1] = t.y;
}
int oracle_intersect_sphere(const float *o, const float *d, const float *center, float radius, float tmin,
                            float tmax, float *t) {
    return sph_hit(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), v3(center[0], center[1], center[2]), radius, tmin,
                   tmax, *t) ? 1 : 0;
}

int oracle_emitter_table(const or_scene *scene, float *sel, float *area, uint32_t capacity, uint32_t *count,
                         float *env_prob) {
    World w;
    w.s = scene;
    build_world(*scene, w);
    *count = (uint32_t)w.emitters.size();
    for (uint32_t i = 0; i < *count && i < capacity; ++i) {
        sel[i] = w.emitters[i].sel;
        area[i] = w.emitters[i].area;
    }
    *env_prob = w.env_sel;
    return 0;
}

}  // extern "C"
