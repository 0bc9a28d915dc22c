// Scalar/vector math, RNG and sampling helpers shared by the host C++ code and
// the gfx950 kernels of the wavefront path tracer.
//
// Every function here reproduces the floating-point operation order of the
// reference device code it cites, so that a host build and the GPU build of the
// same expression round identically.  Both builds use -ffp-contract=off
// (no fused multiply-add), IEEE division and square root.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PT_HD __host__ __device__ __forceinline__
#define PT_DEV __device__ __forceinline__
#else
#include <math.h>
#define PT_HD inline
#define PT_DEV inline
#endif

namespace pupil {

// Float constants of optix/util.h (M_PIf, M_1_PIf) and EPS (util.h:9).
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kEps = 0.000001f;
constexpr float kMaxDistance = 1e16f;

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

PT_HD f2 mk2(float x, float y) { return f2{x, y}; }
PT_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
PT_HD f3 mk3(float s) { return f3{s, s, s}; }
PT_HD f4 mk4(float x, float y, float z, float w) { return f4{x, y, z, w}; }

// Component-wise operators with the semantics of the reference vector library:
// vector / scalar multiplies by the (correctly rounded) reciprocal.
PT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PT_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
PT_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
PT_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
PT_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PT_HD f3 operator/(f3 a, float s) {
    float inv = 1.0f / s;
    return a * inv;
}
PT_HD f3 operator-(float s, f3 a) { return f3{s - a.x, s - a.y, s - a.z}; }
PT_HD f2 operator+(f2 a, f2 b) { return f2{a.x + b.x, a.y + b.y}; }
PT_HD f2 operator*(f2 a, float s) { return f2{a.x * s, a.y * s}; }
PT_HD f2 operator*(float s, f2 a) { return f2{s * a.x, s * a.y}; }

PT_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_HD float dot(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
PT_HD f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PT_HD float length(f3 v) { return sqrtf(dot(v, v)); }
PT_HD f3 normalize(f3 v) {
    float inv_len = 1.0f / sqrtf(dot(v, v));
    return v * inv_len;
}
PT_HD f4 normalize(f4 v) {
    float inv_len = 1.0f / sqrtf(dot(v, v));
    return f4{v.x * inv_len, v.y * inv_len, v.z * inv_len, v.w * inv_len};
}
PT_HD f3 lerp(f3 a, f3 b, float t) { return a + t * (b - a); }
PT_HD float absf(float v) { return fabsf(v); }
PT_HD float maxf(float a, float b) { return fmaxf(a, b); }
PT_HD float minf(float a, float b) { return fminf(a, b); }

// IsZero / MISWeight / GetLuminance: optix/util.h:206-220, 202-204.
PT_HD bool is_zero(float v) { return fabsf(v) < kEps; }
PT_HD bool is_zero(f3 v) { return fabsf(v.x) < kEps && fabsf(v.y) < kEps && fabsf(v.z) < kEps; }
// A BSDF sample the path can continue with: its value and pdf are finite.
// (Degenerate directions, e.g. a GGX reflection exactly in the tangent plane,
// evaluate 0 / 0.)
PT_HD bool finite_sample(f3 f, float pdf) { return isfinite(f.x + f.y + f.z) && isfinite(pdf); }
PT_HD float mis_weight(float x, float y) { return x / (x + y); }
PT_HD float luminance(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

PT_HD uint32_t f2u(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    return v.u;
}
PT_HD float u2f(uint32_t u) {
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}

// sin/cos for the sampling routines.  Cody-Waite reduction by pi/2 with a
// three-part constant whose leading part has 8 significant bits (so q*A is
// exact for |q| < 2^15), followed by the Taylor polynomials of sin and cos on
// |r| <= pi/4 (truncation error below 2e-9).  One operation order on host and
// GPU, so both round identically; max error is about 1.5 ulp.
PT_HD void sincos_pt(float x, float *s_out, float *c_out) {
    const float two_over_pi = 0.636619772367581343076f;
    const float pio2_a = 1.5703125f;
    const float pio2_b = 4.837512969970703125e-4f;
    const float pio2_c = 7.54978995489188216e-8f;
    float q = rintf(x * two_over_pi);
    float r = ((x - q * pio2_a) - q * pio2_b) - q * pio2_c;
    float r2 = r * r;
    float sp = -1.984126984126984e-4f + r2 * 2.755731922398589e-6f;
    sp = 8.333333333333333e-3f + r2 * sp;
    sp = -1.666666666666667e-1f + r2 * sp;
    float sin_r = r + r * (r2 * sp);
    float cp = 2.480158730158730e-5f + r2 * -2.755731922398589e-7f;
    cp = -1.388888888888889e-3f + r2 * cp;
    cp = 4.166666666666667e-2f + r2 * cp;
    cp = -0.5f + r2 * cp;
    float cos_r = 1.0f + r2 * cp;
    int qi = ((int)q) & 3;
    float s, c;
    if (qi == 0) { s = sin_r; c = cos_r; }
    else if (qi == 1) { s = cos_r; c = -sin_r; }
    else if (qi == 2) { s = -sin_r; c = -cos_r; }
    else { s = -cos_r; c = sin_r; }
    *s_out = s;
    *c_out = c;
}
PT_HD float sin_pt(float x) { float s, c; sincos_pt(x, &s, &c); return s; }
PT_HD float cos_pt(float x) { float s, c; sincos_pt(x, &s, &c); return c; }

// ---------------------------------------------------------------------------
// RNG: cuda/random.h:14-36.  TEA-style seed scramble over (pixel, sample),
// then a 32-bit LCG returning the low 24 bits as a float in [0, 1).
struct Rng {
    uint32_t seed;
};
PT_HD uint32_t rng_tea(uint32_t rounds, uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
    for (uint32_t n = 0; n < rounds; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
PT_HD Rng rng_init(uint32_t pixel, uint32_t sample) { return Rng{rng_tea(4u, pixel, sample)}; }
PT_HD float rng_next(Rng &r) {
    r.seed = 1664525u * r.seed + 1013904223u;
    return (float)(r.seed & 0x00FFFFFFu) / 16777216.0f;
}

// ---------------------------------------------------------------------------
// Sampling helpers: optix/util.h:74-156.
PT_HD f3 uniform_sample_triangle(float u1, float u2) {
    float su = sqrtf(u1);
    return f3{1.f - su, su * (1.f - u2), u2 * su};
}
PT_HD f3 cosine_sample_hemisphere(float u1, float u2) {
    float sin_theta = sqrtf(u1);
    float phi = 2.0f * kPi * u2;
    float s, c;
    sincos_pt(phi, &s, &c);
    f3 p;
    p.x = sin_theta * c;
    p.y = sin_theta * s;
    p.z = sqrtf(fmaxf(0.f, 1.f - sin_theta * sin_theta));
    return p;
}
PT_HD float cosine_hemisphere_pdf(f3 v) { return v.z > 0.f ? kInvPi * v.z : 0.f; }
PT_HD f3 uniform_sample_hemisphere(float u1, float u2) {
    float z = 1.f - 2.f * u1;
    float sin_theta = sqrtf(fmaxf(0.f, 1.f - z * z));
    float phi = 2.0f * kPi * u2;
    float s, c;
    sincos_pt(phi, &s, &c);
    return f3{sin_theta * c, sin_theta * s, fabsf(z)};
}
PT_HD float uniform_hemisphere_pdf(f3 v) { return v.z > 0.f ? kInvPi * 0.5f : 0.f; }
// UniformSampleSphere (optix/util.h:38-43).
PT_HD f3 uniform_sample_sphere(float u1, float u2) {
    float z = 1.f - 2.f * u1;
    float sin_theta = sqrtf(fmaxf(0.f, 1.f - z * z));
    float phi = 2.0f * kPi * u2;
    float s, c;
    sincos_pt(phi, &s, &c);
    return f3{sin_theta * c, sin_theta * s, z};
}

// atan on [0, 1]: for a > tan(pi/8) the identity atan(a) = pi/4 +
// atan((a - 1) / (a + 1)) moves the argument into |u| <= tan(pi/8), where the
// Taylor series u - u^3/3 + u^5/5 - ... up to u^19 leaves a truncation error
// below 5e-10.  Like sincos_pt, one operation order on host and GPU.
PT_HD float atan_unit_pt(float a) {
    float base = 0.f, u = a;
    if (a > 0.41421356f) {
        u = (a - 1.f) / (a + 1.f);
        base = 0.785398163397448f;
    }
    const float u2 = u * u;
    float p = -1.f / 19.f;
    p = 1.f / 17.f + u2 * p;
    p = -1.f / 15.f + u2 * p;
    p = 1.f / 13.f + u2 * p;
    p = -1.f / 11.f + u2 * p;
    p = 1.f / 9.f + u2 * p;
    p = -1.f / 7.f + u2 * p;
    p = 1.f / 5.f + u2 * p;
    p = -1.f / 3.f + u2 * p;
    return base + (u + u * (u2 * p));
}
// Four-quadrant arctangent in (-pi, pi]; atan2(0, 0) = 0.
PT_HD float atan2_pt(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    if (mx == 0.f) return 0.f;
    float r = atan_unit_pt(mn / mx);
    if (ay > ax) r = 1.57079632679490f - r;
    if (x < 0.f) r = 3.14159265358979f - r;
    return y < 0.f ? -r : r;
}
// acos(z) = atan2(sqrt((1 - z)(1 + z)), z) for z clamped to [-1, 1].
PT_HD float acos_pt(float z) {
    z = fminf(fmaxf(z, -1.f), 1.f);
    return atan2_pt(sqrtf((1.f - z) * (1.f + z)), z);
}
// GetSphereTexcoord (optix/util.h:117-128): (phi / 2pi, theta / pi) of a unit
// direction in the sphere's object frame, phi in [0, 2pi).
PT_HD f2 sphere_texcoord(f3 p) {
    float phi = atan2_pt(p.y, p.x);
    if (phi < 0.f) phi = phi + kPi * 2.f;
    const float theta = acos_pt(p.z);
    return f2{phi * kInvPi * 0.5f, theta * kInvPi};
}
PT_HD f3 reflect_z(f3 v) { return f3{-v.x, -v.y, v.z}; }
PT_HD f3 reflect_n(f3 v, f3 n) { return -v + (2.f * dot(v, n)) * n; }
PT_HD f3 refract_z(f3 v, float cos_theta_t, float eta) {
    float scale = -(cos_theta_t < 0.f ? 1.f / eta : eta);
    return normalize(f3{scale * v.x, scale * v.y, cos_theta_t});
}
PT_HD f3 refract_n(f3 v, f3 n, float cos_theta_t, float eta) {
    if (cos_theta_t < 0.f) eta = 1.f / eta;
    return n * (dot(v, n) * eta + cos_theta_t) - v * eta;
}
// Orthonormal basis of Duff et al. 2017 (optix/util.h:136-156).
PT_HD void build_onb(f3 n, f3 &b1, f3 &b2) {
    float sign = copysignf(1.f, n.z);
    float a = -1.f / (sign + n.z);
    float b = n.x * n.y * a;
    b1 = f3{1.f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    b2 = f3{b, sign + n.y * n.y * a, -n.y};
}
PT_HD f3 to_local(f3 v, f3 n) {
    f3 b1, b2;
    build_onb(n, b1, b2);
    return f3{dot(v, b1), dot(v, b2), dot(v, n)};
}
PT_HD f3 to_world(f3 v, f3 n) {
    f3 b1, b2;
    build_onb(n, b1, b2);
    return b1 * v.x + b2 * v.y + n * v.z;
}

}  // namespace pupil
