// Denoiser for the path tracer's output: the DenoisePass of the reference
// (system/denoise_pass.cpp:31-93) runs the closed OptiX AI denoiser
// (optix/denoiser.cpp) on the noisy buffer with optional albedo / normal
// guides and writes the "final result" buffer.  Here the filter is the
// edge-avoiding a-trous wavelet transform (Dammertz, Sewtz, Hanika, Lensch,
// HPG 2010) with the variance-guided luminance edge-stopping function of
// SVGF (Schied et al., HPG 2017), on albedo-demodulated color:
//
//   prepare   c = color / albedo (when albedo guides), l = luminance(c), and a
//             spatial luminance variance over a 5 x 5 window of pixels with a
//             similar normal / albedo;
//   k levels  5 x 5 B3-spline taps (1/16, 1/4, 3/8, 1/4, 1/16) at stride 2^k,
//             weight = h(x) h(y) w_l w_n w_a with
//               w_l = exp(-|l_p - l_q| / (sigma_l sqrt(g3x3(var_p)) + 1e-4)),
//               w_n = max(0, n_p . n_q)^sigma_n,
//               w_a = exp(-|a_p - a_q|^2 / sigma_a^2),
//             color' = sum w c / sum w, variance' = sum w^2 var / (sum w)^2;
//   finish    color x albedo.
//
// Each level is one stencil pass over HBM-resident float4 images; the taps
// of a 256-thread block (16 x 16 pixels) hit L2, so a 1080p pass moves about
// (16 + 12 + 12) B per pixel per tap through the cache and ~48 B per pixel
// from HBM.
#include <hip/hip_runtime.h>

#include <cmath>

#include "pt_denoise.h"
#include "pt_math.h"

namespace pupil {
namespace {

constexpr int kTile = 16;

struct Guides {
    const float *albedo;  // 3 floats per pixel, or null
    const float *normal;  // 3 floats per pixel, or null
};

__device__ __forceinline__ f3 load3(const float *p, size_t i) { return mk3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

__device__ __forceinline__ float guide_weight(const Guides &g, size_t p, size_t q, const DenoiseParams &prm) {
    float w = 1.f;
    if (g.normal) {
        // Pixels without a surface carry a zero normal: they only blend with
        // each other.
        const f3 np = load3(g.normal, p), nq = load3(g.normal, q);
        const bool zp = dot(np, np) == 0.f, zq = dot(nq, nq) == 0.f;
        if (zp || zq) {
            if (zp != zq) return 0.f;
        } else {
            w *= powf(fmaxf(0.f, dot(np, nq)), prm.sigma_normal);
        }
    }
    if (g.albedo) {
        const f3 da = load3(g.albedo, p) - load3(g.albedo, q);
        w *= expf(-dot(da, da) / (prm.sigma_albedo * prm.sigma_albedo));
    }
    return w;
}

// Demodulated color (xyz) and the spatial luminance variance (w).
__global__ void k_dn_prepare(const float4 *in, Guides g, float4 *out, uint32_t w, uint32_t h, DenoiseParams prm) {
    const int x = blockIdx.x * kTile + threadIdx.x % kTile, y = blockIdx.y * kTile + threadIdx.x / kTile;
    if (x >= (int)w || y >= (int)h) return;
    const size_t p = (size_t)y * w + x;
    float s1 = 0.f, s2 = 0.f, sw = 0.f;
    f3 cp = mk3(0.f);
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
            const int qx = x + dx, qy = y + dy;
            if (qx < 0 || qy < 0 || qx >= (int)w || qy >= (int)h) continue;
            const size_t q = (size_t)qy * w + qx;
            const float4 c4 = in[q];
            f3 c = mk3(c4.x, c4.y, c4.z);
            if (g.albedo) {
                const f3 a = load3(g.albedo, q);
                c = mk3(c.x / fmaxf(a.x, kDenoiseAlbedoEps), c.y / fmaxf(a.y, kDenoiseAlbedoEps),
                        c.z / fmaxf(a.z, kDenoiseAlbedoEps));
            }
            if (q == p) cp = c;
            const float wq = guide_weight(g, p, q, prm);
            const float l = luminance(c);
            s1 += wq * l;
            s2 += wq * l * l;
            sw += wq;
        }
    const float m1 = s1 / sw, m2 = s2 / sw;
    out[p] = make_float4(cp.x, cp.y, cp.z, fmaxf(0.f, m2 - m1 * m1));
}

// One a-trous level at stride `step`.
__global__ void k_dn_atrous(const float4 *in, Guides g, float4 *out, uint32_t w, uint32_t h, int step,
                            DenoiseParams prm) {
    const int x = blockIdx.x * kTile + threadIdx.x % kTile, y = blockIdx.y * kTile + threadIdx.x / kTile;
    if (x >= (int)w || y >= (int)h) return;
    const size_t p = (size_t)y * w + x;
    const float4 c0 = in[p];
    // 3 x 3 Gaussian (1/4, 1/2, 1/4) of the variance at the center.
    float gv = 0.f, gw = 0.f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int qx = x + dx, qy = y + dy;
            if (qx < 0 || qy < 0 || qx >= (int)w || qy >= (int)h) continue;
            const float k = (dx == 0 ? 0.5f : 0.25f) * (dy == 0 ? 0.5f : 0.25f);
            gv += k * in[(size_t)qy * w + qx].w;
            gw += k;
        }
    const float sigma = prm.sigma_luminance * sqrtf(gv / gw) + 1e-4f;
    const float l0 = luminance(mk3(c0.x, c0.y, c0.z));
    const float kern[5] = {1.f / 16.f, 1.f / 4.f, 3.f / 8.f, 1.f / 4.f, 1.f / 16.f};
    f3 sc = mk3(0.f);
    float sv = 0.f, sw = 0.f;
    for (int j = -2; j <= 2; ++j)
        for (int i = -2; i <= 2; ++i) {
            const int qx = x + i * step, qy = y + j * step;
            if (qx < 0 || qy < 0 || qx >= (int)w || qy >= (int)h) continue;
            const size_t q = (size_t)qy * w + qx;
            const float4 cq = in[q];
            const float lq = luminance(mk3(cq.x, cq.y, cq.z));
            const float wl = expf(-fabsf(l0 - lq) / sigma);
            const float wq = kern[i + 2] * kern[j + 2] * wl * guide_weight(g, p, q, prm);
            sc = sc + wq * mk3(cq.x, cq.y, cq.z);
            sv += wq * wq * cq.w;
            sw += wq;
        }
    out[p] = make_float4(sc.x / sw, sc.y / sw, sc.z / sw, sv / (sw * sw));
}

__global__ void k_dn_finish(const float4 *in, Guides g, float4 *out, uint32_t w, uint32_t h) {
    const int x = blockIdx.x * kTile + threadIdx.x % kTile, y = blockIdx.y * kTile + threadIdx.x / kTile;
    if (x >= (int)w || y >= (int)h) return;
    const size_t p = (size_t)y * w + x;
    const float4 c = in[p];
    f3 r = mk3(c.x, c.y, c.z);
    if (g.albedo) {
        const f3 a = load3(g.albedo, p);
        r = mk3(r.x * fmaxf(a.x, kDenoiseAlbedoEps), r.y * fmaxf(a.y, kDenoiseAlbedoEps),
                r.z * fmaxf(a.z, kDenoiseAlbedoEps));
    }
    out[p] = make_float4(r.x, r.y, r.z, 1.f);
}

}  // namespace

hipError_t denoise_launch(hipStream_t stream, uint32_t w, uint32_t h, const float4 *in, const float *albedo,
                          const float *normal, float4 *out, float4 *scratch, const DenoiseParams &prm) {
    const dim3 grid((w + kTile - 1) / kTile, (h + kTile - 1) / kTile), block(kTile * kTile);
    const Guides g{albedo, normal};
    float4 *a = scratch, *b = scratch + (size_t)w * h;
    hipLaunchKernelGGL(k_dn_prepare, grid, block, 0, stream, in, g, a, w, h, prm);
    for (uint32_t k = 0; k < prm.iterations; ++k) {
        hipLaunchKernelGGL(k_dn_atrous, grid, block, 0, stream, a, g, b, w, h, 1 << k, prm);
        float4 *t = a;
        a = b;
        b = t;
    }
    hipLaunchKernelGGL(k_dn_finish, grid, block, 0, stream, a, g, out, w, h);
    return hipGetLastError();
}

}  // namespace pupil
