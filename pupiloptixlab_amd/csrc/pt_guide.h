// Neural path-guiding proposal: model constants and the host/device math
// around the network (hash indexing, von Mises-Fisher mixture evaluation and
// sampling, one-sample MIS with a learned selection probability).
//
// No reference counterpart (PupilOptixLab has no guiding).  The model:
//   multiresolution hash grid (16 levels x 2 features, 2^log2T entries per
//   level, FP16) of the hit position -> MLP 32-64-64-64-64 (ReLU, no biases,
//   FP16 MFMA with FP32 accumulation) -> 41 outputs: 8 lobes of
//   (mean xyz, log kappa, log weight) and the selection logit.
// The radiance model is Lhat(w) = sum_k weight_k vMF(w; mean_k, kappa_k); the
// guiding pdf is q = Lhat / sum_k weight_k; the selection probability is
// c = 0.95 sigmoid(o40 - 2), about 0.11 for a fresh network so that early
// frames sample mostly from the BSDF.
#pragma once

#include "pt_math.h"

namespace pupil {

constexpr int kGuideLevels = 16;
constexpr int kGuideFeatures = 2;
constexpr int kGuideInputs = kGuideLevels * kGuideFeatures;  // 32
constexpr int kGuideHidden = 64;
constexpr int kGuideOutputs = 64;  // 41 used
constexpr int kGuideUsed = 41;
constexpr int kGuideLobes = 8;
constexpr float kKappaMinLog = -5.f, kKappaMaxLog = 7.f;
constexpr float kWeightMinLog = -15.f, kWeightMaxLog = 15.f;
constexpr float kSelectMax = 0.95f;
constexpr float kSelectBias = -2.f;
constexpr float kRelEps = 0.01f;

struct GuideLevel {
    uint32_t res;    // grid cells per axis
    uint32_t dense;  // 1 when (res + 1)^3 fits the table
};

// Table entry of grid vertex (x, y, z): dense while (res + 1)^3 fits the
// table; above that the hash of the vertex's 4x4x4 brick selects a 64-entry
// block and the position inside the brick the entry, so the 8 corners of a
// cell mostly fall into one 256-byte block (one or two cache lines instead of
// up to eight).
PT_HD uint32_t guide_corner_index(uint32_t res, uint32_t dense, uint32_t mask, uint32_t x, uint32_t y,
                                  uint32_t z) {
    if (dense) {
        const uint32_t side = res + 1u;
        return (x + y * side + z * side * side) & mask;
    }
    const uint32_t h = (x >> 2) * 1u ^ (y >> 2) * 2654435761u ^ (z >> 2) * 805459861u;
    const uint32_t local = (x & 3u) | ((y & 3u) << 2) | ((z & 3u) << 4);
    return ((h << 6) | local) & mask;
}

// Feature slot (8g + s) of a lane in MFMA group g holds network input/hidden
// index (2c + s/4) * 16 + 4g + (s % 4) for K-chunk c.  The same mapping lets
// one layer's output registers feed the next layer without data movement.
PT_HD int guide_perm(int chunk, int g, int s) { return (2 * chunk + (s >> 2)) * 16 + 4 * g + (s & 3); }

// Mixture parameters decoded from the 41 raw outputs.
struct GuideMixture {
    f3 mu[kGuideLobes];
    float kappa[kGuideLobes];
    float weight[kGuideLobes];
    float wsum;
    float c;
};

PT_HD float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }

PT_HD void guide_decode(const float *o, GuideMixture &m) {
    m.wsum = 0.f;
    for (int k = 0; k < kGuideLobes; ++k) {
        f3 raw = mk3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
        float n2 = dot(raw, raw);
        m.mu[k] = n2 > 1e-24f ? raw * (1.f / sqrtf(n2)) : mk3(0.f, 0.f, 1.f);
        m.kappa[k] = expf(clampf(o[5 * k + 3], kKappaMinLog, kKappaMaxLog));
        m.weight[k] = expf(clampf(o[5 * k + 4], kWeightMinLog, kWeightMaxLog));
        m.wsum += m.weight[k];
    }
    m.c = kSelectMax / (1.f + expf(-(o[40] + kSelectBias)));
}

// vMF density kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (cos - 1)).
PT_HD float vmf_pdf(float kappa, float cosang) {
    return kappa / (2.f * kPi * (-expm1f(-2.f * kappa))) * expf(kappa * (cosang - 1.f));
}

PT_HD float guide_radiance(const GuideMixture &m, f3 w) {
    float s = 0.f;
    for (int k = 0; k < kGuideLobes; ++k) s += m.weight[k] * vmf_pdf(m.kappa[k], dot(m.mu[k], w));
    return s;
}

PT_HD float guide_pdf(const GuideMixture &m, f3 w) { return guide_radiance(m, w) / m.wsum; }

// Samples a lobe with probability weight_k / wsum, then the vMF lobe with
// Jakob's numerically stable inversion.
PT_HD f3 guide_sample(const GuideMixture &m, float u0, float u1, float u2) {
    // Linear scan of the lobe CDF written with constant indices only, so the
    // mixture stays in registers on the GPU.
    float t = u0 * m.wsum;
    f3 mu = m.mu[kGuideLobes - 1];
    float kappa = m.kappa[kGuideLobes - 1];
    bool found = false;
#pragma unroll
    for (int k = 0; k < kGuideLobes - 1; ++k) {
        if (!found) {
            if (t < m.weight[k]) {
                mu = m.mu[k];
                kappa = m.kappa[k];
                found = true;
            } else {
                t -= m.weight[k];
            }
        }
    }
    float wz = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * kappa)) / kappa;
    wz = clampf(wz, -1.f, 1.f);
    const float r = sqrtf(fmaxf(0.f, 1.f - wz * wz));
    float s, c;
    sincos_pt(2.f * kPi * u2, &s, &c);
    f3 local = mk3(r * c, r * s, wz);
    return normalize(to_world(local, mu));
}

}  // namespace pupil
