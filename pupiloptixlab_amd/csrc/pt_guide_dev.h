// Wave-level guiding network on gfx950 matrix cores.
//
// A wave evaluates the network for 64 queries as four 16-query tiles.  Each
// layer is Y^T = W^T X^T computed with v_mfma_f32_16x16x32_f16 (FP16 inputs,
// FP32 accumulation).  Lane l = 16g + i holds A[i][8g..8g+7], B[8g..8g+7][i]
// and D[4g..4g+3][i]; input/hidden index k of slot (8g + s) in K-chunk c is
// guide_perm(c, g, s), so the D registers of one layer are the B registers
// of the next and activations never leave the lane.  Weight operands are
// stored pre-permuted (k_guide_pack) so every lane reads 16 contiguous bytes.
//
// The same lane split serves the hash-grid encoding: lane (g, i) computes
// levels {2g, 2g+1, 8+2g, 9+2g} of query i of the tile, which are exactly
// the input slots it feeds to the first layer, and it receives the gradient
// of the same four levels in the backward pass.
#pragma once

#include <hip/hip_runtime.h>

#include "pt_guide.h"

namespace pupil {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct GuideView {
    const uint32_t *table;     // [grids][levels][T] packed FP16 feature pairs
    const half8 *fwd_a;        // [layer 4][rt 4][chunk 2][lane 64]
    const half8 *bwd_a;        // [layer 4][kt 4][chunk 2][lane 64]
    GuideLevel levels[kGuideLevels];
    uint32_t mask;             // T - 1
    uint32_t log2t;
    f3 lo, inv_ext;            // bounds of grid 0, used to normalize positions
    uint32_t grids;            // hash grids (one per scene); the MLP is shared
    const float4 *grid_bounds; // per grid: (lo.xyz, 0), (inv_ext.xyz, 0)
};

// First entry of grid `grid`, level `level` in the table.
__device__ __forceinline__ size_t guide_level_base(const GuideView &G, uint32_t grid, int level) {
    return ((size_t)grid * kGuideLevels + (size_t)level) << G.log2t;
}

__device__ __forceinline__ floatx4 mfma16(half8 a, half8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f3 guide_normalize_pos(const GuideView &G, f3 p) {
    f3 u = (p - G.lo) * G.inv_ext;
    return mk3(clampf(u.x, 0.f, 1.f), clampf(u.y, 0.f, 1.f), clampf(u.z, 0.f, 1.f));
}

// Position normalized by the bounds of hash grid `grid` (< G.grids).
__device__ __forceinline__ f3 guide_grid_pos(const GuideView &G, uint32_t grid, f3 p) {
    if (grid == 0u) return guide_normalize_pos(G, p);
    const float4 lo = G.grid_bounds[2 * grid], inv = G.grid_bounds[2 * grid + 1];
    const f3 u = (p - mk3(lo.x, lo.y, lo.z)) * mk3(inv.x, inv.y, inv.z);
    return mk3(clampf(u.x, 0.f, 1.f), clampf(u.y, 0.f, 1.f), clampf(u.z, 0.f, 1.f));
}

__device__ __forceinline__ float half_lo(uint32_t v) {
    union { uint16_t u; _Float16 h; } c;
    c.u = (uint16_t)(v & 0xFFFFu);
    return (float)c.h;
}
__device__ __forceinline__ float half_hi(uint32_t v) {
    union { uint16_t u; _Float16 h; } c;
    c.u = (uint16_t)(v >> 16);
    return (float)c.h;
}

// Trilinear corner set of one level.
struct LevelCorners {
    uint32_t idx[8];
    float w[8];
};

__device__ __forceinline__ int guide_lane_level(int g, int t) { return (t >> 1) * 8 + 2 * g + (t & 1); }

__device__ __forceinline__ void level_corners(const GuideView &G, int level, f3 u, LevelCorners &lc) {
    const uint32_t res = G.levels[level].res, dense = G.levels[level].dense;
    const float x = u.x * (float)res, y = u.y * (float)res, z = u.z * (float)res;
    const uint32_t cx = min((uint32_t)x, res - 1u), cy = min((uint32_t)y, res - 1u), cz = min((uint32_t)z, res - 1u);
    const float fx = x - (float)cx, fy = y - (float)cy, fz = z - (float)cz;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        lc.idx[k] = guide_corner_index(res, dense, G.mask, cx + dx, cy + dy, cz + dz);
        lc.w[k] = (dx ? fx : 1.f - fx) * (dy ? fy : 1.f - fy) * (dz ? fz : 1.f - fz);
    }
}

// First-layer B operand of lane group g for the query at normalized position
// u in hash grid `grid`.
__device__ __forceinline__ half8 guide_encode(const GuideView &G, f3 u, int g, uint32_t grid = 0u) {
    half8 b;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int level = guide_lane_level(g, t);
        LevelCorners lc;
        level_corners(G, level, u, lc);
        const uint32_t *tab = G.table + guide_level_base(G, grid, level);
        float a0 = 0.f, a1 = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t v = tab[lc.idx[k]];
            a0 += lc.w[k] * half_lo(v);
            a1 += lc.w[k] * half_hi(v);
        }
        b[2 * t] = (_Float16)a0;
        b[2 * t + 1] = (_Float16)a1;
    }
    return b;
}

__device__ __forceinline__ half8 fwd_operand(const GuideView &G, int layer, int rt, int chunk, int lane) {
    return G.fwd_a[((layer * 4 + rt) * 2 + chunk) * 64 + lane];
}
__device__ __forceinline__ half8 bwd_operand(const GuideView &G, int layer, int kt, int chunk, int lane) {
    return G.bwd_a[((layer * 4 + kt) * 2 + chunk) * 64 + lane];
}

// ReLU of a layer's D registers packed as the next layer's two B chunks.
__device__ __forceinline__ void relu_pack(const floatx4 d[4], half8 h[2]) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int s = 0; s < 8; ++s) h[c][s] = (_Float16)fmaxf(d[2 * c + (s >> 2)][s & 3], 0.f);
}

// Forward pass of one 16-query tile; `act` receives each layer's input
// operands (layer 0: act[0][0]; layers 1..3: act[l][0..1]) for training.
__device__ __forceinline__ void guide_mlp_tile(const GuideView &G, int lane, half8 in0, floatx4 out[4],
                                               half8 (*act)[2]) {
    floatx4 d[4];
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
    if (act) act[0][0] = in0;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) d[rt] = mfma16(fwd_operand(G, 0, rt, 0, lane), in0, zero);
    half8 h[2];
    relu_pack(d, h);
#pragma unroll
    for (int l = 1; l < 4; ++l) {
        if (act) {
            act[l][0] = h[0];
            act[l][1] = h[1];
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            d[rt] = mfma16(fwd_operand(G, l, rt, 0, lane), h[0], zero);
            d[rt] = mfma16(fwd_operand(G, l, rt, 1, lane), h[1], d[rt]);
        }
        if (l < 3) relu_pack(d, h);
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) out[rt] = d[rt];
}

// Forward pass of one 16-query tile with the weight operands read from a
// shared (LDS) copy of fwd_a (same layout).
__device__ __forceinline__ void guide_mlp_tile_shared(const half8 *w, int lane, half8 in0, floatx4 out[4]) {
    floatx4 d[4];
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) d[rt] = mfma16(w[(rt * 2) * 64 + lane], in0, zero);
    half8 h[2];
    relu_pack(d, h);
#pragma unroll
    for (int l = 1; l < 4; ++l) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            d[rt] = mfma16(w[((l * 4 + rt) * 2) * 64 + lane], h[0], zero);
            d[rt] = mfma16(w[((l * 4 + rt) * 2 + 1) * 64 + lane], h[1], d[rt]);
        }
        if (l < 3) relu_pack(d, h);
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) out[rt] = d[rt];
}

// Inference for the 64 queries of a wave in hash grid `grid` (wave-uniform):
// lane L supplies the position of query L (any value when `active` is false)
// and receives its 41 outputs in `stage` (LDS, 64 x 41 floats per wave).
// Must be called by all 64 lanes.
__device__ __forceinline__ void guide_infer_wave(const GuideView &G, f3 pos, float *stage, uint32_t grid = 0u) {
    const int lane = (int)__lane_id();
    const int g = lane >> 4, i = lane & 15;
    const f3 u = guide_grid_pos(G, grid, pos);
#pragma unroll
    for (int qt = 0; qt < 4; ++qt) {
        const int q = qt * 16 + i;
        const f3 uq = mk3(__shfl(u.x, q), __shfl(u.y, q), __shfl(u.z, q));
        const half8 in0 = guide_encode(G, uq, g, grid);
        floatx4 out[4];
        guide_mlp_tile(G, lane, in0, out, nullptr);
#pragma unroll
        for (int rt = 0; rt < 3; ++rt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int n = rt * 16 + 4 * g + j;
                if (n < kGuideUsed) stage[q * kGuideUsed + n] = out[rt][j];
            }
    }
}

}  // namespace pupil
