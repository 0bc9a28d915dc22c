// Online training of the guiding network on gfx950.
//
// One optimizer step over a batch of samples is a chain of kernels on the
// tracer's stream, all driven by the device control block (GuideCtl):
//   k_guide_begin  : batch size of this step from the collected sample count
//   k_guide_train  : per wave, 16 samples: hash encoding, MLP forward (MFMA),
//                    loss and output gradient, MLP backward (MFMA), hash-table
//                    gradient scatter; stores every layer's input and output
//                    gradient (FP16) for the weight gradient
//   k_guide_dw     : dW_l = X_l^T dY_l per batch slice (MFMA, FP32 atomics)
//   k_guide_adam_table : sparse Adam on the FP32 table; refreshes the FP16 copy
//   k_guide_finish : Adam on the MLP weights, re-packing of the FP16 MFMA
//                    operands, step count, dynamic loss scale, bookkeeping
// The backward pass runs in FP16 with FP32 accumulation; output gradients are
// multiplied by a dynamic loss scale, and a step whose gradients are not
// finite is dropped and the scale halved.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pt_guide_state.h"

using namespace pupil;

namespace {

constexpr int kTrainBlock = 256;  // 4 waves

__device__ __forceinline__ uint16_t hbits(_Float16 h) {
    union {
        _Float16 h;
        uint16_t u;
    } c;
    c.h = h;
    return c.u;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// all_slots: every slot of [0, capacity) holds a sample or a zero-weight
// placeholder (interleaved sample regions), else the last collection's
// samples are packed at [0, n_samples).
__global__ void k_guide_begin(GuideCtl *ctl, uint32_t first, uint32_t batch, uint32_t capacity, float fp_explicit,
                              uint32_t all_slots) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t n = all_slots ? capacity : min(ctl->n_samples, capacity);
    const uint32_t valid = n > first ? min(batch, n - first) : 0u;
    ctl->batch_valid = valid;
    ctl->active = valid > 0u ? 1u : 0u;
    ctl->loss_sum = 0.f;
    ctl->loss_cnt = 0.f;
    // Selection-loss normalization: the mean F/p of the samples that carry
    // a contribution (given explicitly by host-side training).
    float norm = fp_explicit;
    if (!(norm > 0.f)) norm = (ctl->fp_cnt > 0.f && ctl->fp_sum > 0.f) ? ctl->fp_sum / ctl->fp_cnt : 1.f;
    ctl->fp_norm = norm;
}

struct TrainBuffers {
    uint16_t *xt;       // [layer][k 64][batch] layer inputs (FP16 bits)
    uint16_t *dyt;      // [layer][n 64][batch] scaled layer output gradients (FP16 bits)
    float2 *table_grad; // [level][T]
    uint32_t *touched;  // [grid]: set to 1 by every step sample of the grid (Adam visits only those grids)
};

struct LossParams {
    uint32_t flags;     // pt_guide_loss_flags
    float l2_weight;
};

__device__ __forceinline__ float xor_sum4(float v) {
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}
__device__ __forceinline__ float xor_max4(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}

// One optimization pass over 16 samples per wave (one MFMA tile).  Lane
// (g, i) = 16 g + i belongs to sample i: it encodes levels {2g, 2g+1, 8+2g,
// 9+2g} of the sample's position, and in the loss it owns lobes 2g and 2g+1;
// sums over lobes are reduced across the sample's four lanes (xor 16, 32).
//
// Loss per sample (w = min(F/p / fp_norm, 10), Lhat = W q, W = sum_k w_k):
//   relative L2          l2 (W sg(q) - L)^2 / (sg(Lhat)^2 + 0.01)  (trains W)
//   guide cross-entropy  -w log q                                  (shapes the lobes)
//   mixture              -w log(c sg(q) + (1 - c) p_bsdf)          (trains c)
// The guide term is differentiated through the responsibilities
// r_k = w_k v_k / Lhat, formed in the log domain, so directions no lobe covers
// still get finite gradients.  Per sample, the relative-L2 part of the output
// gradient is clipped in max norm to kGuideL2Clip and the cross-entropy part
// to kGuideGradClip.  Gradients are weighted by the sample weight and summed;
// the Adam kernels divide by the weight total.
//
// samples: 3 float4 per sample (position, direction, {target, F/p, BSDF pdf,
// weight}); weight 0 marks an unused slot.
// Batch element q reads slot first + q, or with a ring of ring_n = m B
// slots slot (first + q m) mod ring_n: every m-th slot, so one batch spans
// the whole ring (the samples of several launches).
__global__ __launch_bounds__(kTrainBlock) void k_guide_train(GuideView G, const float4 *samples, uint32_t batch,
                                                             GuideCtl *ctl, TrainBuffers B, LossParams lp,
                                                             uint32_t first, uint32_t ring_n, uint32_t ring_m) {
    __shared__ float stage[kTrainBlock / 64][16 * kGuideUsed];
    if (!ctl->active) return;
    const uint32_t nvalid = ctl->batch_valid;
    const int wave = threadIdx.x >> 6;
    const int lane = (int)__lane_id();
    const int g = lane >> 4, i = lane & 15;
    const uint32_t base = (blockIdx.x * (kTrainBlock / 64) + wave) * 16u;
    if (base >= nvalid) return;  // wave-uniform; no block barriers below
    float *st = stage[wave];
    const float scale = ctl->loss_scale;
    const float fp_norm = ctl->fp_norm;

    const uint32_t me = base + i;
    const bool in_range = me < nvalid;
    // Lanes past the batch's valid samples read slot `first` (always filled)
    // and carry weight 0, so no lane reads past the collected samples.
    const size_t slot = ring_n ? ((size_t)first + (size_t)(in_range ? me : 0u) * ring_m) % ring_n
                               : (size_t)first + (in_range ? me : 0u);
    const float4 p4 = samples[3 * slot];
    const float4 d4 = samples[3 * slot + 1];
    const float4 aux = samples[3 * slot + 2];
    // With several hash grids, position.w names the sample's grid; a sample
    // naming no grid is unused.
    uint32_t grid = 0u;
    bool grid_ok = true;
    if (G.grids > 1u) {
        grid_ok = p4.w >= 0.f && p4.w < (float)G.grids;
        grid = grid_ok ? (uint32_t)p4.w : 0u;
    }
    const float weight = (in_range && grid_ok) ? aux.w : 0.f;
    const bool valid = weight > 0.f;
    if (valid) B.touched[grid] = 1u;  // same value from every writer
    const f3 u = guide_grid_pos(G, grid, mk3(p4.x, p4.y, p4.z));

    // Forward (inputs of every layer kept for the backward pass).
    const half8 in0 = guide_encode(G, u, g, grid);
    half8 act[4][2];
    floatx4 out[4];
    guide_mlp_tile(G, lane, in0, out, act);
#pragma unroll
    for (int rt = 0; rt < 3; ++rt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = rt * 16 + 4 * g + j;
            if (n < kGuideUsed) st[i * kGuideUsed + n] = out[rt][j];
        }
    wave_sync_lds();

    // Loss and output gradient; this lane holds lobes 2g and 2g + 1.
    {
        const float *o = st + i * kGuideUsed;
        const f3 w = mk3(d4.x, d4.y, d4.z);
        float lt[2], kap[2], wt[2], cs[2], lwv[2], lkv[2];
        f3 mu[2];
        float lhat_p = 0.f, wsum_p = 0.f, lmax_p = -3.0e38f;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int k = 2 * g + t;
            const f3 raw = mk3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
            const float n2 = dot(raw, raw);
            mu[t] = n2 > 1e-24f ? raw * (1.f / sqrtf(n2)) : mk3(0.f, 0.f, 1.f);
            lkv[t] = o[5 * k + 3];
            lwv[t] = o[5 * k + 4];
            const float lkc = clampf(lkv[t], kKappaMinLog, kKappaMaxLog);
            const float lwc = clampf(lwv[t], kWeightMinLog, kWeightMaxLog);
            kap[t] = expf(lkc);
            wt[t] = expf(lwc);
            cs[t] = dot(mu[t], w);
            lhat_p += wt[t] * vmf_pdf(kap[t], cs[t]);
            wsum_p += wt[t];
            lt[t] = lwc + lkc - logf(2.f * kPi) - logf(-expm1f(-2.f * kap[t])) + kap[t] * (cs[t] - 1.f);
            lmax_p = fmaxf(lmax_p, lt[t]);
        }
        const float lhat = xor_sum4(lhat_p), wsum = xor_sum4(wsum_p), lmax = xor_max4(lmax_p);
        const float r0 = expf(lt[0] - lmax), r1 = expf(lt[1] - lmax);
        const float rsum = xor_sum4(r0 + r1);
        const float target = aux.x, fs = aux.y, pb = aux.z;
        const float q = lhat / wsum;
        const float den = lhat * lhat + kRelEps;
        const float c = kSelectMax / (1.f + expf(-(o[40] + kSelectBias)));
        const float pmix = fmaxf(c * q + (1.f - c) * pb, 1e-30f);
        const float wgt = fminf(fs / fp_norm, 10.f);
        const float l2w = (lp.flags & PT_GUIDE_NO_RADIANCE_L2) ? 0.f : lp.l2_weight;
        const float cew = (lp.flags & PT_GUIDE_NO_GUIDE_CE) ? 0.f : wgt;
        const float mxw = (lp.flags & PT_GUIDE_NO_MIXTURE_CE) ? 0.f : wgt;
        const float dl = 2.f * (lhat - target) / den;
        const float loss = l2w * (lhat - target) * (lhat - target) / den - cew * logf(fmaxf(q, 1e-30f)) -
                           mxw * logf(pmix);
        const float rr[2] = {r0 / rsum, r1 / rsum};
        float ge[2][5], g2[2];
        float mx = 0.f, mx2 = 0.f;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const bool in_w = lwv[t] > kWeightMinLog && lwv[t] < kWeightMaxLog;
            const bool in_k = lkv[t] > kKappaMinLog && lkv[t] < kKappaMaxLog;
            g2[t] = in_w ? dl * wt[t] * q : 0.f;
            const float b = -cew * rr[t];
            ge[t][4] = in_w ? b + cew * wt[t] / wsum : 0.f;
            const float dlogv = 1.f / kap[t] - 2.f / expm1f(2.f * kap[t]) + cs[t] - 1.f;
            ge[t][3] = in_k ? b * dlogv * kap[t] : 0.f;
            const int k = 2 * g + t;
            const f3 raw = mk3(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
            const float nr = sqrtf(dot(raw, raw));
            f3 dr = mk3(0.f);
            if (nr > 1e-12f) dr = (w - cs[t] * mu[t]) * (b * kap[t] / nr);
            ge[t][0] = dr.x;
            ge[t][1] = dr.y;
            ge[t][2] = dr.z;
#pragma unroll
            for (int e = 0; e < 5; ++e) mx = fmaxf(mx, fabsf(ge[t][e]));
            mx2 = fmaxf(mx2, fabsf(g2[t]));
        }
        const float sig = c / kSelectMax;
        const float g40 = -mxw * (q - pb) / pmix * kSelectMax * sig * (1.f - sig);
        mx = fmaxf(xor_max4(mx), fabsf(g40));
        mx2 = xor_max4(mx2);
        const float s1 = mx > kGuideGradClip ? kGuideGradClip / mx : 1.f;
        const float s2 = l2w * (mx2 > kGuideL2Clip ? kGuideL2Clip / mx2 : 1.f);
        const float ws = valid ? weight * scale : 0.f;
        wave_sync_lds();
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int k = 2 * g + t;
#pragma unroll
            for (int e = 0; e < 4; ++e) st[i * kGuideUsed + 5 * k + e] = ge[t][e] * s1 * ws;
            st[i * kGuideUsed + 5 * k + 4] = (ge[t][4] * s1 + g2[t] * s2) * ws;
        }
        if (g == 0) st[i * kGuideUsed + 40] = g40 * s1 * ws;
        // weighted loss and weight total (counted once per sample, by g == 0)
        float ls = (valid && g == 0) ? loss * weight : 0.f, vc = (valid && g == 0) ? weight : 0.f;
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) {
            ls += __shfl_xor(ls, off);
            vc += __shfl_xor(vc, off);
        }
        if (lane == 0 && vc > 0.f) {
            unsafeAtomicAdd(&ctl->loss_sum, ls);
            unsafeAtomicAdd(&ctl->loss_cnt, vc);
        }
    }
    wave_sync_lds();

    // Backward.
    const float inv_scale = 1.f / scale;
    const uint32_t qg = base + i;
    half8 dy[2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int n = guide_perm(c, g, s);
            dy[c][s] = (_Float16)(n < kGuideUsed ? st[i * kGuideUsed + n] : 0.f);
        }
    const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
    bool bad = false;
#pragma unroll
    for (int l = 3; l >= 0; --l) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int s = 0; s < 8; ++s)
                B.dyt[((size_t)l * 64 + guide_perm(c, g, s)) * batch + qg] = hbits(dy[c][s]);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (l == 0 && c == 1) break;
#pragma unroll
            for (int s = 0; s < 8; ++s)
                B.xt[((size_t)l * 64 + guide_perm(c, g, s)) * batch + qg] = hbits(act[l][c][s]);
        }
        // dX_l^T = W_l dY_l^T
        floatx4 dx[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            if (l == 0 && kt >= 2) break;
            dx[kt] = mfma16(bwd_operand(G, l, kt, 0, lane), dy[0], zero);
            dx[kt] = mfma16(bwd_operand(G, l, kt, 1, lane), dy[1], dx[kt]);
        }
        if (l > 0) {
            // ReLU of the layer input (the previous layer's output).
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const float gv = dx[2 * c + (s >> 2)][s & 3];
                    dy[c][s] = (_Float16)((float)act[l][c][s] > 0.f ? gv : 0.f);
                }
        } else if (valid) {
            // Hash-table gradient of this lane's four levels.
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float g0 = dx[t >> 1][2 * (t & 1)] * inv_scale;
                const float g1 = dx[t >> 1][2 * (t & 1) + 1] * inv_scale;
                if (g0 == 0.f && g1 == 0.f) continue;
                if (!isfinite(g0) || !isfinite(g1)) {
                    bad = true;
                    continue;
                }
                const int level = guide_lane_level(g, t);
                LevelCorners lc;
                level_corners(G, level, u, lc);
                float2 *tg = B.table_grad + guide_level_base(G, grid, level);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    unsafeAtomicAdd(&tg[lc.idx[k]].x, lc.w[k] * g0);
                    unsafeAtomicAdd(&tg[lc.idx[k]].y, lc.w[k] * g1);
                }
            }
        }
    }
    if (bad) atomicOr(&ctl->bad, 1u);
}

// dW_l[k][n] += sum_q X_l[q][k] dY_l[q][n] over one batch slice (kGuideDwSlice
// samples), one wave per (slice, layer); partial sums are added atomically.
// A[i][8g+s] = X_l[q + 8g + s][16 kt + i], B[8g+s][i] = dY_l[q + 8g + s][16 nt + i],
// D[4g+j][i] = dW[16 kt + 4g + j][16 nt + i].
__global__ __launch_bounds__(64) void k_guide_dw(const uint16_t *xt, const uint16_t *dyt, uint32_t batch,
                                                 GuideCtl *ctl, float *w_grad) {
    if (!ctl->active) return;
    const uint32_t rows = (ctl->batch_valid + 63u) & ~63u;
    const uint32_t q0 = blockIdx.x * kGuideDwSlice;
    if (q0 >= rows) return;
    const uint32_t q1 = min(rows, q0 + kGuideDwSlice);
    const int layer = blockIdx.y;
    const int lane = (int)__lane_id();
    const int g = lane >> 4, i = lane & 15;
    const int ktiles = layer == 0 ? 2 : 4;
    floatx4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *X = xt + (size_t)layer * 64 * batch;
    const uint16_t *Y = dyt + (size_t)layer * 64 * batch;
    for (uint32_t q = q0; q < q1; q += 32) {
        half8 bv[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            bv[nt] = *reinterpret_cast<const half8 *>(Y + (size_t)(nt * 16 + i) * batch + q + 8 * g);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            if (kt >= ktiles) break;
            const half8 av = *reinterpret_cast<const half8 *>(X + (size_t)(kt * 16 + i) * batch + q + 8 * g);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[kt][nt] = mfma16(av, bv[nt], acc[kt][nt]);
        }
    }
    const float s = 1.f / ctl->loss_scale;
    float *W = w_grad + (size_t)layer * 64 * 64;
    bool nonfinite = false;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
        if (kt >= ktiles) break;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = acc[kt][nt][j] * s;
                if (!isfinite(v)) {
                    nonfinite = true;
                    continue;
                }
                if (v != 0.f) unsafeAtomicAdd(&W[(kt * 16 + 4 * g + j) * 64 + nt * 16 + i], v);
            }
    }
    if (nonfinite) atomicOr(&ctl->bad, 1u);
}

struct AdamParams {
    float lr, b1, b2, eps;
};

// 1 / (sum of sample weights of the step).
__device__ __forceinline__ float inv_valid(const GuideCtl *ctl) {
    return ctl->loss_cnt > 0.f ? 1.f / ctl->loss_cnt : 0.f;
}

// Whether this step applies, and its Adam bias corrections (step count t =
// ctl->step + 1; the count itself is advanced by k_guide_finish).
__device__ __forceinline__ bool adam_step_of(const GuideCtl *ctl, uint32_t apply, const AdamParams &a, float &c1,
                                            float &c2) {
    const float t = (float)(ctl->step + 1u);
    c1 = 1.f - powf(a.b1, t);
    c2 = 1.f - powf(a.b2, t);
    return apply && !ctl->bad;
}

// Sparse Adam on the hash table (FP32 master + FP16 copy) with the
// batch-summed gradient divided by the sample-weight total: entries no sample
// touched (gradient exactly zero) keep their values and moments.  Gradients
// are cleared.
// Grids [grid_lo, grid_hi) of grid_entries entries each; grids that no
// sample of the step touched have an all-zero gradient and are skipped whole
// (with one grid per batch view, a step's samples name a few of them).
__global__ void k_guide_adam_table(float2 *p, float2 *m, float2 *v, float2 *g, uint32_t *p16, size_t grid_entries,
                                   uint32_t grid_lo, uint32_t grid_hi, const uint32_t *touched, AdamParams a,
                                   const GuideCtl *ctl, uint32_t apply) {
    if (!ctl->active) return;
    float c1, c2;
    const bool step = adam_step_of(ctl, apply, a, c1, c2);
    const float inv = inv_valid(ctl);
    for (uint32_t gi = grid_lo; gi < grid_hi; ++gi) {
        if (!touched[gi]) continue;
        const size_t e0 = (size_t)gi * grid_entries, e1 = e0 + grid_entries;
        const size_t stride = (size_t)gridDim.x * blockDim.x;
        for (size_t e = e0 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < e1; e += stride) {
            float2 gg = g[e];
            if (gg.x == 0.f && gg.y == 0.f) continue;
            g[e] = make_float2(0.f, 0.f);
            if (!step) continue;
            gg.x *= inv;
            gg.y *= inv;
            float2 mm = m[e], vv = v[e], pp = p[e];
            mm.x = a.b1 * mm.x + (1.f - a.b1) * gg.x;
            mm.y = a.b1 * mm.y + (1.f - a.b1) * gg.y;
            vv.x = a.b2 * vv.x + (1.f - a.b2) * gg.x * gg.x;
            vv.y = a.b2 * vv.y + (1.f - a.b2) * gg.y * gg.y;
            pp.x -= a.lr * (mm.x / c1) / (sqrtf(vv.x / c2) + a.eps);
            pp.y -= a.lr * (mm.y / c1) / (sqrtf(vv.y / c2) + a.eps);
            m[e] = mm;
            v[e] = vv;
            p[e] = pp;
            p16[e] = (uint32_t)hbits((_Float16)pp.x) | ((uint32_t)hbits((_Float16)pp.y) << 16);
        }
    }
}

constexpr int kFinishBlock = 1024;

// Single block: Adam on the MLP weights (gradient = w_grad / weight total;
// w_grad is cleared), then the FP16 MFMA operands are re-packed and the step
// is closed (step count, dynamic loss scale, loss bookkeeping, the grids'
// touched flags cleared).
//   fwd[layer][rt][chunk][lane][s] = W[perm(chunk, g, s)][16 rt + i]
//   bwd[layer][kt][chunk][lane][s] = W[16 kt + i][perm(chunk, g, s)]
// apply = 0 only reports the gradient (grad_out) and closes the step.
__global__ __launch_bounds__(kFinishBlock) void k_guide_finish(float *p, float *m, float *v, float *w_grad,
                                                              AdamParams a, GuideCtl *ctl, uint32_t apply,
                                                              float *grad_out, _Float16 *fwd, _Float16 *bwd,
                                                              uint32_t *touched, uint32_t grids) {
    for (uint32_t gi = threadIdx.x; gi < grids; gi += kFinishBlock) touched[gi] = 0u;
    const bool active = ctl->active != 0u;
    float c1, c2;
    const bool step = active && adam_step_of(ctl, apply, a, c1, c2);
    const float inv = inv_valid(ctl);
    const uint32_t n = 4 * 64 * 64;
    if (active) {
        for (uint32_t e = threadIdx.x; e < n; e += kFinishBlock) {
            const float gg = w_grad[e] * inv;
            w_grad[e] = 0.f;
            if (grad_out) grad_out[e] = gg;
            if (!step) continue;
            const float mm = a.b1 * m[e] + (1.f - a.b1) * gg;
            const float vv = a.b2 * v[e] + (1.f - a.b2) * gg * gg;
            m[e] = mm;
            v[e] = vv;
            p[e] -= a.lr * (mm / c1) / (sqrtf(vv / c2) + a.eps);
        }
    }
    __syncthreads();
    if (step) {
        for (uint32_t t = threadIdx.x; t < 4 * 4 * 2 * 64 * 8; t += kFinishBlock) {
            const int s = t & 7, lane = (t >> 3) & 63, chunk = (t >> 9) & 1, tile = (t >> 10) & 3, layer = t >> 12;
            const int g = lane >> 4, i = lane & 15;
            const int in_dim = layer == 0 ? kGuideInputs : kGuideHidden;
            const float *W = p + layer * 64 * 64;
            const int kf = guide_perm(chunk, g, s);
            fwd[t] = (_Float16)(kf < in_dim ? W[kf * 64 + tile * 16 + i] : 0.f);
            const int kb = tile * 16 + i;
            bwd[t] = (_Float16)(kb < in_dim ? W[kb * 64 + kf] : 0.f);
        }
    }
    if (threadIdx.x == 0) {
        if (active && apply) {
            if (ctl->bad) {
                ctl->loss_scale = fmaxf(1.f, ctl->loss_scale * 0.5f);
                ctl->good = 0u;
                ctl->skipped += 1u;
            } else {
                ctl->step += 1u;
                if (++ctl->good >= kGuideScaleGrowth) {
                    ctl->loss_scale = fminf(kGuideScaleMax, ctl->loss_scale * 2.f);
                    ctl->good = 0u;
                }
            }
        }
        if (active) {
            ctl->last_loss = ctl->loss_cnt > 0.f ? ctl->loss_sum / ctl->loss_cnt : 0.f;
            ctl->last_count = ctl->batch_valid;
        }
        ctl->bad = 0u;
        ctl->active = 0u;
    }
}

// Stand-alone packing of the FP16 operands (parameter initialization and
// pt_guide_set_params).
__global__ void k_guide_pack(const float *w, _Float16 *fwd, _Float16 *bwd) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * 4 * 2 * 64 * 8) return;
    const int s = t & 7, lane = (t >> 3) & 63, chunk = (t >> 9) & 1, tile = (t >> 10) & 3, layer = t >> 12;
    const int g = lane >> 4, i = lane & 15;
    const int in_dim = layer == 0 ? kGuideInputs : kGuideHidden;
    const float *W = w + layer * 64 * 64;
    const int kf = guide_perm(chunk, g, s);
    fwd[t] = (_Float16)(kf < in_dim ? W[kf * 64 + tile * 16 + i] : 0.f);
    const int kb = tile * 16 + i;
    bwd[t] = (_Float16)(kb < in_dim ? W[kb * 64 + kf] : 0.f);
}

// Table initialization U(-1e-4, 1e-4), a counter-based draw per entry
// (SplitMix64 of seed and index), so tables of any size start on the device.
__global__ void k_guide_init_table(float2 *p, size_t n, uint64_t seed) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        uint64_t z = seed ^ ((uint64_t)e * 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const float a = (float)(uint32_t)(z >> 40) * (1.f / 16777216.f);
        const float b = (float)(uint32_t)(z & 0xFFFFFFu) * (1.f / 16777216.f);
        p[e] = make_float2((a * 2.f - 1.f) * 1e-4f, (b * 2.f - 1.f) * 1e-4f);
    }
}

__global__ void k_guide_fp16_table(const float2 *p, uint32_t *p16, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const float2 v = p[e];
        p16[e] = (uint32_t)hbits((_Float16)v.x) | ((uint32_t)hbits((_Float16)v.y) << 16);
    }
}

__global__ __launch_bounds__(256) void k_guide_query(GuideView G, const float *pos, uint32_t n, float *out,
                                                     uint32_t grid) {
    __shared__ float stage[4][64 * kGuideUsed];
    const int wave = threadIdx.x >> 6;
    const int lane = (int)__lane_id();
    const uint32_t base = (blockIdx.x * 4 + wave) * 64;
    if (base >= n) return;
    const uint32_t me = min(base + lane, n - 1);
    guide_infer_wave(G, mk3(pos[3 * me], pos[3 * me + 1], pos[3 * me + 2]), stage[wave], grid);
    wave_sync_lds();
    if (base + lane < n)
        for (int k = 0; k < kGuideUsed; ++k)
            out[(size_t)(base + lane) * kGuideUsed + k] = stage[wave][lane * kGuideUsed + k];
}

__global__ void k_guide_collect_begin(GuideCtl *ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->ring_start = ctl->ring_head;
        ctl->n_samples = 0u;
        ctl->n_nonzero = 0u;
        ctl->n_zero = 0u;
        ctl->fp_sum = 0.f;
        ctl->fp_cnt = 0.f;
    }
}

__global__ void k_guide_collect_end(GuideCtl *ctl, uint32_t capacity, uint32_t ring_n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t kept = min(ctl->n_samples, capacity);
        ctl->total_samples += kept;
        if (ring_n) ctl->ring_head = (uint32_t)(((uint64_t)ctl->ring_head + kept) % ring_n);
    }
}

// Ring collections: F/p of the samples just added is divided by their own
// collection's mean (the selection-loss normalization), so the samples of
// different launches -- views with very different brightness -- can share
// one optimizer step with the normalization 1.
__global__ void k_guide_ring_normalize(float4 *samples, const GuideCtl *ctl, uint32_t capacity, uint32_t ring_n) {
    const uint32_t kept = min(ctl->n_samples, capacity);
    const float norm = (ctl->fp_cnt > 0.f && ctl->fp_sum > 0.f) ? ctl->fp_sum / ctl->fp_cnt : 1.f;
    const float inv = 1.f / norm;
    const uint32_t start = ctl->ring_start;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < kept; k += gridDim.x * blockDim.x) {
        const size_t slot = ((size_t)start + k) % ring_n;
        samples[3 * slot + 2].y *= inv;
    }
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// Training sample of record idx, if the record was completed in this launch.
// Incident radiance along the sampled direction = (L_final - L_adj) / T_{k+1}
// per channel (L_adj already holds the unweighted emission of the next
// vertex); the radiance target is its luminance times |cos| (the density a
// diffuse vertex would ideally sample), and F/p = luminance of
// (f |cos| / p) * incident radiance.
__device__ __forceinline__ bool build_sample(const GuideRecords &r, uint32_t idx, uint32_t total, const float4 *out_l,
                                             float4 &s0, float4 &s1, float &target, float &fpv, float &pb) {
    const size_t n = r.capacity;
    const float4 r4 = r.rec[4 * n + idx];
    const uint32_t path = (idx / r.max_depth) * r.stride + r.phase;
    if (__float_as_uint(r4.w) != r.epoch || path >= total) return false;
    const float4 r0 = r.rec[idx], r1 = r.rec[n + idx], r2 = r.rec[2 * n + idx], r3 = r.rec[3 * n + idx];
    const float4 lf = out_l[path];
    const float ix = r2.x > 0.f ? fmaxf(lf.x - r4.x, 0.f) / r2.x : 0.f;
    const float iy = r2.y > 0.f ? fmaxf(lf.y - r4.y, 0.f) / r2.y : 0.f;
    const float iz = r2.z > 0.f ? fmaxf(lf.z - r4.z, 0.f) / r2.z : 0.f;
    target = luminance(mk3(ix, iy, iz)) * r1.w;
    fpv = luminance(mk3(r3.x * ix, r3.y * iy, r3.z * iz));
    if (!isfinite(target) || !isfinite(fpv)) return false;
    s0 = make_float4(r0.x, r0.y, r0.z, r2.w);  // r2.w: the vertex's hash grid
    s1 = make_float4(r1.x, r1.y, r1.z, 0.f);
    pb = r0.w;
    return true;
}

__global__ __launch_bounds__(256) void k_guide_collect_count(const float4 *out_l, uint32_t total, GuideRecords r,
                                                             GuideCtl *ctl) {
    uint32_t nz = 0, z = 0;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < r.capacity; idx += gridDim.x * blockDim.x) {
        float4 s0, s1;
        float target, fpv, pb;
        if (build_sample(r, idx, total, out_l, s0, s1, target, fpv, pb)) {
            if (target > 0.f || fpv > 0.f) ++nz;
            else ++z;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        nz += __shfl_xor(nz, off);
        z += __shfl_xor(z, off);
    }
    if (__lane_id() == 0) {
        if (nz) atomicAdd(&ctl->n_nonzero, nz);
        if (z) atomicAdd(&ctl->n_zero, z);
    }
}

// Turns the completed records of one launch into training samples.  Samples
// that carry a contribution are kept first (up to 3/4 of the capacity), the
// rest of the capacity is filled with zero-contribution samples; each class
// is a uniform subsample, and a kept sample's weight (aux.w) is 1 / its keep
// probability, so weighted batch means estimate means over all records.
// Kept sample k (k < capacity) goes to slot k, or with a ring of ring_n
// slots to slot (ring_head + k) mod ring_n.
__global__ __launch_bounds__(256) void k_guide_collect(const float4 *out_l, uint32_t total, GuideRecords r,
                                                       float4 *samples, uint32_t capacity, GuideCtl *ctl,
                                                       uint32_t ring_n) {
    const uint32_t head = ring_n ? ctl->ring_head : 0u;
    __shared__ uint32_t wcount[4], wbase[4];
    const uint32_t n = r.capacity;
    const uint32_t nnz = ctl->n_nonzero, nz = ctl->n_zero;
    const uint32_t cap_nz = capacity - capacity / 4u;
    const float keep_nz = nnz > cap_nz ? (float)cap_nz / (float)nnz : 1.f;
    const uint32_t kept_nz = min(nnz, cap_nz);
    const uint32_t cap_z = capacity - kept_nz;
    const float keep_z = nz > cap_z ? (float)cap_z / (float)nz : 1.f;
    float fs_acc = 0.f, fc_acc = 0.f;  // summed per lane, one atomic per wave at the end
    for (uint32_t b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
        const uint32_t idx = b + threadIdx.x;
        bool ok = false;
        float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0, s2 = s0;
        float fpv = 0.f;
        if (idx < n) {
            float target, pb;
            if (build_sample(r, idx, total, out_l, s0, s1, target, fpv, pb)) {
                const float keep = (target > 0.f || fpv > 0.f) ? keep_nz : keep_z;
                const float u = (float)(mix32(idx * 0x9E3779B1u ^ (r.epoch * 0x85EBCA77u)) >> 8) * (1.f / 16777216.f);
                if (u < keep) {
                    ok = true;
                    s2 = make_float4(target, fpv, pb, 1.f / keep);
                }
            }
        }
        // Slots for the block's kept samples: one global atomic per block
        // and iteration (wave counts combined in LDS).
        const unsigned long long mask = __ballot(ok);
        const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
        if (lane == 0) wcount[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t t = wcount[0] + wcount[1] + wcount[2] + wcount[3];
            const uint32_t b0 = t ? atomicAdd(&ctl->n_samples, t) : 0u;
            wbase[0] = b0;
            wbase[1] = b0 + wcount[0];
            wbase[2] = b0 + wcount[0] + wcount[1];
            wbase[3] = b0 + wcount[0] + wcount[1] + wcount[2];
        }
        __syncthreads();
        const unsigned long long below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane)));
        const uint32_t k = wbase[wave] + (uint32_t)__popcll(below);
        const bool kept = ok && k < capacity;
        fs_acc += kept ? fpv * s2.w : 0.f;
        fc_acc += (kept && fpv > 0.f) ? s2.w : 0.f;
        if (ok && k < capacity) {
            const size_t slot = ring_n ? ((size_t)head + k) % ring_n : (size_t)k;
            samples[3 * slot] = s0;
            samples[3 * slot + 1] = s1;
            samples[3 * slot + 2] = s2;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        fs_acc += __shfl_xor(fs_acc, off);
        fc_acc += __shfl_xor(fc_acc, off);
    }
    if (__lane_id() == 0 && fs_acc != 0.f) {
        unsafeAtomicAdd(&ctl->fp_sum, fs_acc);
        unsafeAtomicAdd(&ctl->fp_cnt, fc_acc);
    }
}

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform01(uint64_t &s) { return (float)((splitmix(s) >> 40) * (1.0 / 16777216.0)); }

template <typename T>
void gfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

namespace pupil {

#define GUIDE_HIP(call)                                              \
    do {                                                             \
        hipError_t e_ = (call);                                      \
        if (e_ != hipSuccess) {                                      \
            err = std::string(#call) + ": " + hipGetErrorString(e_); \
            return false;                                            \
        }                                                            \
    } while (0)

void GuideState::release() {
    gfree(table16);
    gfree(table32);
    gfree(table_m);
    gfree(table_v);
    gfree(table_g);
    gfree(grid_touched);
    gfree(w32);
    gfree(w_m);
    gfree(w_v);
    gfree(w_grad);
    gfree(fwd_a);
    gfree(bwd_a);
    gfree(ctl);
    gfree(xt);
    gfree(dyt);
    gfree(samples);
    gfree(rec);
    gfree(d_grid_bounds);
    batch_cap = sample_cap = sample_fill = rec_cap = 0;
    grids = 1;
    grid_bounds.clear();
    enabled = false;
    for_batch = false;
}

bool GuideState::init(const pt_guide_config &c, const float *lo, const float *hi, uint32_t n_grids, hipStream_t s,
                      std::string &err) {
    release();
    if (n_grids == 0) {
        err = "pt_guide_enable: no hash grids";
        return false;
    }
    stream = s;
    cfg = c;
    if (cfg.log2_table == 0) cfg.log2_table = 19;
    if (cfg.log2_table < 10 || cfg.log2_table > 24) {
        err = "pt_guide_enable: log2_table must be in [10, 24]";
        return false;
    }
    if (cfg.n_min == 0) cfg.n_min = 16;
    if (cfg.n_max == 0) cfg.n_max = 2048;
    if (cfg.n_max < cfg.n_min) {
        err = "pt_guide_enable: n_max below n_min";
        return false;
    }
    if (!(cfg.learning_rate > 0.f)) cfg.learning_rate = 1e-2f;
    if (cfg.batch_size == 0) cfg.batch_size = 16384;
    if (cfg.batch_size % 64u != 0u || cfg.batch_size > (1u << 22)) {
        err = "pt_guide_enable: batch_size must be a multiple of 64, at most 2^22";
        return false;
    }
    if (cfg.steps_per_frame == 0) cfg.steps_per_frame = 8;
    if (cfg.train_stride == 0) cfg.train_stride = 1;
    if (!(cfg.l2_weight > 0.f)) cfg.l2_weight = 1e-3f;
    // Level resolutions floor(n_min * b^l), b = (n_max / n_min)^(1 / 15), in double.
    const double b = std::exp((std::log((double)cfg.n_max) - std::log((double)cfg.n_min)) / (kGuideLevels - 1));
    const uint64_t T = 1ull << cfg.log2_table;
    for (int l = 0; l < kGuideLevels; ++l) {
        levels[l].res = (uint32_t)std::floor((double)cfg.n_min * std::pow(b, (double)l));
        const uint64_t side = levels[l].res + 1ull;
        levels[l].dense = side * side * side <= T ? 1u : 0u;
    }
    // Per grid: the scene box padded by 1e-3 of its extent (at least 1e-4).
    grid_bounds.assign((size_t)n_grids * 8, 0.f);
    for (uint32_t gi = 0; gi < n_grids; ++gi)
        for (int k = 0; k < 3; ++k) {
            const float l = lo[3 * gi + k], h = hi[3 * gi + k];
            if (!(h >= l) || !std::isfinite(l) || !std::isfinite(h)) {
                err = "pt_guide_enable: invalid scene bounds";
                return false;
            }
            const float ext = h - l;
            const float pad = std::max(ext * 1e-3f, 1e-4f);
            grid_bounds[8 * gi + k] = l - pad;
            grid_bounds[8 * gi + 4 + k] = 1.f / (ext + 2.f * pad);
        }
    for (int k = 0; k < 3; ++k) {
        bounds_lo[k] = grid_bounds[k];
        bounds_inv[k] = grid_bounds[4 + k];
    }
    grids = n_grids;
    grid_entries = (size_t)kGuideLevels * T;
    const size_t nt = grid_entries * n_grids;
    table_entries = nt;
    GUIDE_HIP(hipMalloc((void **)&d_grid_bounds, grid_bounds.size() * sizeof(float)));
    GUIDE_HIP(hipMemcpy(d_grid_bounds, grid_bounds.data(), grid_bounds.size() * sizeof(float), hipMemcpyHostToDevice));
    const size_t nw = 4 * 64 * 64;
    GUIDE_HIP(hipMalloc((void **)&table16, nt * sizeof(uint32_t)));
    GUIDE_HIP(hipMalloc((void **)&table32, nt * sizeof(float2)));
    GUIDE_HIP(hipMalloc((void **)&table_m, nt * sizeof(float2)));
    GUIDE_HIP(hipMalloc((void **)&table_v, nt * sizeof(float2)));
    GUIDE_HIP(hipMalloc((void **)&table_g, nt * sizeof(float2)));
    GUIDE_HIP(hipMalloc((void **)&grid_touched, n_grids * sizeof(uint32_t)));
    GUIDE_HIP(hipMalloc((void **)&w32, nw * sizeof(float)));
    GUIDE_HIP(hipMalloc((void **)&w_m, nw * sizeof(float)));
    GUIDE_HIP(hipMalloc((void **)&w_v, nw * sizeof(float)));
    GUIDE_HIP(hipMalloc((void **)&w_grad, nw * sizeof(float)));
    GUIDE_HIP(hipMalloc((void **)&fwd_a, 4 * 4 * 2 * 64 * 8 * sizeof(_Float16)));
    GUIDE_HIP(hipMalloc((void **)&bwd_a, 4 * 4 * 2 * 64 * 8 * sizeof(_Float16)));
    GUIDE_HIP(hipMalloc((void **)&ctl, sizeof(GuideCtl)));
    // Initialization: table U(-1e-4, 1e-4); weights U(+-sqrt(6 / (fan_in + fan_out)))
    // on the used part of each matrix, zero elsewhere.
    uint64_t st = (uint64_t)cfg.seed * 0x2545F4914F6CDD1Dull + 1;
    hipLaunchKernelGGL(k_guide_init_table, dim3(2048), dim3(256), 0, stream, table32, nt, splitmix(st));
    GUIDE_HIP(hipGetLastError());
    std::vector<float> hw(nw, 0.f);
    for (int l = 0; l < 4; ++l) {
        const int in = l == 0 ? kGuideInputs : kGuideHidden;
        const int out = l == 3 ? kGuideUsed : kGuideHidden;
        const float a = std::sqrt(6.f / (float)(in + out));
        for (int k = 0; k < in; ++k)
            for (int n = 0; n < out; ++n) hw[(size_t)l * 4096 + k * 64 + n] = (uniform01(st) * 2.f - 1.f) * a;
    }
    GUIDE_HIP(hipMemcpy(w32, hw.data(), nw * sizeof(float), hipMemcpyHostToDevice));
    GUIDE_HIP(hipMemset(table_m, 0, nt * sizeof(float2)));
    GUIDE_HIP(hipMemset(table_v, 0, nt * sizeof(float2)));
    GUIDE_HIP(hipMemset(table_g, 0, nt * sizeof(float2)));
    GUIDE_HIP(hipMemset(grid_touched, 0, n_grids * sizeof(uint32_t)));
    GUIDE_HIP(hipMemset(w_m, 0, nw * sizeof(float)));
    GUIDE_HIP(hipMemset(w_v, 0, nw * sizeof(float)));
    GUIDE_HIP(hipMemset(w_grad, 0, nw * sizeof(float)));
    GuideCtl hc;
    std::memset(&hc, 0, sizeof(hc));
    hc.loss_scale = kGuideScaleInit;
    hc.fp_norm = 1.f;
    GUIDE_HIP(hipMemcpy(ctl, &hc, sizeof(hc), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_guide_fp16_table, dim3(2048), dim3(256), 0, stream, table32, table16, nt);
    GUIDE_HIP(hipGetLastError());
    if (!pack(err)) return false;
    GUIDE_HIP(hipStreamSynchronize(stream));
    steps_launched = 0;
    epoch = 0;
    phase = 0;
    enabled = true;
    return true;
}

bool GuideState::pack(std::string &err) {
    hipLaunchKernelGGL(k_guide_pack, dim3(64), dim3(256), 0, stream, w32, fwd_a, bwd_a);
    GUIDE_HIP(hipGetLastError());
    return true;
}

GuideView GuideState::view() const {
    GuideView v;
    v.table = table16;
    v.fwd_a = reinterpret_cast<const half8 *>(fwd_a);
    v.bwd_a = reinterpret_cast<const half8 *>(bwd_a);
    for (int l = 0; l < kGuideLevels; ++l) v.levels[l] = levels[l];
    v.mask = (uint32_t)((1ull << cfg.log2_table) - 1ull);
    v.log2t = cfg.log2_table;
    v.lo = mk3(bounds_lo[0], bounds_lo[1], bounds_lo[2]);
    v.inv_ext = mk3(bounds_inv[0], bounds_inv[1], bounds_inv[2]);
    v.grids = grids;
    v.grid_bounds = reinterpret_cast<const float4 *>(d_grid_bounds);
    return v;
}

bool GuideState::ensure_batch(uint32_t batch, std::string &err) {
    if (batch <= batch_cap) return true;
    gfree(xt);
    gfree(dyt);
    const size_t n = (size_t)4 * 64 * batch;
    GUIDE_HIP(hipMalloc((void **)&xt, n * sizeof(uint16_t)));
    GUIDE_HIP(hipMalloc((void **)&dyt, n * sizeof(uint16_t)));
    GUIDE_HIP(hipMemset(xt, 0, n * sizeof(uint16_t)));
    GUIDE_HIP(hipMemset(dyt, 0, n * sizeof(uint16_t)));
    batch_cap = batch;
    return true;
}

bool GuideState::ensure_samples(uint32_t n, std::string &err) {
    sample_fill = n;
    if (n <= sample_cap) return true;
    gfree(samples);
    GUIDE_HIP(hipMalloc((void **)&samples, (size_t)n * 3 * sizeof(float4)));
    GUIDE_HIP(hipMemset(samples, 0, (size_t)n * 3 * sizeof(float4)));
    sample_cap = n;
    return true;
}

bool GuideState::ensure_records(uint32_t n, std::string &err) {
    if (n <= rec_cap) return true;
    gfree(rec);
    GUIDE_HIP(hipMalloc((void **)&rec, (size_t)n * 5 * sizeof(float4)));
    GUIDE_HIP(hipMemset(rec, 0, (size_t)n * 5 * sizeof(float4)));
    rec_cap = n;
    return true;
}

bool GuideState::launch_step(uint32_t first, uint32_t batch, float fp_norm, bool apply, uint32_t grid_lo,
                             uint32_t grid_hi, std::string &err, uint32_t ring_n) {
    const GuideView G = view();
    hipLaunchKernelGGL(k_guide_begin, dim3(1), dim3(64), 0, stream, ctl, first, batch, ring_n ? ring_n : sample_fill,
                       fp_norm, ring_n ? 1u : 0u);
    TrainBuffers B{xt, dyt, table_g, grid_touched};
    const uint32_t waves = batch / 16;
    hipLaunchKernelGGL(k_guide_train, dim3((waves + 3) / 4), dim3(kTrainBlock), 0, stream, G, samples, batch, ctl, B,
                       LossParams{cfg.loss_flags, cfg.l2_weight}, first, ring_n, ring_n ? ring_n / batch : 0u);
    hipLaunchKernelGGL(k_guide_dw, dim3((batch + kGuideDwSlice - 1) / kGuideDwSlice, 4), dim3(64), 0, stream, xt,
                       dyt, batch, ctl, w_grad);
    const AdamParams at{cfg.learning_rate, 0.9f, 0.99f, 1e-15f};
    const AdamParams aw{cfg.learning_rate, 0.9f, 0.99f, 1e-8f};
    // Sparse Adam over the grids the step's samples can touch.
    hipLaunchKernelGGL(k_guide_adam_table, dim3(2048), dim3(256), 0, stream, table32, table_m, table_v, table_g,
                       table16, grid_entries, grid_lo, grid_hi, (const uint32_t *)grid_touched, at, ctl,
                       apply ? 1u : 0u);
    hipLaunchKernelGGL(k_guide_finish, dim3(1), dim3(kFinishBlock), 0, stream, w32, w_m, w_v, w_grad, aw, ctl,
                       apply ? 1u : 0u, (float *)nullptr, fwd_a, bwd_a, grid_touched, grids);
    GUIDE_HIP(hipGetLastError());
    ++steps_launched;
    return true;
}

bool GuideState::collect(const float4 *out_l, uint32_t total_paths, const GuideRecords &r, std::string &err,
                         uint32_t ring_n, uint32_t ring_cap) {
    const uint32_t cap = ring_n ? std::min(ring_cap, ring_n) : sample_fill;
    if (ring_n > sample_cap) {
        err = "guide: sample ring exceeds the sample buffer";
        return false;
    }
    hipLaunchKernelGGL(k_guide_collect_begin, dim3(1), dim3(64), 0, stream, ctl);
    const uint32_t grid = std::max(1u, std::min<uint32_t>((r.capacity + 255) / 256, 4096u));
    hipLaunchKernelGGL(k_guide_collect_count, dim3(grid), dim3(256), 0, stream, out_l, total_paths, r, ctl);
    hipLaunchKernelGGL(k_guide_collect, dim3(grid), dim3(256), 0, stream, out_l, total_paths, r, samples, cap, ctl,
                       ring_n);
    if (ring_n)
        hipLaunchKernelGGL(k_guide_ring_normalize, dim3(std::max(1u, std::min((cap + 255u) / 256u, 2048u))),
                           dim3(256), 0, stream, samples, ctl, cap, ring_n);
    hipLaunchKernelGGL(k_guide_collect_end, dim3(1), dim3(64), 0, stream, ctl, cap, ring_n);
    GUIDE_HIP(hipGetLastError());
    return true;
}

bool GuideState::train_ring(uint32_t steps, float fp_norm, uint32_t first_step, uint32_t ring_n, std::string &err) {
    const uint32_t B = cfg.batch_size;
    if (!ensure_batch(B, err)) return false;
    if (ring_n < B || ring_n % B != 0 || ring_n > sample_cap) return true;
    const uint32_t m = ring_n / B;
    // Ring samples carry F/p already divided by their collection's mean.
    (void)fp_norm;
    for (uint32_t s = 0; s < steps; ++s)
        if (!launch_step((uint32_t)(((uint64_t)first_step + s) % m), B, 1.f, true, 0u, grids, err, ring_n))
            return false;
    return true;
}

bool GuideState::train(uint32_t steps, float fp_norm, std::string &err) { return train(steps, fp_norm, 0u, grids, err); }

bool GuideState::train(uint32_t steps, float fp_norm, uint32_t grid_lo, uint32_t grid_hi, std::string &err) {
    const uint32_t B = cfg.batch_size;
    if (!ensure_batch(B, err)) return false;
    if ((uint64_t)steps * B > sample_fill) steps = sample_fill / B;
    grid_hi = std::min(grid_hi, grids);
    if (grid_lo >= grid_hi) return true;
    for (uint32_t s = 0; s < steps; ++s)
        if (!launch_step(s * B, B, fp_norm, true, grid_lo, grid_hi, err)) return false;
    return true;
}

bool GuideState::upload_samples(const float *host, uint32_t n, std::string &err) {
    if (!ensure_samples(n, err)) return false;
    GUIDE_HIP(hipMemcpyAsync(samples, host, (size_t)n * 12 * sizeof(float), hipMemcpyHostToDevice, stream));
    GUIDE_HIP(hipMemcpyAsync(&ctl->n_samples, &n, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    GUIDE_HIP(hipStreamSynchronize(stream));
    return true;
}

bool GuideState::gradients(uint32_t n, float fp_norm, float *table_grad, float *weight_grad, float *loss,
                           std::string &err) {
    if (!ensure_batch(n, err)) return false;
    const GuideView G = view();
    hipLaunchKernelGGL(k_guide_begin, dim3(1), dim3(64), 0, stream, ctl, 0u, n, sample_fill, fp_norm, 0u);
    TrainBuffers B{xt, dyt, table_g, grid_touched};
    hipLaunchKernelGGL(k_guide_train, dim3((n / 16 + 3) / 4), dim3(kTrainBlock), 0, stream, G, samples, n, ctl, B,
                       LossParams{cfg.loss_flags, cfg.l2_weight}, 0u, 0u, 0u);
    hipLaunchKernelGGL(k_guide_dw, dim3((n + kGuideDwSlice - 1) / kGuideDwSlice, 4), dim3(64), 0, stream, xt, dyt,
                       n, ctl, w_grad);
    GUIDE_HIP(hipGetLastError());
    GUIDE_HIP(hipStreamSynchronize(stream));
    GuideCtl hc;
    GUIDE_HIP(hipMemcpy(&hc, ctl, sizeof(hc), hipMemcpyDeviceToHost));
    if (table_grad) {
        GUIDE_HIP(hipMemcpy(table_grad, table_g, table_entries * sizeof(float2), hipMemcpyDeviceToHost));
        const float inv = hc.loss_cnt > 0.f ? 1.f / hc.loss_cnt : 0.f;
        for (size_t k = 0; k < 2 * table_entries; ++k) table_grad[k] *= inv;
    }
    if (loss) *loss = hc.loss_cnt > 0.f ? hc.loss_sum / hc.loss_cnt : 0.f;
    // Report the weight gradient and close the step without applying it;
    // the table gradient is cleared, the parameters are untouched.
    float *d_wg = nullptr;
    GUIDE_HIP(hipMalloc((void **)&d_wg, 4 * 64 * 64 * sizeof(float)));
    const AdamParams aw{cfg.learning_rate, 0.9f, 0.99f, 1e-8f};
    hipLaunchKernelGGL(k_guide_finish, dim3(1), dim3(kFinishBlock), 0, stream, w32, w_m, w_v, w_grad, aw, ctl, 0u,
                       d_wg, fwd_a, bwd_a, grid_touched, grids);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemsetAsync(table_g, 0, table_entries * sizeof(float2), stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess && weight_grad)
        e = hipMemcpy(weight_grad, d_wg, 4 * 64 * 64 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d_wg);
    if (e != hipSuccess) {
        err = std::string("pt_guide_gradients: ") + hipGetErrorString(e);
        return false;
    }
    if (hc.bad) {
        err = "pt_guide_gradients: non-finite gradient";
        return false;
    }
    return true;
}

bool GuideState::get_params(float *table, float *weights, std::string &err) {
    GUIDE_HIP(hipStreamSynchronize(stream));
    if (table) GUIDE_HIP(hipMemcpy(table, table32, table_entries * sizeof(float2), hipMemcpyDeviceToHost));
    if (weights) GUIDE_HIP(hipMemcpy(weights, w32, 4 * 64 * 64 * sizeof(float), hipMemcpyDeviceToHost));
    return true;
}

bool GuideState::set_params(const float *table, const float *weights, std::string &err) {
    GUIDE_HIP(hipStreamSynchronize(stream));
    if (table) {
        GUIDE_HIP(hipMemcpy(table32, table, table_entries * sizeof(float2), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_guide_fp16_table, dim3(2048), dim3(256), 0, stream, table32, table16, table_entries);
    }
    if (weights) {
        std::vector<float> w(weights, weights + 4 * 64 * 64);
        // Keep the padding of the first layer (inputs 32..63) and of the
        // last layer (outputs 41..63) at zero.
        for (int k = kGuideInputs; k < 64; ++k)
            for (int n = 0; n < 64; ++n) w[k * 64 + n] = 0.f;
        for (int k = 0; k < 64; ++k)
            for (int n = kGuideUsed; n < 64; ++n) w[3 * 4096 + k * 64 + n] = 0.f;
        GUIDE_HIP(hipMemcpy(w32, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
        if (!pack(err)) return false;
    }
    GUIDE_HIP(hipGetLastError());
    GUIDE_HIP(hipStreamSynchronize(stream));
    return true;
}

bool GuideState::query(const float *pos, uint32_t n, float *out, uint32_t grid, std::string &err) {
    if (n == 0) return true;
    if (grid >= grids) {
        err = "pt_guide_query: no such hash grid";
        return false;
    }
    float *d_pos = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc((void **)&d_pos, (size_t)n * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&d_out, (size_t)n * kGuideUsed * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d_pos, pos, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_guide_query, dim3((n + 255) / 256), dim3(256), 0, stream, view(), d_pos, n, d_out, grid);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, (size_t)n * kGuideUsed * sizeof(float), hipMemcpyDeviceToHost);
    if (d_pos) (void)hipFree(d_pos);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) {
        err = std::string("pt_guide_query: ") + hipGetErrorString(e);
        return false;
    }
    return true;
}

bool GuideState::read_samples(float *out, uint32_t max_samples, uint32_t &count, std::string &err) {
    GUIDE_HIP(hipStreamSynchronize(stream));
    uint32_t n = 0;
    GUIDE_HIP(hipMemcpy(&n, &ctl->n_samples, sizeof(n), hipMemcpyDeviceToHost));
    count = std::min(std::min(n, sample_fill), max_samples);
    if (count) GUIDE_HIP(hipMemcpy(out, samples, (size_t)count * 3 * sizeof(float4), hipMemcpyDeviceToHost));
    return true;
}

bool GuideState::info(pt_guide_info &o, std::string &err) {
    std::memset(&o, 0, sizeof(o));
    o.enabled = enabled ? 1u : 0u;
    if (!enabled) return true;
    GUIDE_HIP(hipStreamSynchronize(stream));
    GuideCtl hc;
    GUIDE_HIP(hipMemcpy(&hc, ctl, sizeof(hc), hipMemcpyDeviceToHost));
    o.log2_table = cfg.log2_table;
    o.steps = hc.step;
    o.skipped = hc.skipped;
    o.loss_scale = hc.loss_scale;
    o.last_loss = hc.last_loss;
    o.last_samples = hc.last_count;
    for (int l = 0; l < kGuideLevels; ++l) {
        o.resolution[l] = levels[l].res;
        o.dense[l] = levels[l].dense;
    }
    for (int k = 0; k < 3; ++k) {
        o.bounds_lo[k] = bounds_lo[k];
        o.bounds_inv[k] = bounds_inv[k];
    }
    o.grids = grids;
    return true;
}

bool GuideState::grid_box(uint32_t grid, float lo[3], float inv[3], std::string &err) {
    if (grid >= grids) {
        err = "pt_guide_grid_bounds: no such hash grid";
        return false;
    }
    for (int k = 0; k < 3; ++k) {
        lo[k] = grid_bounds[8 * grid + k];
        inv[k] = grid_bounds[8 * grid + 4 + k];
    }
    return true;
}

}  // namespace pupil
