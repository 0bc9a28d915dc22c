"""Reference (float64 numpy) of the neural path-guiding proposal -- TEST
INFRASTRUCTURE ONLY.

The reference has no path guiding; this module defines the model the GPU
kernels implement, and is the checker for their forward pass, gradients and
Adam step:

* multiresolution hash-grid encoding of the hit position (Mueller et al. 2022,
  "Instant neural graphics primitives"): L levels, F features per level, T
  entries per level, resolutions N_l = floor(N_min * b^l), trilinear
  interpolation, dense indexing while (N_l + 1)^3 <= T; above that a
  spatial hash of 4x4x4-vertex bricks, h = (X * 1) ^ (Y * 2654435761) ^
  (Z * 805459861) of the brick (X, Y, Z) = (x, y, z) >> 2, entry
  ((h << 6) | local) mod T with local = (x & 3) | (y & 3) << 2 | (z & 3) << 4,
  so the 8 corners of a cell mostly share one 64-entry (256-byte) brick;
* MLP 32 -> 64 -> 64 -> 64 -> 64 (41 outputs used), ReLU, no biases;
* outputs -> 8 von Mises-Fisher lobes (mean = normalize(o[5k:5k+3]),
  kappa = exp(clamp(o[5k+3])), weight = exp(clamp(o[5k+4]))) and a selection
  probability c = 0.95 * sigmoid(o[40] - 2) (about 0.11 for a fresh network,
  so early frames sample mostly from the BSDF);
* radiance model Lhat(w) = sum_k weight_k vMF(w; mean_k, kappa_k); guiding pdf
  q = Lhat / sum_k weight_k;
* the radiance model factorizes as Lhat = W q, W = sum_k weight_k (the
  magnitude) and q the normalized guide density; loss per sample, with
  w = min(F / p_s / norm, 10) and F = f L |cos| of the sample:
  - relative L2 of the magnitude, l2 (W sg(q) - L)^2 / (sg(Lhat)^2 + 0.01), on
    the target L (incident radiance times |cos| at the vertex), l2 = L2_WEIGHT
    by default;
  - the cross-entropy of the guide, -w log q (shapes the lobes toward the
    contribution-weighted sample distribution, as in neural importance
    sampling);
  - the cross-entropy of the one-sample mixture, -w log(c sg(q) + (1 - c) p_bsdf)
    (trains the selection probability c);
  (the relative L2 loss fit to the lobes directly collapses them onto single
  noisy samples, so it reaches the network only through the magnitude W);
* per-sample output gradients clipped in max norm (direction kept): the
  relative-L2 part to L2_CLIP (so single bright samples cannot swamp the
  shared network), the cross-entropy part to GRAD_CLIP; the batch mean is
  weighted by the sample weights;
* Adam (beta1 0.9, beta2 0.99, eps 1e-15 on the hash table / 1e-8 on weights);
  hash-table entries with a zero gradient in a step are skipped (sparse Adam).
"""
from __future__ import annotations

import numpy as np

LEVELS, FEATURES, HIDDEN, OUT_USED, OUT_PAD, LOBES = 16, 2, 64, 41, 64, 8
PRIMES = (1, 2654435761, 805459861)
KAPPA_MIN_LOG, KAPPA_MAX_LOG = -5.0, 7.0
WEIGHT_MIN_LOG, WEIGHT_MAX_LOG = -15.0, 15.0
C_MAX = 0.95
C_BIAS = -2.0
REL_EPS = 0.01
GRAD_CLIP = 100.0
L2_CLIP = 1.0
L2_WEIGHT = 0.001


def fp16(x):
    """Rounds to the nearest IEEE half (the storage type of the GPU table,
    weights, features and hidden activations)."""
    return np.asarray(x, np.float64).astype(np.float16).astype(np.float64)


def resolutions(levels=LEVELS, n_min=16, n_max=2048):
    b = np.exp((np.log(n_max) - np.log(n_min)) / (levels - 1))
    return [int(np.floor(n_min * b ** l)) for l in range(levels)]


def corner_index(level_res, log2_t, cx, cy, cz):
    t = 1 << log2_t
    side = level_res + 1
    if side ** 3 <= t:
        return (cx + cy * side + cz * side * side) % t
    h = ((cx >> 2) * PRIMES[0]) ^ (((cy >> 2) * PRIMES[1]) & 0xFFFFFFFF) ^ (((cz >> 2) * PRIMES[2]) & 0xFFFFFFFF)
    local = (cx & 3) | ((cy & 3) << 2) | ((cz & 3) << 4)
    return (((h << 6) & 0xFFFFFFFF) | local) & (t - 1)


def corner_indices(level_res, log2_t, cx, cy, cz):
    """corner_index over arrays of integer corner coordinates."""
    t = 1 << log2_t
    side = level_res + 1
    cx, cy, cz = (np.asarray(a, np.int64) for a in (cx, cy, cz))
    if side ** 3 <= t:
        return (cx + cy * side + cz * side * side) % t
    m32 = np.uint64(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        h = ((cx >> 2).astype(np.uint64) * np.uint64(PRIMES[0]) & m32) ^ \
            ((cy >> 2).astype(np.uint64) * np.uint64(PRIMES[1]) & m32) ^ \
            ((cz >> 2).astype(np.uint64) * np.uint64(PRIMES[2]) & m32)
    local = ((cx & 3) | ((cy & 3) << 2) | ((cz & 3) << 4)).astype(np.uint64)
    return ((((h << np.uint64(6)) & m32) | local) & np.uint64(t - 1)).astype(np.int64)


def encode(table, u, res, log2_t, quantize=False):
    """table: [L, T, F]; u: [N, 3] in [0, 1].  Returns features [N, L*F] and the
    (indices, weights) needed by the backward pass.  quantize: round the
    interpolated features to half precision, as the GPU feeds them to MFMA."""
    n = u.shape[0]
    feats = np.zeros((n, LEVELS * FEATURES))
    cache = []
    for l, r in enumerate(res):
        x = np.clip(u, 0.0, 1.0) * r
        c0 = np.floor(x).astype(np.int64)
        c0 = np.minimum(c0, r - 1) if r > 0 else c0
        fr = x - c0
        idx = np.zeros((n, 8), np.int64)
        w = np.zeros((n, 8))
        for k in range(8):
            dx, dy, dz = k & 1, (k >> 1) & 1, (k >> 2) & 1
            idx[:, k] = corner_indices(r, log2_t, c0[:, 0] + dx, c0[:, 1] + dy, c0[:, 2] + dz)
            w[:, k] = (fr[:, 0] if dx else 1 - fr[:, 0]) * (fr[:, 1] if dy else 1 - fr[:, 1]) * \
                (fr[:, 2] if dz else 1 - fr[:, 2])
        feats[:, l * FEATURES:(l + 1) * FEATURES] = np.einsum("nk,nkf->nf", w, table[l][idx])
        cache.append((idx, w))
    if quantize:
        feats = fp16(feats)
    return feats, cache


def mlp_forward(x, weights, quantize=False):
    """weights: list of [in, out] matrices.  Returns outputs and activations.
    quantize: round hidden activations to half precision (the GPU keeps them
    as FP16 MFMA operands; the output layer stays FP32)."""
    acts = [x]
    h = x
    for i, w in enumerate(weights):
        z = h @ w
        if i < len(weights) - 1:
            h = np.maximum(z, 0.0)
            if quantize:
                h = fp16(h)
        else:
            h = z
        acts.append(h)
    return h, acts


def mixture(o):
    """o: [N, 64] -> dict of lobe parameters and c (all [N, ...])."""
    raw = np.stack([o[:, 5 * k:5 * k + 3] for k in range(LOBES)], 1)  # [N, 8, 3]
    norm = np.linalg.norm(raw, axis=-1, keepdims=True)
    mu = np.where(norm > 1e-12, raw / np.maximum(norm, 1e-30), np.array([0, 0, 1.0]))
    lk = np.stack([o[:, 5 * k + 3] for k in range(LOBES)], 1)
    lw = np.stack([o[:, 5 * k + 4] for k in range(LOBES)], 1)
    kappa = np.exp(np.clip(lk, KAPPA_MIN_LOG, KAPPA_MAX_LOG))
    weight = np.exp(np.clip(lw, WEIGHT_MIN_LOG, WEIGHT_MAX_LOG))
    c = C_MAX / (1.0 + np.exp(-(o[:, 40] + C_BIAS)))
    return {"raw": raw, "norm": norm, "mu": mu, "lk": lk, "lw": lw, "kappa": kappa, "weight": weight, "c": c}


def vmf(kappa, cosang):
    """vMF density on the sphere, stable for small and large kappa."""
    return kappa / (2 * np.pi * (-np.expm1(-2 * kappa))) * np.exp(kappa * (cosang - 1.0))


def radiance_and_pdf(m, w):
    cosang = np.einsum("nkd,nd->nk", m["mu"], w)
    v = vmf(m["kappa"], cosang)
    lhat = np.sum(m["weight"] * v, axis=1)
    q = lhat / np.sum(m["weight"], axis=1)
    return lhat, q, cosang, v


def vmf_log(kappa, cosang):
    """log of the vMF density."""
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (cosang - 1.0)


def loss_and_output_grad(o, w, target, fs_over_p, p_bsdf, norm_fp=1.0, lam_c=1.0, clip=GRAD_CLIP, l2=L2_WEIGHT):
    """Per-sample loss and dLoss/do (the mean over the batch is taken by the
    caller).  target: radiance target of the sample; fs_over_p: F / p_s
    (luminance of f L |cos| / sampling pdf); p_bsdf: BSDF pdf of w.
    Each row of the gradient is scaled down to max-norm `clip` if larger."""
    m = mixture(o)
    lhat, q, cosang, v = radiance_and_pdf(m, w)
    wsum = np.sum(m["weight"], axis=1)
    den = lhat ** 2 + REL_EPS
    loss = l2 * (lhat - target) ** 2 / den
    dl = 2 * (lhat - target) / den  # d(relative L2)/dLhat (denominator held fixed), scaled by l2 below
    wgt = lam_c * np.minimum(fs_over_p / norm_fp, 10.0)
    # lobe responsibilities r_k = w_k v_k / Lhat, from log-domain terms
    lt = np.log(m["weight"]) + vmf_log(m["kappa"], cosang)
    lt = lt - lt.max(axis=1, keepdims=True)
    r = np.exp(lt)
    r /= r.sum(axis=1, keepdims=True)
    c = m["c"]
    pmix = np.maximum(c * q + (1 - c) * p_bsdf, 1e-30)
    loss = loss - wgt * np.log(np.maximum(q, 1e-30)) - wgt * np.log(pmix)
    dc = -wgt * (q - p_bsdf) / pmix
    g = np.zeros_like(o)
    g2 = np.zeros_like(o)  # relative-L2 part
    for k in range(LOBES):
        wk, kk, vk, ck = m["weight"][:, k], m["kappa"][:, k], v[:, k], cosang[:, k]
        # relative L2 through W only: d(W q0)/dlw_k = w_k q0
        # d(-log q)/d(log-term of lobe k) = -(r_k); the weight also enters 1/wsum
        b = -wgt * r[:, k]
        inside_w = (m["lw"][:, k] > WEIGHT_MIN_LOG) & (m["lw"][:, k] < WEIGHT_MAX_LOG)
        g2[:, 5 * k + 4] = dl * wk * q * inside_w
        g[:, 5 * k + 4] = (b + wgt * wk / wsum) * inside_w
        # kappa: d log v / d kappa = 1/kappa - 2/(e^{2 kappa} - 1) + cos - 1
        with np.errstate(over="ignore"):
            dlogv = 1.0 / kk - 2.0 / np.expm1(2 * kk) + ck - 1.0
        inside_k = (m["lk"][:, k] > KAPPA_MIN_LOG) & (m["lk"][:, k] < KAPPA_MAX_LOG)
        g[:, 5 * k + 3] = b * dlogv * kk * inside_k
        # mean: d cos / d raw = (w - cos * mu) / |raw|
        mu = m["mu"][:, k]
        nr = np.maximum(m["norm"][:, k, 0], 1e-30)
        dr = (w - ck[:, None] * mu) / nr[:, None]
        g[:, 5 * k:5 * k + 3] = ((b * kk)[:, None] * dr) * (m["norm"][:, k, 0] > 1e-12)[:, None]
    sig = c / C_MAX
    g[:, 40] = dc * C_MAX * sig * (1 - sig)
    if clip is not None:
        for part, bound in ((g, clip), (g2, L2_CLIP)):
            mx = np.max(np.abs(part), axis=1, keepdims=True)
            part *= np.where(mx > bound, bound / np.maximum(mx, 1e-300), 1.0)
    return loss, g + l2 * g2


def backward(acts, weights, g_out):
    """Returns dW list and d(input features)."""
    dws = [None] * len(weights)
    d = g_out
    for i in reversed(range(len(weights))):
        dws[i] = acts[i].T @ d
        d = d @ weights[i].T
        if i > 0:
            d = d * (acts[i] > 0)
    return dws, d


def encode_backward(table_shape, cache, dfeat):
    g = np.zeros(table_shape)
    for l, (idx, w) in enumerate(cache):
        df = dfeat[:, l * FEATURES:(l + 1) * FEATURES]
        for k in range(8):
            np.add.at(g[l], idx[:, k], w[:, k:k + 1] * df)
    return g


class Adam:
    """Adam; sparse=True skips entries whose gradient is exactly zero (their
    moments and values stay as they are), as for hash-table entries no sample
    touched in the step."""

    def __init__(self, lr=1e-2, b1=0.9, b2=0.99, eps=1e-15, sparse=False):
        self.lr, self.b1, self.b2, self.eps, self.sparse = lr, b1, b2, eps, sparse

    def step(self, p, g, m, v, t):
        sel = (g != 0) if self.sparse else np.ones(g.shape, bool)
        m[sel] = self.b1 * m[sel] + (1 - self.b1) * g[sel]
        v[sel] = self.b2 * v[sel] + (1 - self.b2) * g[sel] * g[sel]
        mh = m[sel] / (1 - self.b1 ** t)
        vh = v[sel] / (1 - self.b2 ** t)
        p[sel] = p[sel] - self.lr * mh / (np.sqrt(vh) + self.eps)


def full_loss(table, weights, u, w, target, fs_over_p, p_bsdf, res, log2_t, norm_fp=1.0, clip=GRAD_CLIP,
              valid=None, quantize=False, l2=L2_WEIGHT):
    """valid: per-sample weights (or booleans); the loss is the weighted mean
    and g holds the weighted per-sample gradients."""
    feats, cache = encode(table, u, res, log2_t, quantize)
    o, acts = mlp_forward(feats, weights, quantize)
    loss, g = loss_and_output_grad(o, w, target, fs_over_p, p_bsdf, norm_fp, clip=clip, l2=l2)
    wts = np.ones(len(loss)) if valid is None else np.asarray(valid, np.float64)
    g = g * wts[:, None]
    m = wts > 0
    return float(np.sum(loss[m] * wts[m]) / np.sum(wts[m])), (feats, cache, o, acts, g)


def gradients(table, weights, u, w, target, fs_over_p, p_bsdf, res, log2_t, norm_fp=1.0, clip=GRAD_CLIP,
              valid=None, quantize=False, l2=L2_WEIGHT):
    """Gradients of the weighted mean loss (valid: per-sample weights, 0 for
    unused samples; q inside the mixture term and the relative-L2
    denominator held fixed, as in training).  quantize: evaluate with
    half-precision parameters, features and activations as the GPU does."""
    if quantize:
        table = fp16(table)
        weights = [fp16(x) for x in weights]
    loss, (feats, cache, o, acts, g) = full_loss(table, weights, u, w, target, fs_over_p, p_bsdf, res, log2_t,
                                                 norm_fp, clip, valid, quantize, l2)
    n = u.shape[0] if valid is None else float(np.sum(np.asarray(valid, np.float64)))
    dws, dfeat = backward(acts, weights, g / n)
    dtable = encode_backward(table.shape, cache, dfeat)
    return dws, dtable, o, loss


def sample_vmf(mu, kappa, u1, u2):
    """Jakob's stable vMF sampler around mu."""
    wz = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa
    wz = np.clip(wz, -1.0, 1.0)
    r = np.sqrt(np.maximum(0.0, 1.0 - wz * wz))
    phi = 2 * np.pi * u2
    # orthonormal basis around mu (Duff et al.)
    sign = np.copysign(1.0, mu[2])
    a = -1.0 / (sign + mu[2])
    b = mu[0] * mu[1] * a
    b1 = np.array([1.0 + sign * mu[0] * mu[0] * a, sign * b, -sign * mu[0]])
    b2 = np.array([b, sign + mu[1] * mu[1] * a, -mu[1]])
    return r * np.cos(phi) * b1 + r * np.sin(phi) * b2 + wz * mu
