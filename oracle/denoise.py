"""Float64 restatement of the denoising filter -- TEST INFRASTRUCTURE ONLY.

Used by tests/ as the checker of the GPU denoiser (pupiloptixlab_amd/csrc/
pt_denoise.hip).  The reference's own denoiser is the closed OptiX AI model
(optix/denoiser.cpp), so nothing here restates reference code: this is the
edge-avoiding a-trous wavelet transform (Dammertz et al., HPG 2010) with the
variance-guided luminance weight of SVGF (Schied et al., HPG 2017), as the
GPU implements it:

  prepare: c = color / max(albedo, 1e-3); variance = guide-weighted 5x5
           luminance variance;
  level k: 5x5 B3-spline taps at stride 2^k, weight h_i h_j w_l w_n w_a,
           w_l = exp(-|l_p - l_q| / (sigma_l sqrt(g3x3(var_p)) + 1e-4)),
           w_n = max(0, n_p.n_q)^sigma_n (zero normals only match zero normals),
           w_a = exp(-|a_p - a_q|^2 / sigma_a^2);
           color' = sum w c / sum w, var' = sum w^2 var / (sum w)^2;
  finish:  color x max(albedo, 1e-3).

Pinned by closed forms in tests/test_denoise.py: constant images are fixed
points, albedo steps and normal edges do not bleed, flat noise shrinks.
"""
import numpy as np

EPS = 1e-3
KERN = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])
LUM = np.array([0.2126, 0.7152, 0.0722])


def _shift(a, dy, dx):
    """a shifted so that out[y, x] = a[y + dy, x + dx]; invalid where outside."""
    h, w = a.shape[:2]
    out = np.zeros_like(a)
    valid = np.zeros((h, w), bool)
    ys, yd = (slice(dy, h), slice(0, h - dy)) if dy >= 0 else (slice(0, h + dy), slice(-dy, h))
    xs, xd = (slice(dx, w), slice(0, w - dx)) if dx >= 0 else (slice(0, w + dx), slice(-dx, w))
    if ys.start < ys.stop and xs.start < xs.stop:
        out[yd, xd] = a[ys, xs]
        valid[yd, xd] = True
    return out, valid


def _guide(albedo, normal, dy, dx, sn, sa):
    h, w = (albedo if albedo is not None else normal).shape[:2]
    wgt = np.ones((h, w))
    if normal is not None:
        nq, _ = _shift(normal, dy, dx)
        zp = (normal ** 2).sum(-1) == 0
        zq = (nq ** 2).sum(-1) == 0
        d = np.maximum(0.0, (normal * nq).sum(-1))
        wn = np.where(zp | zq, np.where(zp == zq, 1.0, 0.0), d ** sn)
        wgt = wgt * wn
    if albedo is not None:
        aq, _ = _shift(albedo, dy, dx)
        wgt = wgt * np.exp(-((albedo - aq) ** 2).sum(-1) / (sa * sa))
    return wgt


def denoise(rgba, albedo=None, normal=None, iterations=5, sigma_luminance=4.0, sigma_normal=64.0, sigma_albedo=0.1):
    c = np.asarray(rgba, np.float64)[..., :3]
    a = None if albedo is None else np.asarray(albedo, np.float64)
    n = None if normal is None else np.asarray(normal, np.float64)
    h, w = c.shape[:2]
    if a is not None:
        c = c / np.maximum(a, EPS)
    ones = np.ones((h, w))
    gw = (lambda dy, dx: _guide(a, n, dy, dx, sigma_normal, sigma_albedo)) if (a is not None or n is not None) \
        else (lambda dy, dx: ones)
    lum = c @ LUM
    s1 = np.zeros((h, w))
    s2 = np.zeros((h, w))
    sw = np.zeros((h, w))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            lq, valid = _shift(lum, dy, dx)
            wq = gw(dy, dx) * valid
            s1 += wq * lq
            s2 += wq * lq * lq
            sw += wq
    var = np.maximum(0.0, s2 / sw - (s1 / sw) ** 2)
    for k in range(iterations):
        step = 1 << k
        gv = np.zeros((h, w))
        gws = np.zeros((h, w))
        for dy in range(-1, 2):
            for dx in range(-1, 2):
                vq, valid = _shift(var, dy, dx)
                kk = (0.5 if dx == 0 else 0.25) * (0.5 if dy == 0 else 0.25)
                gv += kk * vq * valid
                gws += kk * valid
        sigma = sigma_luminance * np.sqrt(gv / gws) + 1e-4
        l0 = c @ LUM
        sc = np.zeros_like(c)
        sv = np.zeros((h, w))
        sw = np.zeros((h, w))
        for j in range(-2, 3):
            for i in range(-2, 3):
                cq, valid = _shift(c, j * step, i * step)
                vq, _ = _shift(var, j * step, i * step)
                wl = np.exp(-np.abs(l0 - cq @ LUM) / sigma)
                wq = KERN[i + 2] * KERN[j + 2] * wl * gw(j * step, i * step) * valid
                sc += wq[..., None] * cq
                sv += wq * wq * vq
                sw += wq
        c = sc / sw[..., None]
        var = sv / (sw * sw)
    if a is not None:
        c = c * np.maximum(a, EPS)
    return np.concatenate([c, np.ones((h, w, 1))], -1)
