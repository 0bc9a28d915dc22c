"""Pins the float64 guiding reference (oracle/guide.py): vMF normalization and
sampling, finite-difference gradients of every stage, Adam."""
import numpy as np
import pytest

from oracle import guide as G


def fib_sphere(n):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    theta = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], -1)


@pytest.mark.parametrize("kappa", [1e-4, 0.5, 5.0, 80.0])
def test_vmf_integrates_to_one(kappa):
    d = fib_sphere(400000)
    mu = np.array([0.3, -0.4, np.sqrt(1 - 0.25)])
    v = G.vmf(np.full(len(d), kappa), d @ mu)
    assert v.mean() * 4 * np.pi == pytest.approx(1.0, rel=2e-3)


def test_mixture_pdf_normalized_and_sampler_matches():
    rng = np.random.default_rng(0)
    o = rng.normal(size=(1, 64))
    m = G.mixture(o)
    d = fib_sphere(200000)
    lhat, q, _, _ = G.radiance_and_pdf({k: np.repeat(v, len(d), 0) for k, v in m.items()}, d)
    assert q.mean() * 4 * np.pi == pytest.approx(1.0, rel=3e-3)
    # sample one lobe, compare the cos(angle to mean) histogram with the vMF law
    mu, kappa = m["mu"][0, 0], 3.0
    s = np.array([G.sample_vmf(mu, kappa, a, b) for a, b in rng.random((40000, 2))])
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-9)
    c = s @ mu
    # cos theta of a vMF sample has density kappa e^{kappa (t - 1)} / (1 - e^{-2 kappa}) on [-1, 1]
    hist, edges = np.histogram(c, bins=20, range=(-1, 1))
    cdf = lambda t: (np.exp(kappa * (t - 1)) - np.exp(-2 * kappa)) / (1 - np.exp(-2 * kappa))
    exp = (cdf(edges[1:]) - cdf(edges[:-1])) * len(c)
    m_ = exp > 20
    assert np.sum((hist[m_] - exp[m_]) ** 2 / exp[m_]) < 3 * m_.sum() + 20


def setup_small(seed=1, n=24, log2_t=8):
    rng = np.random.default_rng(seed)
    res = G.resolutions(G.LEVELS, 2, 64)
    table = rng.normal(0, 0.3, (G.LEVELS, 1 << log2_t, G.FEATURES))
    dims = [32, 64, 64, 64, 64]
    weights = [rng.normal(0, 1.0 / np.sqrt(dims[i]), (dims[i], dims[i + 1])) for i in range(4)]
    u = rng.random((n, 3))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    target = rng.gamma(1.0, 1.0, n)
    fs = rng.gamma(1.0, 1.0, n)
    pb = rng.uniform(0.05, 1.0, n)
    return table, weights, u, w, target, fs, pb, res, log2_t


def surrogate(table, weights, u, w, target, fs, pb, res, log2_t, frozen):
    """Batch-mean loss with the relative-L2 denominator and q frozen at `frozen`."""
    feats, _ = G.encode(table, u, res, log2_t)
    o, _ = G.mlp_forward(feats, weights)
    m = G.mixture(o)
    lhat, q, _, _ = G.radiance_and_pdf(m, w)
    den, q0 = frozen
    loss = (lhat - target) ** 2 / den
    c = m["c"]
    pmix = c * q0 + (1 - c) * pb
    loss = loss - np.minimum(fs, 10.0) * np.log(pmix)
    return float(loss.mean())


def test_output_gradient_matches_finite_differences():
    table, weights, u, w, target, fs, pb, res, log2_t = setup_small()
    feats, _ = G.encode(table, u, res, log2_t)
    o, _ = G.mlp_forward(feats, weights)
    _, g = G.loss_and_output_grad(o, w, target, fs, pb, clip=None)
    m = G.mixture(o)
    lhat, q, _, _ = G.radiance_and_pdf(m, w)
    frozen_den, frozen_q = lhat ** 2 + G.REL_EPS, q

    def per_sample(oo):
        mm = G.mixture(oo)
        lh, _, _, _ = G.radiance_and_pdf(mm, w)
        c = mm["c"]
        return (lh - target) ** 2 / frozen_den - np.minimum(fs, 10.0) * np.log(c * frozen_q + (1 - c) * pb)

    eps = 1e-6
    for j in list(range(0, 41, 3)) + [40]:
        op, om = o.copy(), o.copy()
        op[:, j] += eps
        om[:, j] -= eps
        fd = (per_sample(op) - per_sample(om)) / (2 * eps)
        assert np.allclose(g[:, j], fd, rtol=1e-4, atol=1e-7), j


def test_network_and_table_gradients_match_finite_differences():
    table, weights, u, w, target, fs, pb, res, log2_t = setup_small(n=12)
    dws, dtable, o, _ = G.gradients(table, weights, u, w, target, fs, pb, res, log2_t, clip=None)
    m = G.mixture(o)
    lhat, q, _, _ = G.radiance_and_pdf(m, w)
    frozen = (lhat ** 2 + G.REL_EPS, q)
    f = lambda tb, ws: surrogate(tb, ws, u, w, target, fs, pb, res, log2_t, frozen)
    rng = np.random.default_rng(9)
    eps = 1e-6
    for layer in range(4):
        for _ in range(6):
            a, b = rng.integers(0, weights[layer].shape[0]), rng.integers(0, weights[layer].shape[1])
            wp = [x.copy() for x in weights]
            wm = [x.copy() for x in weights]
            wp[layer][a, b] += eps
            wm[layer][a, b] -= eps
            fd = (f(table, wp) - f(table, wm)) / (2 * eps)
            assert dws[layer][a, b] == pytest.approx(fd, rel=1e-4, abs=1e-8), (layer, a, b)
    nz = np.argwhere(np.abs(dtable) > 1e-6)
    for l, e, k in nz[rng.choice(len(nz), 12, replace=False)]:
        tp, tm = table.copy(), table.copy()
        tp[l, e, k] += eps
        tm[l, e, k] -= eps
        fd = (f(tp, weights) - f(tm, weights)) / (2 * eps)
        assert dtable[l, e, k] == pytest.approx(fd, rel=1e-4, abs=1e-8)


def test_hash_index_known_values():
    # dense while (res+1)^3 <= T, hashed above; the hash is (x*1)^(y*2654435761)^(z*805459861) mod T
    assert G.corner_index(3, 8, 1, 2, 3) == 1 + 2 * 4 + 3 * 16
    x, y, z = 5, 7, 11
    h = (x * 1) ^ ((y * 2654435761) & 0xFFFFFFFF) ^ ((z * 805459861) & 0xFFFFFFFF)
    assert G.corner_index(100, 10, x, y, z) == h & 1023


def test_adam_matches_closed_form_first_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    m, v = np.zeros(2), np.zeros(2)
    G.Adam(lr=0.01, eps=0.0).step(p, g, m, v, 1)
    assert np.allclose(p, [1.0 - 0.01, -2.0 + 0.01])  # first Adam step moves lr * sign(g)


def test_gradient_clip_keeps_direction_and_bounds_max_norm():
    table, weights, u, w, target, fs, pb, res, log2_t = setup_small(n=32)
    feats, _ = G.encode(table, u, res, log2_t)
    o, _ = G.mlp_forward(feats, weights)
    _, g_raw = G.loss_and_output_grad(o, w, target * 1e4, fs, pb, clip=None)
    _, g_clip = G.loss_and_output_grad(o, w, target * 1e4, fs, pb, clip=5.0)
    mx = np.abs(g_raw).max(axis=1)
    assert (mx > 5.0).any()
    assert np.all(np.abs(g_clip).max(axis=1) <= 5.0 + 1e-9)
    for r in range(len(mx)):
        if mx[r] > 5.0:
            assert np.allclose(g_clip[r], g_raw[r] * (5.0 / mx[r]))
        else:
            assert np.array_equal(g_clip[r], g_raw[r])


def test_invalid_samples_do_not_contribute():
    table, weights, u, w, target, fs, pb, res, log2_t = setup_small(n=16)
    valid = np.arange(16) % 3 != 0
    dws, dt, _, loss = G.gradients(table, weights, u, w, target, fs, pb, res, log2_t, valid=valid)
    keep = np.flatnonzero(valid)
    dws2, dt2, _, loss2 = G.gradients(table, weights, u[keep], w[keep], target[keep], fs[keep], pb[keep], res,
                                      log2_t)
    assert loss == pytest.approx(loss2)
    for a, b in zip(dws, dws2):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-15)
    assert np.allclose(dt, dt2, rtol=1e-12, atol=1e-15)
