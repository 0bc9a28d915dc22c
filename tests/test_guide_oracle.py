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


def surrogate(table, weights, u, w, target, fs, pb, res, log2_t, frozen, frozen_q=None):
    """Batch-mean loss with the relative-L2 denominator frozen at `frozen`
    and q frozen at `frozen_q` inside the magnitude and mixture terms."""
    feats, _ = G.encode(table, u, res, log2_t)
    o, _ = G.mlp_forward(feats, weights)
    m = G.mixture(o)
    lhat, q, _, _ = G.radiance_and_pdf(m, w)
    qf = frozen_q if frozen_q is not None else q
    wsum = m["weight"].sum(axis=1)
    loss = (wsum * qf - target) ** 2 / frozen
    c = m["c"]
    pmix = c * qf + (1 - c) * pb
    loss = loss - np.minimum(fs, 10.0) * (np.log(q) + np.log(pmix))
    return float(loss.mean())


def test_selection_probability_starts_small():
    """A fresh network (outputs ~ 0) samples the guide with c = 0.95 sigmoid(-2)."""
    m = G.mixture(np.zeros((1, 64)))
    assert m["c"][0] == pytest.approx(0.95 / (1 + np.exp(2.0)))


def test_output_gradient_matches_finite_differences():
    table, weights, u, w, target, fs, pb, res, log2_t = setup_small()
    feats, _ = G.encode(table, u, res, log2_t)
    o, _ = G.mlp_forward(feats, weights)
    _, g = G.loss_and_output_grad(o, w, target, fs, pb, clip=None, l2=1.0)
    m = G.mixture(o)
    lhat, q, _, _ = G.radiance_and_pdf(m, w)
    frozen_den, frozen_q = lhat ** 2 + G.REL_EPS, q

    def per_sample(oo):
        mm = G.mixture(oo)
        lh, qq, _, _ = G.radiance_and_pdf(mm, w)
        c = mm["c"]
        wgt = np.minimum(fs, 10.0)
        mag = mm["weight"].sum(axis=1) * frozen_q
        return (mag - target) ** 2 / frozen_den - wgt * np.log(qq) - wgt * np.log(c * frozen_q + (1 - c) * pb)

    eps = 1e-6
    for j in range(41):
        op, om = o.copy(), o.copy()
        op[:, j] += eps
        om[:, j] -= eps
        fd = (per_sample(op) - per_sample(om)) / (2 * eps)
        assert np.allclose(g[:, j], fd, rtol=1e-4, atol=1e-7), j


def test_network_and_table_gradients_match_finite_differences():
    table, weights, u, w, target, fs, pb, res, log2_t = setup_small(n=12)
    dws, dtable, o, _ = G.gradients(table, weights, u, w, target, fs, pb, res, log2_t, clip=None, l2=1.0)
    m = G.mixture(o)
    lhat, q, _, _ = G.radiance_and_pdf(m, w)
    frozen = lhat ** 2 + G.REL_EPS
    f = lambda tb, ws: surrogate(tb, ws, u, w, target, fs, pb, res, log2_t, frozen, q)
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
    # dense while (res+1)^3 <= T
    assert G.corner_index(3, 8, 1, 2, 3) == 1 + 2 * 4 + 3 * 16
    # hashed: brick (1, 1, 2) of vertex (5, 7, 11), local (1, 3, 3)
    h = 1 ^ ((1 * 2654435761) & 0xFFFFFFFF) ^ ((2 * 805459861) & 0xFFFFFFFF)
    assert h == 0xFE33D69A  # 0x9E3779B1 ^ 0x6004AF2A ^ 1
    local = 1 | (3 << 2) | (3 << 4)
    assert G.corner_index(100, 10, 5, 7, 11) == (((h << 6) & 0xFFFFFFFF) | local) & 1023
    assert G.corner_index(200, 20, 5, 7, 11) == (((h << 6) & 0xFFFFFFFF) | local) & ((1 << 20) - 1)


def test_hash_bricks_keep_cell_corners_together():
    """Inside a brick the 8 corners of a cell are distinct entries of one
    64-entry block; vectorized and scalar indexing agree."""
    rng = np.random.default_rng(3)
    for r in (3, 10, 100, 2048):
        c = rng.integers(0, r + 1, (300, 3))
        a = G.corner_indices(r, 12, c[:, 0], c[:, 1], c[:, 2])
        assert np.array_equal(a, [G.corner_index(r, 12, *map(int, x)) for x in c]), r
    x, y, z = 4 * 7 + 1, 4 * 3 + 2, 4 * 9  # cell (29..30, 14..15, 36..37) inside one brick
    idx = {G.corner_index(2048, 19, x + dx, y + dy, z + dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)}
    assert len(idx) == 8 and len({i >> 6 for i in idx}) == 1


def test_sparse_adam_skips_untouched_entries():
    p = np.array([1.0, 2.0, 3.0])
    m = np.array([0.1, 0.2, 0.0])
    v = np.array([0.01, 0.04, 0.0])
    G.Adam(sparse=True).step(p, np.array([0.5, 0.0, -1.0]), m, v, 3)
    assert p[1] == 2.0 and m[1] == 0.2 and v[1] == 0.04
    assert p[0] < 1.0 and p[2] > 3.0


def test_adam_matches_closed_form_first_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    m, v = np.zeros(2), np.zeros(2)
    G.Adam(lr=0.01, eps=0.0).step(p, g, m, v, 1)
    assert np.allclose(p, [1.0 - 0.01, -2.0 + 0.01])  # first Adam step moves lr * sign(g)


def test_gradient_clip_keeps_direction_and_bounds_max_norm():
    """Each loss part is clipped on its own: the relative-L2 part (target terms
    only) to L2_CLIP, the cross-entropy part (F/p terms only) to `clip`."""
    table, weights, u, w, target, fs, pb, res, log2_t = setup_small(n=32)
    feats, _ = G.encode(table, u, res, log2_t)
    o, _ = G.mlp_forward(feats, weights)
    zero = np.zeros_like(target)
    # cross-entropy part alone (relative L2 made inert by target = Lhat)
    lhat = G.radiance_and_pdf(G.mixture(o), w)[0]
    _, g_raw = G.loss_and_output_grad(o, w, lhat, fs * 1e3, pb, clip=None)
    _, g_clip = G.loss_and_output_grad(o, w, lhat, fs * 1e3, pb, clip=5.0)
    mx = np.abs(g_raw).max(axis=1)
    assert (mx > 5.0).any()
    assert np.all(np.abs(g_clip).max(axis=1) <= 5.0 + 1e-9)
    for r in range(len(mx)):
        if mx[r] > 5.0:
            assert np.allclose(g_clip[r], g_raw[r] * (5.0 / mx[r]))
        else:
            assert np.allclose(g_clip[r], g_raw[r], rtol=1e-12, atol=0)
    # relative-L2 part alone (F/p = 0): bounded by L2_CLIP
    _, g2_raw = G.loss_and_output_grad(o, w, target * 1e4, zero, pb, clip=None, l2=1.0)
    _, g2 = G.loss_and_output_grad(o, w, target * 1e4, zero, pb, l2=1.0)
    assert np.abs(g2_raw).max() > G.L2_CLIP
    assert np.all(np.abs(g2).max(axis=1) <= G.L2_CLIP + 1e-12)


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


def test_guide_cross_entropy_gradient_is_stable_far_from_every_lobe():
    """A direction no lobe covers (q underflows) still gets a finite gradient
    that moves weight toward the lobe nearest to it."""
    o = np.zeros((1, 64))
    for k in range(G.LOBES):
        o[0, 5 * k:5 * k + 3] = [0.0, 0.0, 1.0]
        o[0, 5 * k + 3] = 7.0  # kappa = e^7
    o[0, 0:3] = [0.0, 0.2, 1.0]  # lobe 0 slightly tilted toward +y
    w = np.array([[0.0, 1.0, 0.0]])
    _, g = G.loss_and_output_grad(o, w, np.array([0.0]), np.array([1.0]), np.array([0.1]), clip=None)
    assert np.all(np.isfinite(g))
    assert g[0, 4] < 0 and np.all(g[0, 9:40:5] > 0)  # raise lobe 0's weight, lower the others
