"""Guiding network on the GPU vs the float64 guide oracle (oracle/guide.py),
and guided rendering.

Tolerances (FP16 storage, FP32 accumulation on MFMA):
* forward outputs vs the half-precision-faithful oracle: |d| <= 1e-2 * (rms + |ref|),
  mean 5e-4 (a rounding flip of one FP16 activation shows up as ~1e-3);
* weight and table gradients vs the oracle: relative Frobenius error <= 2e-2
  per layer / per table;
* first Adam step: every parameter with a clear gradient moves by exactly
  -lr * sign(g) (table) or -lr * g / (|g| + 1e-8) (weights).
"""
import numpy as np
import pytest

import pupiloptixlab_amd as pt
from oracle import guide as G

pytestmark = pytest.mark.gpu

LOG2T = 14


def setup_guide(tracer, seed=0, log2_t=LOG2T, **kw):
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=64, height=64)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.guide_enable(log2_table=log2_t, seed=seed, **kw)
    info = tracer.guide_info()
    assert info["enabled"] == 1 and info["log2_table"] == log2_t
    assert info["resolution"] == G.resolutions(G.LEVELS, 16, 2048)
    return scene, info


def random_params(seed, log2_t=LOG2T, table_sd=0.5):
    rng = np.random.default_rng(seed)
    table = G.fp16(rng.normal(0, table_sd, (G.LEVELS, 1 << log2_t, 2))).astype(np.float32)
    dims = [32, 64, 64, 64, 41]
    w = np.zeros((4, 64, 64), np.float32)
    for l in range(4):
        w[l, :dims[l], :dims[l + 1]] = G.fp16(rng.normal(0, 1.2 / np.sqrt(dims[l]), (dims[l], dims[l + 1])))
    return table, w


def oracle_weights(w):
    dims = [32, 64, 64, 64, 64]
    return [w[l, :dims[l], :].astype(np.float64) for l in range(4)]


def world_positions(scene, n, seed):
    p = scene.positions()
    lo, hi = p.min(0), p.max(0)
    return (lo + np.random.default_rng(seed).random((n, 3)) * (hi - lo)).astype(np.float32)


def normalized(info, pos):
    lo = np.array(info["bounds_lo"], np.float32)
    inv = np.array(info["bounds_inv"], np.float32)
    return np.clip((pos - lo) * inv, 0.0, 1.0).astype(np.float64)


def test_guide_forward_matches_oracle(tracer):
    print("case: guide forward, 4096 positions, log2T 14", flush=True)
    scene, info = setup_guide(tracer)
    table, w = random_params(1)
    tracer.guide_set_params(table, w)
    tb, wb = tracer.guide_get_params()
    assert np.array_equal(tb, table) and np.array_equal(wb, w)
    pos = world_positions(scene, 4096, 2)
    out = tracer.guide_query(pos)
    u = normalized(info, pos)
    feats, _ = G.encode(table.astype(np.float64), u, info["resolution"], LOG2T, quantize=True)
    ref, _ = G.mlp_forward(feats, oracle_weights(w), quantize=True)
    ref = ref[:, :41]
    scale = np.sqrt(np.mean(ref ** 2, axis=0)) + np.abs(ref)
    err = np.abs(out - ref) / scale
    print(f"max rel err {err.max():.2e} mean {err.mean():.2e}")
    assert err.max() <= 1e-2 and err.mean() <= 5e-4
    # float64 model without rounding: the FP16 pipeline stays within 3% of it
    feats64, _ = G.encode(table.astype(np.float64), u, info["resolution"], LOG2T)
    ref64, _ = G.mlp_forward(feats64, oracle_weights(w))
    err64 = np.abs(out - ref64[:, :41]) / (np.sqrt(np.mean(ref64[:, :41] ** 2, axis=0)) + np.abs(ref64[:, :41]))
    assert err64.max() <= 3e-2


def make_samples(scene, n, seed, invalid_every=0):
    rng = np.random.default_rng(seed)
    pos = world_positions(scene, n, seed + 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    target = rng.gamma(1.0, 1.0, n)
    fs = rng.gamma(1.0, 1.0, n)
    pb = rng.uniform(0.05, 1.0, n)
    valid = np.ones(n, np.float32)
    if invalid_every:
        valid[::invalid_every] = 0.0
    return pt.PathTracer.guide_samples(pos, d, target, fs, pb, valid), pos, d, target, fs, pb, valid


def oracle_grads(info, table, w, pos, d, target, fs, pb, valid, fp_norm=1.0, l2=G.L2_WEIGHT):
    u = normalized(info, pos)
    samples32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    dws, dtable, o, loss = G.gradients(table.astype(np.float64), oracle_weights(w), u, samples32(d),
                                       samples32(target), samples32(fs), samples32(pb), info["resolution"], LOG2T,
                                       norm_fp=fp_norm, valid=samples32(valid), quantize=True, l2=l2)
    return dws, dtable, loss


def rel_fro(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("l2", [G.L2_WEIGHT, 1.0])
def test_guide_gradients_match_oracle(tracer, l2):
    print(f"case: guide gradients, 2048 weighted samples (every 7th unused), l2 weight {l2}", flush=True)
    scene, info = setup_guide(tracer, seed=3, l2_weight=l2)
    table, w = random_params(4, table_sd=0.3)
    tracer.guide_set_params(table, w)
    samples, pos, d, target, fs, pb, valid = make_samples(scene, 2048, 5, invalid_every=7)
    # sample weights as the collector assigns them (1 / keep probability)
    valid = valid * np.where(np.arange(2048) % 3 == 0, 4.0, 1.0).astype(np.float32)
    samples[:, 11] = valid
    tg, wg, loss = tracer.guide_gradients(samples, fp_norm=1.3)
    dws, dtable, loss_ref = oracle_grads(info, table, w, pos, d, target, fs, pb, valid, fp_norm=1.3, l2=l2)
    print(f"loss gpu {loss:.6f} oracle {loss_ref:.6f}")
    assert loss == pytest.approx(loss_ref, rel=5e-3)
    dims = [32, 64, 64, 64, 41]
    for l in range(4):
        e = rel_fro(wg[l, :dims[l], :dims[l + 1]], dws[l][:, :dims[l + 1]])
        print(f"layer {l} weight-gradient rel err {e:.2e}")
        assert e <= 2e-2, l
        # padding (layer-0 inputs 32..63, output-layer columns 41..63) gets no gradient
        assert np.all(wg[l, dims[l]:, :] == 0) and np.all(wg[l, :, dims[l + 1]:] == 0)
    e = rel_fro(tg, dtable)
    print(f"table-gradient rel err {e:.2e}")
    assert e <= 2e-2
    # Gradients are consumed: parameters unchanged, a second call gives the same.
    tb, wb = tracer.guide_get_params()
    assert np.array_equal(tb, table) and np.array_equal(wb, w)
    tg2, wg2, _ = tracer.guide_gradients(samples, fp_norm=1.3)
    assert rel_fro(wg2, wg) <= 1e-5 and rel_fro(tg2, tg) <= 1e-5


def test_guide_first_adam_step(tracer):
    print("case: guide first Adam step, 4096 samples", flush=True)
    scene, info = setup_guide(tracer, seed=6, learning_rate=1e-2)
    table, w = random_params(7, table_sd=0.3)
    tracer.guide_set_params(table, w)
    samples, pos, d, target, fs, pb, valid = make_samples(scene, 4096, 8)
    tg, wg, _ = tracer.guide_gradients(samples)
    tracer.guide_train(samples)
    info2 = tracer.guide_info()
    assert info2["steps"] == 1 and info2["skipped"] == 0
    tb, wb = tracer.guide_get_params()
    lr = 1e-2
    # table: Adam's first step moves every touched entry by lr * sign(g)
    clear = np.abs(tg) > 1e-3 * np.abs(tg).max()
    expect = (table - lr * np.sign(tg)).astype(np.float32)
    assert np.allclose(tb[clear], expect[clear], rtol=0, atol=1e-6)
    assert np.array_equal(tb[tg == 0], table[tg == 0])
    # weights: eps 1e-8
    m = np.abs(wg) > 1e-3 * np.abs(wg).max()
    expect_w = w - lr * wg / (np.abs(wg) + 1e-8)
    assert np.allclose(wb[m], expect_w[m], rtol=0, atol=1e-6)
    assert np.array_equal(wb[wg == 0], w[wg == 0])


def test_guide_learns_a_radiance_field(tracer):
    """Online fit of a synthetic field: radiance from a vMF lobe (kappa 20)
    pointing from the hit position to a fixed point, sampled uniformly.  After
    300 steps the learned guiding density is far higher toward the lobe than
    away from it."""
    print("case: guide fit synthetic field, 300 steps x 4096", flush=True)
    scene, info = setup_guide(tracer, seed=9)
    p = scene.positions()
    lo, hi = p.min(0), p.max(0)
    c = (lo + hi) / 2 + np.array([0.0, 0.3, 0.0]) * (hi - lo)
    rng = np.random.default_rng(10)
    losses = []
    for it in range(300):
        pos = (lo + rng.random((4096, 3)) * (hi - lo)).astype(np.float32)
        mu = c - pos
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        d = rng.normal(size=(4096, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        target = 2.0 * G.vmf(np.full(4096, 20.0), np.sum(d * mu, axis=1))
        # uniform directions: F/p is proportional to the target
        s = pt.PathTracer.guide_samples(pos, d, target, target, np.full(4096, 1 / (4 * np.pi)))
        losses.append(tracer.guide_train(s, fp_norm=float(target.mean())))
    info = tracer.guide_info()
    print(f"loss first {np.mean(losses[:10]):.4f} last {np.mean(losses[-10:]):.4f} steps {info['steps']}")
    assert info["steps"] + info["skipped"] == 300 and info["skipped"] <= 3
    assert np.mean(losses[-10:]) < np.mean(losses[:10]) - 1.0
    pos = (lo + 0.1 * (hi - lo) + rng.random((512, 3)) * 0.8 * (hi - lo)).astype(np.float32)
    mu = c - pos
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    m = G.mixture(tracer.guide_query(pos).astype(np.float64))
    lh = lambda dd: G.radiance_and_pdf(m, dd)[1]
    ratio = lh(mu) / np.maximum(lh(-mu), 1e-30)
    print(f"median q(toward)/q(away) {np.median(ratio):.1f}")
    assert np.mean(ratio > 5.0) > 0.9


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def relmse_trimmed(img, ref, drop=1e-3):
    """relMSE without the largest `drop` fraction of per-pixel errors: the
    gain checks compare the bulk of the image, since a single rare firefly
    (one sample carrying up to 20x weight where the learned selection
    probability is high) can outweigh every other pixel at these sizes.
    Bias is checked on the plain image mean."""
    e = np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2), axis=-1).ravel()
    keep = e.size - int(np.ceil(drop * e.size))
    return float(np.mean(np.partition(e, keep - 1)[:keep]))


def test_guided_furnace_is_unbiased(tracer):
    """Closed emissive diffuse cube: the guided estimator converges to the
    closed form Le * sum_{k<depth} albedo^k whatever the network outputs."""
    print("case: guided furnace 64x64 depth 6, train 32 + render 96 spp", flush=True)
    albedo, depth = 0.7, 6
    scene = pt.HostScene.generate(pt.SCENE_FURNACE, width=64, height=64, max_depth=depth, albedo=albedo)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.guide_enable(log2_table=16, batch_size=4096, train_stride=4)
    tracer.render(spp=32, mis_mode=pt.MIS_CONSISTENT, guide=True, guide_train=True)
    tracer.render(spp=96, mis_mode=pt.MIS_CONSISTENT, guide=True)
    img = tracer.read(pt.BUF_ACCUM)[..., :3]
    expected = sum(albedo ** k for k in range(depth))
    st = tracer.stats()
    print(f"mean {img.mean():.5f} expected {expected:.5f} guide steps {st['guide_steps']} "
          f"samples {st['guide_samples']}")
    assert np.all(np.isfinite(img))
    assert float(img.mean()) == pytest.approx(expected, rel=5e-3)
    assert st["guide_samples"] > 0 and tracer.guide_info()["steps"] > 0


def test_guided_cornell_matches_plain(tracer):
    """Guided (trained online) and plain renders of the Cornell box converge
    to the same image."""
    print("case: cornell 128x128 guided vs plain consistent", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=128, height=128)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.set_random_seed(1 << 24)  # reference samples independent of frames 0, 1, ...
    tracer.render(spp=1024, mis_mode=pt.MIS_CONSISTENT)
    ref = tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    tracer.reset()
    tracer.guide_enable(log2_table=16, batch_size=8192, train_stride=2)
    tracer.render(spp=64, mis_mode=pt.MIS_CONSISTENT, guide=True, guide_train=True)
    tracer.render(spp=448, mis_mode=pt.MIS_CONSISTENT, guide=True)
    img = tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    print(f"mean guided {img.mean():.5f} plain {ref.mean():.5f} relmse {relmse(img, ref):.2e}")
    assert np.all(np.isfinite(img))
    assert img.mean() == pytest.approx(ref.mean(), rel=1e-2)
    assert relmse(img, ref) < 2e-2


def test_guided_errors(tracer):
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=32, height=32)
    tracer.set_scene(scene)
    tracer.build_bvh()
    with pytest.raises(pt.PupilError) as e:
        tracer.render(spp=1, mis_mode=pt.MIS_CONSISTENT, guide=True)
    assert e.value.code == -10
    tracer.guide_enable(log2_table=12)
    with pytest.raises(pt.PupilError):
        tracer.render(spp=1, mis_mode=pt.MIS_REFERENCE, guide=True)
    with pytest.raises(pt.PupilError):
        tracer.guide_train(np.zeros((100, 12), np.float32))
    with pytest.raises(pt.PupilError):
        tracer.guide_enable(log2_table=30)


def test_guided_door_room_gains_over_plain(tracer):
    """Hard indirect lighting (the room lit through a 5 cm door gap), reduced
    to 480x270: after 64 training frames the guided estimator reaches less
    than half the relMSE of plain BSDF sampling at the same sample count
    (measured against a 4096-spp FP64 plain render).  Training samples are
    well formed."""
    print("case: door room 480x270, plain 256 spp vs guided (train 64) 256 spp, ref 4096", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_DOOR_ROOM, width=480, height=270)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.set_random_seed(1 << 24)  # reference samples independent of frames 0, 1, ...
    tracer.render(spp=4096, mis_mode=pt.MIS_CONSISTENT, fp64=True, aov=False)
    ref = tracer.read(pt.BUF_ACCUM64) / 4096
    tracer.reset()
    tracer.render(spp=256, mis_mode=pt.MIS_CONSISTENT, aov=False)
    plain_img = tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    plain, plain_t = relmse(plain_img, ref), relmse_trimmed(plain_img, ref)
    tracer.reset()
    tracer.guide_enable(seed=1)
    tracer.render(spp=64, mis_mode=pt.MIS_CONSISTENT, aov=False, guide=True, guide_train=True)
    s = tracer.guide_read_samples()
    assert len(s) > 0
    assert np.all(np.isfinite(s)) and np.all(s[:, 8] >= 0) and np.all(s[:, 9] >= 0) and np.all(s[:, 11] > 0)
    assert np.allclose(np.linalg.norm(s[:, 4:7], axis=1), 1.0, atol=1e-5)
    assert np.mean(s[:, 9] > 0) > 0.5  # contribution-carrying samples are kept first
    tracer.render(spp=192, mis_mode=pt.MIS_CONSISTENT, aov=False, guide=True)
    img = tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    guided, guided_t = relmse(img, ref), relmse_trimmed(img, ref)
    info = tracer.guide_info()
    print(f"relmse plain {plain:.4f} guided {guided:.4f} ratio {plain / guided:.2f}; trimmed plain {plain_t:.4f} "
          f"guided {guided_t:.4f} ratio {plain_t / guided_t:.2f}; steps {info['steps']}")
    assert abs(img.mean() / ref.mean() - 1) < 0.02
    assert guided_t < 0.5 * plain_t
    assert guided < plain  # untrimmed: the metric itself, fireflies included
