"""The batched, guided and guided + batched workloads at their own sizes.

- Batched Cornell-box variants (50K-2M triangles, films up to 1920x1080):
  every view of a batch equals the oracle's render of its scene, and a
  full-size view at 1920x1080 equals its single-view render bit for bit.
- Guided door room at 1920x1080 (depth 16): at equal samples per pixel the
  online-trained guide has at most half the relMSE of plain BSDF sampling,
  untrimmed, against an FP64 plain reference drawn from frames that neither
  render uses, and its image mean agrees with the reference within 2%.
- Guided + batched at 1920x1080 (a door-room view and a terrain view in one
  wavefront loop, per-view hash grids, shared MLP): finite, unbiased per view,
  and the door view gains over the plain batch at equal samples.

relMSE(I, R) = mean over pixels and channels of (I - R)^2 / (R^2 + 0.01).
"""
import numpy as np
import pytest

import oracle
import pupiloptixlab_amd as pt

pytestmark = pytest.mark.gpu

REF_SEED = 1 << 24  # reference frames start here: disjoint from the tested frames 0, 1, ...
REL_TOL = 1e-3


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def radiance_outliers(img, ref):
    g, r = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    err = np.abs(g - r) / np.maximum(np.abs(r), 1e-6 / REL_TOL)
    return int(np.count_nonzero(err.max(axis=-1) > REL_TOL)), float(np.mean(np.all(g == r, axis=-1)))


def test_batched_variants_match_oracle(tracer):
    """Three Cornell variants of 50K, 120K and 70K triangles (films reduced so
    the oracle finishes in seconds) in one batch: each view's radiance equals
    the oracle's render of that scene (no pixel beyond 1e-3 relative), and the
    primary-hit ids of each scene's single-view render equal the oracle's."""
    print("case: batch of 3 Cornell variants 50K-120K tris vs oracle", flush=True)
    specs = [(11, 50_000, 96, 64, 4), (12, 120_000, 64, 64, 3), (13, 70_000, 80, 48, 5)]
    views = [(pt.HostScene.generate(pt.SCENE_CORNELL_VARIANT, seed=s, triangles=t, width=w, height=h), spp)
             for s, t, w, h, spp in specs]
    for (scene, _), (_, t, _, _, _) in zip(views, specs):
        assert scene.desc.num_triangles >= 0.9 * t
    tracer.batch_clear()
    for scene, spp in views:
        tracer.batch_add_view(scene, spp)
    tracer.batch_render(mis_mode=pt.MIS_REFERENCE, graph=True, paths_per_launch=20000)
    for k, (scene, spp) in enumerate(views):
        ref = oracle.render(scene.desc, spp=spp)
        bad, exact = radiance_outliers(tracer.batch_read(k), ref["accum"])
        print(f"view {k}: {scene.desc.num_triangles} tris, outliers {bad}, bit-identical {exact:.4f}")
        assert bad == 0
        tracer.set_scene(scene)
        tracer.build_bvh()
        tracer.render(spp=spp)  # primitive ids of the last frame, as the oracle keeps
        assert float(np.mean(tracer.read(pt.BUF_PRIM_ID) != ref["prim"])) <= 1e-4
    tracer.batch_clear()


def test_batched_full_size_view_equals_single_render(tracer):
    """A Cornell variant with its drawn triangle count (50K-2M) at 1920x1080
    next to a 512x512 variant: the batch gives each view exactly its
    single-view image, and the images are finite and lit."""
    print("case: batch 1920x1080 variant (drawn tris) + 512x512 variant, 4 spp", flush=True)
    views = [(pt.HostScene.generate(pt.SCENE_CORNELL_VARIANT, seed=21, width=1920, height=1080), 4),
             (pt.HostScene.generate(pt.SCENE_CORNELL_VARIANT, seed=22, width=512, height=512), 16)]
    assert views[0][0].desc.num_triangles >= 50_000
    single = []
    for scene, spp in views:
        tracer.set_scene(scene)
        tracer.build_bvh()
        tracer.render(spp=spp, aov=False)
        single.append(tracer.read(pt.BUF_ACCUM))
    tracer.batch_clear()
    for scene, spp in views:
        tracer.batch_add_view(scene, spp)
    tracer.batch_render(mis_mode=pt.MIS_REFERENCE, graph=True)
    for k, img in enumerate(single):
        got = tracer.batch_read(k)
        assert np.all(np.isfinite(got)) and got[..., :3].mean() > 0.01
        assert np.array_equal(got, img), f"view {k} differs from its single-view render"
    tracer.batch_clear()


def test_guided_door_room_1080p_halves_relmse(tracer):
    """The door room at 1920x1080, depth 16: plain 128 spp vs guided 128 spp
    (the first 32 frames train the guide online), against a 1024-spp FP64
    plain reference from frames 2^24 on.  The reference's own noise adds to
    both errors, so the measured ratio understates the true one."""
    print("case: door room 1920x1080, plain 128 spp vs guided 128 spp (train 32), ref 1024", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_DOOR_ROOM)
    assert (scene.desc.width, scene.desc.height, scene.desc.max_depth) == (1920, 1080, 16)
    tracer.set_scene(scene)
    tracer.build_bvh()
    mis = pt.MIS_CONSISTENT
    tracer.set_random_seed(REF_SEED)
    tracer.render(spp=1024, mis_mode=mis, fp64=True, aov=False)
    ref = tracer.read(pt.BUF_ACCUM64) / 1024
    tracer.reset()
    tracer.render(spp=128, mis_mode=mis, aov=False)
    plain = relmse(tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64), ref)
    tracer.reset()
    tracer.guide_enable(seed=3)
    tracer.render(spp=32, mis_mode=mis, aov=False, guide=True, guide_train=True)
    tracer.render(spp=96, mis_mode=mis, aov=False, guide=True)
    img = tracer.read(pt.BUF_ACCUM)[..., :3].astype(np.float64)
    guided = relmse(img, ref)
    bias = img.mean() / ref.mean() - 1
    print(f"relmse plain {plain:.4f} guided {guided:.4f} ratio {plain / guided:.2f}; mean {bias:+.4f}")
    tracer.guide_disable()
    assert np.all(np.isfinite(img))
    assert abs(bias) < 0.02
    assert plain / guided >= 2.0


def test_guided_batch_door_and_terrain_1080p(tracer):
    """A door-room view (128 spp) and a terrain view (1024x1024 heightfield,
    256 icospheres, all seven BSDFs; 32 spp) at 1920x1080 in one wavefront
    loop with per-view hash grids and a shared MLP, trained on the first 32
    frames of each view: every image is finite, each view's mean agrees with
    its FP64 reference within 2% (terrain) / 3% (door), and the door view's
    relMSE is below the plain batch's at the same samples."""
    print("case: guided batch door 1080p 128 spp + terrain 1080p 32 spp, train 32", flush=True)
    views = [(pt.HostScene.generate(pt.SCENE_DOOR_ROOM, seed=1), 128),
             (pt.HostScene.generate(pt.SCENE_TERRAIN, seed=101, grid=1024), 32)]
    mis = pt.MIS_CONSISTENT
    refs = []
    for (scene, _), n in zip(views, (1024, 256)):
        tracer.set_scene(scene)
        tracer.build_bvh()
        tracer.set_random_seed(REF_SEED)
        tracer.render(spp=n, mis_mode=mis, fp64=True, aov=False)
        refs.append(tracer.read(pt.BUF_ACCUM64) / n)
    tracer.batch_clear()
    for scene, spp in views:
        tracer.batch_add_view(scene, spp)
    tracer.batch_render(mis_mode=mis, graph=True)
    plain = [relmse(tracer.batch_read(k)[..., :3].astype(np.float64), refs[k]) for k in range(2)]
    tracer.batch_reset()
    tracer.guide_enable_batch(seed=4)
    tracer.batch_render(mis_mode=mis, graph=True, guide=True, guide_train=True, guide_train_spp=32)
    for k, tol in enumerate((0.03, 0.02)):
        img = tracer.batch_read(k)[..., :3].astype(np.float64)
        assert np.all(np.isfinite(img))
        g = relmse(img, refs[k])
        bias = img.mean() / refs[k].mean() - 1
        print(f"view {k}: relmse plain {plain[k]:.5f} guided {g:.5f} ratio {plain[k] / g:.2f}; mean {bias:+.4f}")
        assert abs(bias) < tol
        if k == 0:
            assert g < plain[k]
    tracer.batch_clear()
