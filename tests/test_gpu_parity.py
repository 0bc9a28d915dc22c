"""GPU path tracer vs the CPU oracle at equal seeds.

Tolerances: primary-hit primitive ids must match except at most 1e-4 of the
pixels (grazing edges); every pixel's radiance within 1e-3 relative (absolute
floor 1e-6), with no outliers.  The kernels and the oracle evaluate the same
float expressions in the same order, so the renders are expected to agree bit
for bit; the fraction of bit-identical pixels is asserted too, so a drift
that stays inside the tolerance still shows.
"""
import numpy as np
import pytest

import oracle
import pupiloptixlab_amd as pt

pytestmark = pytest.mark.gpu

PRIM_TOL = 1e-4
REL_TOL = 1e-3
EXACT_MIN = 0.9999  # fraction of pixels whose three channels are bit-identical


def radiance_outliers(gpu, ref):
    g, r = gpu[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    err = np.abs(g - r) / np.maximum(np.abs(r), 1e-6 / REL_TOL)
    return float(np.mean(err.max(axis=-1) > REL_TOL)), float(np.mean(np.all(g == r, axis=-1)))


def render_both(tracer, scene, spp, **kw):
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=spp, **kw)
    out = {b: tracer.read(b) for b in (pt.BUF_ACCUM, pt.BUF_FRAME, pt.BUF_PRIM_ID, pt.BUF_ALBEDO, pt.BUF_NORMAL,
                                       pt.BUF_TEST)}
    okw = {k: v for k, v in kw.items() if k in ("max_depth", "rr_depth", "mis_mode", "accumulate")}
    ref = oracle.render(scene.desc, spp=spp, **okw)
    return out, ref


def test_cornell_primary_ids_and_radiance(tracer):
    """Cornell-box workload (512x512, 16 spp, max depth 8): full parity."""
    print("case: cornell 512x512 16spp depth 8", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL)
    out, ref = render_both(tracer, scene, 16)
    prim_mismatch = float(np.mean(out[pt.BUF_PRIM_ID] != ref["prim"]))
    outliers, exact = radiance_outliers(out[pt.BUF_ACCUM], ref["accum"])
    print(f"prim mismatch {prim_mismatch:.2e} radiance outliers {outliers:.2e} exact {exact:.4f}")
    assert prim_mismatch <= PRIM_TOL
    assert outliers == 0.0
    assert exact >= EXACT_MIN
    assert np.array_equal(out[pt.BUF_ACCUM], out[pt.BUF_FRAME])
    # AOVs of the last frame
    assert np.allclose(out[pt.BUF_ALBEDO], ref["albedo"], atol=0)
    assert float(np.mean(np.abs(out[pt.BUF_NORMAL] - ref["normal"]).max(axis=-1) > 1e-6)) <= PRIM_TOL
    assert float(np.mean(out[pt.BUF_TEST] != ref["test"])) == 0.0


def test_terrain_all_materials_reduced(tracer):
    """Procedural heightfield + icospheres, all seven BSDFs, area lights and a
    constant environment, at reduced size."""
    print("case: terrain 192x108 grid 128 spheres 48 4spp depth 16", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=3, width=192, height=108, grid=128, spheres=48)
    out, ref = render_both(tracer, scene, 4)
    prim_mismatch = float(np.mean(out[pt.BUF_PRIM_ID] != ref["prim"]))
    outliers, exact = radiance_outliers(out[pt.BUF_ACCUM], ref["accum"])
    print(f"prim mismatch {prim_mismatch:.2e} radiance outliers {outliers:.2e} exact {exact:.4f}")
    assert prim_mismatch <= PRIM_TOL
    assert outliers == 0.0
    assert exact >= EXACT_MIN


def test_consistent_mis_mode(tracer):
    print("case: cornell 128x128 consistent MIS", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=128, height=128)
    out, ref = render_both(tracer, scene, 8, mis_mode=pt.MIS_CONSISTENT)
    outliers, exact = radiance_outliers(out[pt.BUF_ACCUM], ref["accum"])
    print(f"radiance outliers {outliers:.2e} exact {exact:.4f}")
    assert outliers == 0.0
    assert exact >= EXACT_MIN


def test_no_accumulation_and_depth_one(tracer):
    print("case: cornell 64x64 accumulate off / max_depth 1", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=64, height=64)
    out, ref = render_both(tracer, scene, 3, accumulate=False)
    assert radiance_outliers(out[pt.BUF_ACCUM], ref["accum"]) == (0.0, 1.0)
    out, ref = render_both(tracer, scene, 2, max_depth=1)
    assert radiance_outliers(out[pt.BUF_ACCUM], ref["accum"]) == (0.0, 1.0)


def test_frames_per_launch_invariance(tracer):
    """Concurrent frames fold in frame order: identical bits for any launch width."""
    print("case: cornell 96x96 spp_per_launch 1 vs 5", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=96, height=96)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=7, spp_per_launch=1)
    a = tracer.read(pt.BUF_ACCUM)
    tracer.reset()
    tracer.render(spp=7, spp_per_launch=5)
    b = tracer.read(pt.BUF_ACCUM)
    tracer.reset()
    tracer.render(spp=3, spp_per_launch=3)
    tracer.render(spp=4, spp_per_launch=2)
    c = tracer.read(pt.BUF_ACCUM)
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)


def test_trace_rays_matches_oracle(tracer):
    print("case: random rays on terrain grid 96", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_TERRAIN, seed=1, width=64, height=64, grid=96, spheres=24)
    tracer.set_scene(scene)
    tracer.build_bvh()
    rng = np.random.default_rng(7)
    n = 50000
    o = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(-0.3, 1.5, n)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, np.full((n, 1), 1e-3, np.float32), d, np.full((n, 1), 1e16, np.float32)], axis=1)
    t_g, p_g, b1_g, b2_g = tracer.trace_rays(rays)
    t_o, p_o, b1_o, b2_o = oracle.trace(scene.desc, rays)
    assert np.array_equal(p_g, p_o)
    assert np.array_equal(t_g.view(np.uint32), t_o.view(np.uint32))
    assert np.array_equal(b1_g.view(np.uint32), b1_o.view(np.uint32))
    rays[:, 7] = rng.uniform(0.01, 2.0, n)
    a_g = tracer.trace_rays(rays, any_hit=True)[1] != pt.MISS
    a_o = oracle.trace(scene.desc, rays, any_hit=True)[1] != pt.MISS
    assert np.array_equal(a_g, a_o)


def test_determinism(tracer):
    print("case: cornell 64x64 rerun", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=64, height=64)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=4)
    a = tracer.read(pt.BUF_ACCUM)
    tracer.reset()
    tracer.render(spp=4)
    assert np.array_equal(a, tracer.read(pt.BUF_ACCUM))


@pytest.mark.parametrize("guided", [False, True])
def test_chained_launch_invariance(tracer, guided):
    """A launch's paths run as up to four concurrent chains of kernels of at
    least 10^6 paths (own queue parts, counters, streams and spill areas):
    34 frames at 1920x1080 (one launch of four chains; then launches of 5
    frames and one of 4, four chains each) and 3 single-frame launches (two
    chains each) equal the same frames rendered with every launch as one
    chain, bit for bit, AOVs and ray counts included; guided renders (network
    inference, per-chain output buffers) too."""
    kind = pt.SCENE_DOOR_ROOM if guided else pt.SCENE_TERRAIN
    print(f"case: {'door room' if guided else 'terrain'} 1920x1080, chained launches vs one chain", flush=True)
    scene = pt.HostScene.generate(kind, seed=2)
    assert scene.desc.width * scene.desc.height >= 2 * 10 ** 6  # a single 1080p frame splits into two chains
    tracer.set_scene(scene)
    tracer.build_bvh()
    kw = {"mis_mode": pt.MIS_CONSISTENT, "guide": True} if guided else {}
    if guided:
        tracer.guide_enable(seed=3)
        tracer.render(spp=4, mis_mode=pt.MIS_CONSISTENT, aov=False, guide=True, guide_train=True)
    bufs = (pt.BUF_ACCUM, pt.BUF_PRIM_ID, pt.BUF_ALBEDO, pt.BUF_NORMAL, pt.BUF_TEST)
    for spp, launch in ((34, 0), (34, 5), (3, 1)):
        out = []
        for one in (True, False):
            tracer.reset()
            before = tracer.stats()
            tracer.render(spp=spp, spp_per_launch=launch, one_chain=one, **kw)
            after = tracer.stats()
            rays = {k: after[k] - before[k] for k in ("rays_extension", "rays_shadow", "guide_queries")}
            out.append(({b: tracer.read(b) for b in bufs}, rays))
        assert out[0][1] == out[1][1]
        assert out[0][1]["rays_extension"] > spp * scene.desc.width * scene.desc.height
        for b in bufs:
            assert np.array_equal(out[0][0][b], out[1][0][b]), (launch, b)
