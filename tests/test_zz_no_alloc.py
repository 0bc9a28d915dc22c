"""pt_render_noalloc: the allocation-free render entry point, called through
the C ABI, must give exactly the images pt_render gives for the same frames,
on consecutive calls whose inputs shrink (the scratch array keeps the last
call's contents), and must refuse a frame larger than its scratch array."""
import ctypes as C

import numpy as np
import pytest

import pupiloptixlab_amd as pt

pytestmark = pytest.mark.gpu

BUFS = (pt.BUF_ACCUM, pt.BUF_FRAME, pt.BUF_ALBEDO, pt.BUF_NORMAL, pt.BUF_TEST, pt.BUF_PRIM_ID)


def render_noalloc(tracer, spp, mis_mode=pt.MIS_REFERENCE, spp_per_launch=0, aov=True):
    flags = pt.RENDER_AOV if aov else 0
    p = pt.RenderParams(spp, 0, 2, mis_mode, 1, flags, spp_per_launch, 0)
    return pt.lib().pt_render_noalloc(tracer._ctx, C.byref(p))


def both(tracer, scene, spp, mis_mode=pt.MIS_REFERENCE, spp_per_launch=0):
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=spp, mis_mode=mis_mode)
    want = {b: tracer.read(b) for b in BUFS}
    tracer.reset()
    rc = render_noalloc(tracer, spp, mis_mode, spp_per_launch)
    assert rc == 0, pt.lib().pt_last_error(tracer._ctx)
    got = {b: tracer.read(b) for b in BUFS}
    return want, got


def test_noalloc_matches_render_on_shrinking_calls(tracer):
    cases = [
        ("cornell 256x192 6 spp", pt.HostScene.generate(pt.SCENE_CORNELL, width=256, height=192), 6,
         pt.MIS_REFERENCE, 0),
        ("terrain 160x90 (spheres, 7 BSDFs) 3 spp", pt.HostScene.generate(pt.SCENE_TERRAIN, seed=2, width=160,
                                                                          height=90, grid=64, spheres=12), 3,
         pt.MIS_CONSISTENT, 0),
        ("cornell 96x64 7 spp, 3 frames per launch", pt.HostScene.generate(pt.SCENE_CORNELL, width=96, height=64), 7,
         pt.MIS_REFERENCE, 3),
    ]
    for name, scene, spp, mode, per_launch in cases:
        print(f"case: noalloc vs pt_render, {name}", flush=True)
        want, got = both(tracer, scene, spp, mode, per_launch)
        for b in BUFS:
            assert np.array_equal(want[b], got[b]), (name, b)
        assert got[pt.BUF_ACCUM][..., :3].mean() > 0.0


def test_noalloc_progressive_and_limits(tracer):
    print("case: noalloc progressive accumulation, then a frame beyond the scratch", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=128, height=128)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=5)
    want = tracer.read(pt.BUF_ACCUM)
    tracer.reset()
    assert render_noalloc(tracer, 2) == 0
    assert render_noalloc(tracer, 3) == 0
    assert np.array_equal(tracer.read(pt.BUF_ACCUM), want)
    # guided rendering needs allocations
    p = pt.RenderParams(1, 0, 2, pt.MIS_CONSISTENT, 1, pt.RENDER_GUIDE, 0, 0)
    assert pt.lib().pt_render_noalloc(tracer._ctx, C.byref(p)) == -8
    big = pt.HostScene.generate(pt.SCENE_CORNELL, width=2560, height=1440)  # 3.7M paths per frame
    tracer.set_scene(big)
    tracer.build_bvh()
    assert render_noalloc(tracer, 1) == -4
    assert b"exceeds" in pt.lib().pt_last_error(tracer._ctx)
