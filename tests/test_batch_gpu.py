"""Batched multi-view rendering: several scenes (each with its own camera,
film size, depth and frame count) traced in one wavefront loop must give, for
every view, exactly the image pt_render gives for that scene alone (same
seeds: bit-identical), with direct launches, with hipGraph replay and after
autotuning."""
import numpy as np
import pytest

import pupiloptixlab_amd as pt

pytestmark = pytest.mark.gpu

VIEWS = [  # (seed, triangles, width, height, spp)
    (1, 2000, 64, 64, 3),
    (2, 9000, 96, 48, 5),
    (3, 4000, 80, 80, 2),
    (4, 20000, 40, 72, 4),
]


def make_views():
    return [(pt.HostScene.generate(pt.SCENE_CORNELL_VARIANT, seed=s, triangles=t, width=w, height=h), spp)
            for s, t, w, h, spp in VIEWS]


def single_renders(tracer, views, mis_mode, calls=1):
    out = []
    for scene, spp in views:
        tracer.set_scene(scene)
        tracer.build_bvh()
        for _ in range(calls):
            tracer.render(spp=spp, mis_mode=mis_mode, aov=False)
        out.append(tracer.read(pt.BUF_ACCUM))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_renders(tracer, graph):
    print(f"case: 4 Cornell variants, batch vs pt_render, graph={graph}", flush=True)
    views = make_views()
    ref = single_renders(tracer, views, pt.MIS_REFERENCE, calls=2)
    tracer.batch_clear()
    for scene, spp in views:
        tracer.batch_add_view(scene, spp)
    # small launches so that views split across launches and launches mix views
    tracer.batch_render(mis_mode=pt.MIS_REFERENCE, graph=graph, paths_per_launch=16384)
    tracer.batch_render(mis_mode=pt.MIS_REFERENCE, graph=graph, paths_per_launch=16384)
    st = tracer.batch_stats()
    assert st["views"] == 4 and st["launches"] > 2 and st["rays_extension"] > 0
    for v, r in enumerate(ref):
        img = tracer.batch_read(v)
        assert img.shape == r.shape
        assert np.array_equal(img, r), f"view {v} differs"
    tracer.batch_clear()


def test_batch_consistent_mode_and_fp64(tracer):
    print("case: batch consistent MIS + FP64 sums", flush=True)
    views = make_views()[:2]
    ref = single_renders(tracer, views, pt.MIS_CONSISTENT)
    tracer.batch_clear()
    for scene, spp in views:
        tracer.batch_add_view(scene, spp)
    tracer.batch_render(mis_mode=pt.MIS_CONSISTENT, fp64=True)
    for v, r in enumerate(ref):
        assert np.array_equal(tracer.batch_read(v), r)
        s64 = tracer.batch_read(v, pt.BUF_ACCUM64)
        spp = views[v][1]
        assert np.allclose(s64 / spp, r[..., :3], rtol=1e-5, atol=1e-6)
    tracer.batch_clear()


def test_batch_autotune_keeps_results(tracer):
    print("case: batch autotune then render", flush=True)
    views = make_views()
    ref = single_renders(tracer, views, pt.MIS_REFERENCE)
    tracer.batch_clear()
    for scene, spp in views:
        tracer.batch_add_view(scene, spp)
    t = tracer.batch_autotune()
    print(t)
    assert t["candidates"] >= 3 and t["best_ns_per_path"] > 0
    assert t["best_ns_per_path"] <= t["default_ns_per_path"] * 1.0001
    assert t["paths_per_launch"] >= 80 * 80 and t["shade_block"] in (64, 128, 256)
    tracer.batch_render(mis_mode=pt.MIS_REFERENCE, graph=True)
    for v, r in enumerate(ref):
        assert np.array_equal(tracer.batch_read(v), r)
    tracer.batch_clear()


def test_batch_errors(tracer):
    tracer.batch_clear()
    with pytest.raises(pt.PupilError):
        tracer.batch_render()
    scene = pt.HostScene.generate(pt.SCENE_CORNELL_VARIANT, seed=5, triangles=2000, width=64, height=64)
    with pytest.raises(pt.PupilError):
        tracer.batch_add_view(scene, 0)
    tracer.batch_add_view(scene, 1)
    with pytest.raises(pt.PupilError):
        tracer.batch_render(paths_per_launch=100)  # smaller than one frame
    with pytest.raises(pt.PupilError):
        tracer.batch_render(shade_block=100)
    tracer.batch_clear()
