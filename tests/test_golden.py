"""Regression pins: the oracle (CPU) and, on a GPU, the product path against
the golden vectors in tests/golden/oracle_vectors.json."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
import pupiloptixlab_amd as pt
from tests.test_oracle_pins import material

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.json")))


def bits(a):
    return [int(x) for x in np.asarray(a, np.float32).view(np.uint32).ravel()]


def test_rng_vectors():
    for case in GOLDEN["rng"]:
        vals, seeds = oracle.rng(case["pixel"], case["sample"], 8)
        assert bits(vals) == case["values"]
        assert [int(s) for s in seeds] == case["seeds"]


def test_bsdf_vectors():
    for case in GOLDEN["bsdf"]:
        m = material(case["type"])
        wo = np.array(case["wo"], np.uint32).view(np.float32)
        seed = 2463534242
        for ref in case["samples"]:
            s = oracle.bsdf_sample(m, wo, seed)
            seed = s["seed"]
            assert bits(s["wi"]) == ref["wi"] and bits(s["f"]) == ref["f"]
            assert bits([s["pdf"]])[0] == ref["pdf"] and s["lobe"] == ref["lobe"] and s["seed"] == ref["seed"]


def test_camera_vectors():
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=32, height=32)
    for case in GOLDEN["camera"]:
        o, d = oracle.camera_ray(scene.desc, case["x"], case["y"], *case["jitter"])
        assert bits(o) == case["origin"] and bits(d) == case["dir"]


def test_render_checksum():
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=32, height=32)
    r = oracle.render(scene.desc, spp=4)
    acc = np.ascontiguousarray(r["accum"])
    ref = GOLDEN["render"]
    assert hashlib.sha256(acc.tobytes()).hexdigest() == ref["sha256"]
    assert [int(v) for v in r["prim"][16]] == ref["prim_row16"]


@pytest.mark.gpu
def test_gpu_render_matches_golden(tracer):
    print("case: golden cornell 32x32 4spp", flush=True)
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=32, height=32)
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=4)
    acc = np.ascontiguousarray(tracer.read(pt.BUF_ACCUM))
    assert hashlib.sha256(acc.tobytes()).hexdigest() == GOLDEN["render"]["sha256"]
    assert [int(v) for v in tracer.read(pt.BUF_PRIM_ID)[16]] == GOLDEN["render"]["prim_row16"]
