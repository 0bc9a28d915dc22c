"""Writes the golden vectors in this directory from the CPU oracle.

The reference ships no test vectors for its path tracer (it has no tests and
cannot be built here), so these fixtures freeze the oracle's outputs after it
was pinned by closed forms and invariants (tests/test_oracle_pins.py); they
catch later regressions of the oracle and give the GPU tests fixed targets.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
import pupiloptixlab_amd as pt  # noqa: E402


def f32bits(a):
    return [int(x) for x in np.asarray(a, np.float32).view(np.uint32).ravel()]


def material(kind):
    from tests.test_oracle_pins import material as mk  # same material builder as the pin tests

    return mk(kind, 0.3)


def main():
    out = {"rng": [], "bsdf": [], "camera": [], "render": {}}
    for pixel, sample in [(0, 0), (7, 3), (262143, 15), (2073599, 63)]:
        vals, seeds = oracle.rng(pixel, sample, 8)
        out["rng"].append({"pixel": pixel, "sample": sample, "values": f32bits(vals),
                           "seeds": [int(s) for s in seeds]})
    wo = np.array([0.48, 0.36, 0.8], np.float32)
    for kind in range(1, 8):
        m = material(kind)
        seed = 2463534242
        samples = []
        for _ in range(4):
            s = oracle.bsdf_sample(m, wo, seed)
            seed = s["seed"]
            samples.append({"wi": f32bits(s["wi"]), "f": f32bits(s["f"]), "pdf": f32bits([s["pdf"]])[0],
                            "lobe": s["lobe"], "seed": s["seed"]})
        out["bsdf"].append({"type": kind, "wo": f32bits(wo), "samples": samples})
    scene = pt.HostScene.generate(pt.SCENE_CORNELL, width=32, height=32)
    for x, y, jx, jy in [(0, 0, 0.5, 0.5), (31, 31, 0.1, 0.9), (16, 8, 0.3, 0.7)]:
        o, d = oracle.camera_ray(scene.desc, x, y, jx, jy)
        out["camera"].append({"x": x, "y": y, "jitter": [jx, jy], "origin": f32bits(o), "dir": f32bits(d)})
    r = oracle.render(scene.desc, spp=4)
    acc = np.ascontiguousarray(r["accum"])
    out["render"] = {"scene": "cornell 32x32", "spp": 4,
                     "sha256": hashlib.sha256(acc.tobytes()).hexdigest(),
                     "mean": float(acc[..., :3].mean()),
                     "center_pixel": f32bits(acc[16, 16, :3]),
                     "prim_row16": [int(v) for v in r["prim"][16]]}
    with open(os.path.join(HERE, "oracle_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
