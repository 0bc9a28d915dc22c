"""The C++ drop-in layer: Pupil::System + Pupil::pt::PTPass (+ DenoisePass),
driven by the pt_render command-line program, must produce exactly what the
C ABI gives for the same scene and frames.

The program adds the passes as example/path_tracer/main.cpp:5-21 does, loads
the scene through System::SetScene (system/system.cpp:136-173), and runs one
PTPass::OnRun per frame (pt_pass.cpp:39-57: one launch of one sample per
pixel, accumulation restarted when dirty); the ctypes side renders all frames
in one pt_render call.  Frames are folded in frame order either way, so the
"final result" buffers must agree bit for bit.  The EXR stores buffer row 0
(the image bottom) last (util/texture.cpp:37-44)."""
import os
import subprocess

import numpy as np
import pytest

import pupiloptixlab_amd as pt

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pupiloptixlab_amd", "pt_render")

SCENE_XML = """<scene version="3.0.0">
  <integrator type="path"><integer name="max_depth" value="6"/></integrator>
  <sensor type="perspective"><float name="fov" value="45"/>
    <transform name="to_world"><lookat origin="0, 1.2, 3.2" target="0, 0.6, 0" up="0, 1, 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="160"/><integer name="height" value="120"/></film>
  </sensor>
  <bsdf type="twosided" id="floor"><bsdf type="diffuse"><rgb name="reflectance" value="0.6, 0.5, 0.4"/></bsdf></bsdf>
  <shape type="rectangle">
    <transform name="to_world"><scale value="3"/><rotate x="1" angle="-90"/></transform>
    <ref id="floor"/>
  </shape>
  <shape type="cube">
    <transform name="to_world"><scale value="0.3"/><rotate y="1" angle="30"/><translate value="-0.5, 0.3, 0"/></transform>
    <bsdf type="roughplastic"><float name="alpha" value="0.2"/><rgb name="diffuse_reflectance" value="0.2, 0.4, 0.8"/></bsdf>
  </shape>
  <shape type="sphere">
    <transform name="to_world"><scale value="0.35"/><translate value="0.6, 0.35, 0.2"/></transform>
    <bsdf type="roughconductor"><string name="material" value="Al"/><float name="alpha" value="0.15"/></bsdf>
  </shape>
  <shape type="rectangle">
    <transform name="to_world"><scale value="0.4"/><rotate x="1" angle="90"/><translate value="0, 2, 0"/></transform>
    <bsdf type="diffuse"><rgb name="reflectance" value="0"/></bsdf>
    <emitter type="area"><rgb name="radiance" value="12, 10, 8"/></emitter>
  </shape>
  <emitter type="constant"><rgb name="radiance" value="0.05"/></emitter>
</scene>
"""


def run_cli(args, out):
    r = subprocess.run([CLI] + args + ["--out", out], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    return np.flipud(pt.load_image(out))[..., :3]


def test_cli_system_matches_c_abi(tracer, tmp_path):
    print("case: pt_render CLI (System + PTPass) vs ctypes pt_render, XML scene 160x120 12 spp", flush=True)
    xml = tmp_path / "scene.xml"
    xml.write_text(SCENE_XML)
    got = run_cli(["--scene", str(xml), "--spp", "12"], str(tmp_path / "a.exr"))
    scene = pt.HostScene.load_xml(str(xml))
    tracer.set_scene(scene)
    tracer.build_bvh()
    tracer.render(spp=12)
    want = tracer.read(pt.BUF_FRAME)[..., :3]
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    assert want.mean() > 0.01


def test_cli_denoise_pass_matches_c_abi(tracer, tmp_path):
    print("case: pt_render CLI --denoise (DenoisePass) vs ctypes render + pt_denoise, Cornell 512x512 4 spp",
          flush=True)
    got = run_cli(["--cornell", "--spp", "4", "--max-depth", "5", "--denoise"], str(tmp_path / "d.exr"))
    tracer.set_scene(pt.HostScene.generate(pt.SCENE_CORNELL))
    tracer.build_bvh()
    tracer.render(spp=4, max_depth=5)
    want = tracer.denoise(albedo=True, normal=True)[..., :3]
    assert np.array_equal(got, want)
    noisy = tracer.read(pt.BUF_ACCUM)[..., :3]
    assert not np.array_equal(want, noisy)
