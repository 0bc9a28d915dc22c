"""How a launch's paths split into concurrent chains of kernels
(pupiloptixlab_amd/csrc/pt_chains.h, used by pt_render, the batched launches
and their raygen kernels), checked on the host by tests/chain_sim.cpp:
every (frame, pixel) of a launch is traced by exactly one chain slot, the
chains' queue parts are disjoint and cover the launch (films of 1 to 2M
pixels, 1 to 64 frames, with and without PT_RENDER_ONE_CHAIN), and every
batched chain's path count fits its fixed share of the queues."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_chain_split(tmp_path):
    exe = str(tmp_path / "chain_sim")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "pupiloptixlab_amd", "csrc"),
                    os.path.join(ROOT, "tests", "chain_sim.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
