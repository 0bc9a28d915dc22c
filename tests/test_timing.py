"""Kernel active time (pupiloptixlab_amd/csrc/pt_timing.h): the union of a
kernel kind's dispatch intervals, which the library reports per kernel kind
when a launch's concurrent chains overlap, checked on the host by
tests/timing_sim.cpp against a brute-force count on fixed and seeded random
interval sets."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_interval_union(tmp_path):
    exe = str(tmp_path / "timing_sim")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "pupiloptixlab_amd", "csrc"),
                    os.path.join(ROOT, "tests", "timing_sim.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
