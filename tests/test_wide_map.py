"""The wide fluid step (boltzmann_amd/csrc/update_wide.hpp) moves populations between lanes
of a wavefront and lets edge lanes load and store single elements.  This compiles
tests/native/wide_map_check.cpp against that header with hipcc and runs its host code,
which replays every wavefront's loads, rotations and stores of the streaming step for
2D and 3D shapes with rows shorter than, equal to and longer than one wavefront, and checks
the bounds, the AA mapping and that every element is written exactly once (no GPU is
used)."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None and not os.path.exists(HIPCC), reason="hipcc not found")
def test_wide_access_map(tmp_path):
    exe = tmp_path / "wide_map_check"
    src = os.path.join(HERE, "native", "wide_map_check.cpp")
    build = subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", src,
                            "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.strip() == "ok", run.stdout + run.stderr
