"""The network kernels split pixel and channel indices by multiply-high with host-computed
constants (nnk::make_divisor / nnk::divide in boltzmann_amd/csrc/nn_kernels.hpp).  This
compiles tests/native/divisor_check.cpp against that header with hipcc and runs its host
code, which compares the split with integer division over an exhaustive and a random range
(no GPU is used)."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None and not os.path.exists(HIPCC), reason="hipcc not found")
def test_multiply_high_division_matches_integer_division(tmp_path):
    exe = tmp_path / "divisor_check"
    src = os.path.join(HERE, "native", "divisor_check.cpp")
    build = subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-Xclang",
                            "-target-feature", "-Xclang", "-packed-fp32-ops", src, "-o", str(exe)],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.strip() == "ok", run.stdout + run.stderr
