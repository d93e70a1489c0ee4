"""Builds the native parts of boltzmann_amd in place.

    python boltzmann_amd/build.py

* libbolt_hip.so   -- HIP kernels for gfx950 and the C ABI of include/bolt.h (hipcc)
* _boltzmann*.so   -- the Python module over that ABI (g++ + pybind11)
* bw_probe         -- memory-ceiling probe for the update's access pattern (bench.py --probe)

Every HIP translation unit is compiled with packed-FP32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-value",
    "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in res.stdout.splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if out:
        print(out, flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {cmd[0]}")


def lib_path() -> str:
    return os.path.join(HERE, "libbolt_hip.so")


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_boltzmann" + suffix)


def build_hip() -> None:
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", os.path.join(CSRC, "bolt_runtime.hip"),
          os.path.join(CSRC, "bolt_nn.hip"), "-o", lib_path()])


def probe_path() -> str:
    return os.path.join(HERE, "bw_probe")


def build_probe() -> None:
    _run([HIPCC, *[f for f in HIP_FLAGS if f != "-fPIC"], os.path.join(CSRC, "bw_probe.hip"),
          "-o", probe_path()])


def build_ext() -> None:
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          f"-I{py_inc}", f"-I{pybind11.get_include()}", f"-I{INCLUDE}",
          os.path.join(CSRC, "bolt_py.cpp"), "-o", ext_path(),
          f"-L{HERE}", "-lbolt_hip", "-Wl,-rpath,$ORIGIN"])


def build() -> None:
    build_hip()
    build_ext()
    build_probe()


if __name__ == "__main__":
    build()
    sys.exit(0)
