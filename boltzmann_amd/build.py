"""Builds the native parts of boltzmann_amd in place.

    python boltzmann_amd/build.py

* libbolt_hip.so   -- HIP kernels for gfx950 and the C ABI of include/bolt.h (hipcc; the
                     translation units compile in parallel into _build/ and are linked)
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


def _report(cmd: list[str], rc: int, output: str) -> None:
    out = "\n".join(line for line in output.splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if out:
        print(out, flush=True)
    if rc != 0:
        raise RuntimeError(f"build step failed ({rc}): {' '.join(cmd)}")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    _report(cmd, res.returncode, res.stdout)


def _run_parallel(cmds: list[list[str]]) -> None:
    """Runs the commands at the same time (one per translation unit); fails if any fails."""
    procs = []
    for cmd in cmds:
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    errors = []
    for cmd, p in procs:
        out, _ = p.communicate()
        try:
            _report(cmd, p.returncode, out)
        except RuntimeError as e:
            errors.append(str(e))
    if errors:
        raise RuntimeError("; ".join(errors))


def lib_path() -> str:
    return os.path.join(HERE, "libbolt_hip.so")


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_boltzmann" + suffix)


HIP_UNITS = ("bolt_runtime.hip", "bolt_wide.hip", "bolt_wide_keep.hip", "bolt_row.hip", "bolt_nn.hip")


def build_hip(extra: list[list[str]] = ()) -> None:
    """Compiles the translation units of libbolt_hip.so in parallel (with any `extra`
    independent commands), then links them."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(u)[0] + ".o") for u in HIP_UNITS]
    _run_parallel([[HIPCC, *HIP_FLAGS, "-c", f"-I{INCLUDE}", os.path.join(CSRC, u), "-o", o]
                   for u, o in zip(HIP_UNITS, objs)] + list(extra))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", *objs, "-o", lib_path()])


def probe_path() -> str:
    return os.path.join(HERE, "bw_probe")


def probe_cmd() -> list[str]:
    return [HIPCC, *[f for f in HIP_FLAGS if f != "-fPIC"], os.path.join(CSRC, "bw_probe.hip"),
            "-o", probe_path()]


def build_probe() -> None:
    _run(probe_cmd())


def build_ext() -> None:
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          f"-I{py_inc}", f"-I{pybind11.get_include()}", f"-I{INCLUDE}",
          os.path.join(CSRC, "bolt_py.cpp"), "-o", ext_path(),
          f"-L{HERE}", "-lbolt_hip", "-Wl,-rpath,$ORIGIN"])


def build() -> None:
    build_hip(extra=[probe_cmd()])
    build_ext()


if __name__ == "__main__":
    build()
    sys.exit(0)
