"""CPU-side checks of the boundary: the C ABI library loads and exports every function
include/bolt.h declares; host logic that needs no GPU (argument checks, the CPU-device
refusal, the Python API surface of boltzmann_amd.core)."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bolt.h")
LIB = os.path.join(ROOT, "boltzmann_amd", "libbolt_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bolt_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libbolt_hip.so is not built (python boltzmann_amd/build.py)")
    return ctypes.CDLL(LIB)


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["bolt_ctx_create", "bolt_lattice_create", "bolt_lattice_step",
                 "bolt_lattice_write", "bolt_lattice_read", "bolt_lattice_add_tracer",
                 "bolt_lattice_add_coupling", "bolt_lattice_set_gravity", "bolt_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_cpu_device_is_refused_without_touching_a_gpu(lib):
    lib.bolt_ctx_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.bolt_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p(1)
    rc = lib.bolt_ctx_create(0, 0, ctypes.byref(out))  # BOLT_DEVICE_CPU
    assert rc == 2  # BOLT_ENOTFOUND
    assert out.value is None
    assert b"no CPU device" in lib.bolt_last_error()


def test_null_arguments_are_rejected(lib):
    lib.bolt_lattice_step.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.bolt_lattice_destroy.argtypes = [ctypes.c_void_p]
    lib.bolt_last_error.restype = ctypes.c_char_p
    assert lib.bolt_lattice_step(None, 2) == 1  # BOLT_EINVAL
    assert lib.bolt_lattice_destroy(None) == 0
    lib.bolt_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.bolt_version()


def test_python_api_surface():
    import boltzmann_amd
    from boltzmann_amd.core import CellFlags, Omega, bgk, trt
    # core.py:40-62
    assert repr(bgk(0.5 + 0.5)) == "Omega.BGK(1)"
    o = trt(0.6)
    assert o.kind == "TRT"
    wp, wn = o.values
    assert abs(wp - 1 / 0.6) < 1e-6
    assert abs(wn - 1 / (0.25 / 0.1 + 0.5)) < 1e-6
    o = trt(0.6, 0.7)
    assert abs(o.values[1] - 1 / 0.7) < 1e-6
    # simulation.rs:80-89 / core.py:68-81
    assert (CellFlags.FLUID, CellFlags.WALL, CellFlags.FIXED_FLUID_VELOCITY,
            CellFlags.FIXED_FLUID_DENSITY, CellFlags.FIXED_FLUID,
            CellFlags.FIXED_SCALAR_VALUE) == (0, 1, 2, 4, 6, 8)
    assert Omega.BGK(1.5).values == pytest.approx((1.5,))
    assert "gfx950" in boltzmann_amd.version()


def test_device_strings():
    import boltzmann_amd
    from boltzmann_amd.core import bgk
    with pytest.raises(boltzmann_amd.BoltError, match="Invalid device type"):
        boltzmann_amd.Simulation("tpu", [8, 8], bgk(0.6))
    with pytest.raises(boltzmann_amd.BoltError, match="no CPU device"):
        boltzmann_amd.Simulation("cpu", [8, 8], bgk(0.6))


def test_bench_refuses_more_than_one_gpu():
    # the measurement is one GPU's; the refusal happens before anything touches a device
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2, res
    assert "one GPU" in res.stderr or "single MI355X" in res.stderr, res.stderr
    assert res.stdout.strip() == ""
