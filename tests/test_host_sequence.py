"""The C ABI driven as the reference's Rust host would drive it (tests/native/host_sequence.cpp,
built with g++ against include/bolt.h and -lbolt_hip, following the call sequence of
src/simulation.rs:254-446 as INTEGRATION.md maps it).

* CPU: the program's first-iterate equilibrium, the product's host_init.hpp, equals the
  oracle's f32 equilibrium (velocities.rs:43-52) bit for bit, for the fluid and tracer
  velocity sets.
* GPU: the whole sequence -- create, tracer, gravity / coupling, host equilibrium, one write
  per host array, bolt_lattice_step, read-back -- gives the oracle's populations, rho, u
  and tracer values bit for bit.
"""

import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "boltzmann_amd")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libbolt_hip.so")):
        pytest.fail("libbolt_hip.so is not built (python boltzmann_amd/build.py)")
    exe = str(tmp_path_factory.mktemp("native") / "host_sequence")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall",
           f"-I{ROOT}/include", f"-I{ROOT}/boltzmann_amd/csrc",
           os.path.join(ROOT, "tests", "native", "host_sequence.cpp"),
           f"-L{LIBDIR}", "-lbolt_hip", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
           f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    oracle.build()
    return exe


CASES = {
    "d2q9_trt": dict(counts=(37, 29), q=9, coll=1, om=(1 / 0.55, 1 / (0.25 / 0.05 + 0.5)), tq=5),
    "d3q27_bgk": dict(counts=(9, 11, 14), q=27, coll=0, om=(1 / 0.7, 1 / 0.7), tq=7),
}


def write_inputs(d, case, gravity, coupling, steps, seed):
    counts = case["counts"]
    dim = len(counts)
    rng = np.random.default_rng(seed)
    flags = np.where(rng.random(counts) < 0.1, 1, 0).astype(np.int32)
    flags[(rng.random(counts) < 0.04) & (flags == 0)] = 6
    flags[(rng.random(counts) < 0.04) & (flags == 0)] |= 8
    rho = (1 + 0.01 * rng.standard_normal(counts)).astype(np.float32)
    vel = (0.03 * rng.standard_normal(counts + (dim,))).astype(np.float32)
    C = rng.random(counts).astype(np.float32)
    omp, omn = (float(np.float32(x)) for x in case["om"])
    tom = float(np.float32(1 / 0.8))
    g = [1e-5, -2e-5, 1e-5][:dim] + [0.0] * (3 - dim)
    cg = [0.0, -1e-3, 0.0][:dim] + [0.0] * (3 - dim)
    n = list(counts) + [1] * (3 - dim)
    meta = [dim, *n, case["q"], case["coll"], omp, omn, case["tq"], tom, int(gravity), *g,
            int(coupling), 0.02, 0.5, *cg, steps]
    with open(os.path.join(d, "meta.txt"), "w") as fh:
        fh.write(" ".join(repr(x) for x in meta))
    flags.tofile(os.path.join(d, "flags.i32"))
    rho.tofile(os.path.join(d, "rho.f32"))
    vel.tofile(os.path.join(d, "vel.f32"))
    C.tofile(os.path.join(d, "C.f32"))
    return dict(flags=flags, rho=rho, vel=vel, C=C, omp=omp, omn=omn, tom=tom,
                g=np.array(g[:dim], np.float32), cg=np.array(cg[:dim], np.float32))


@pytest.mark.parametrize("name", list(CASES))
def test_host_equilibrium_matches_oracle_bit_exactly(harness, tmp_path, name):
    case = CASES[name]
    inp = write_inputs(str(tmp_path), case, False, False, 2, seed=1)
    res = subprocess.run([harness, "equil", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    n = int(np.prod(case["counts"]))
    feq = np.fromfile(str(tmp_path / "feq.bin"), np.float32).reshape(case["counts"] + (case["q"],))
    geq = np.fromfile(str(tmp_path / "geq.bin"), np.float32).reshape(case["counts"] + (case["tq"],))
    assert feq.size == n * case["q"]
    np.testing.assert_array_equal(feq, oracle.equilibrate(case["q"], inp["rho"], inp["vel"]))
    np.testing.assert_array_equal(geq, oracle.equilibrate(case["tq"], inp["C"], inp["vel"]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("forcing", ["gravity", "coupling"])
def test_rust_host_sequence_matches_oracle(harness, tmp_path, name, forcing):
    case = CASES[name]
    steps = 10
    inp = write_inputs(str(tmp_path), case, forcing == "gravity", forcing == "coupling", steps, seed=2)
    res = subprocess.run([harness, "run", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    counts, q, tq = case["counts"], case["q"], case["tq"]
    om = ("bgk", inp["omp"]) if case["coll"] == 0 else ("trt", inp["omp"], inp["omn"])
    ref = oracle.OracleSimulation(counts, om, q=q)
    ref.flags[...] = inp["flags"]
    ref.rho[...] = inp["rho"]
    ref.vel[...] = inp["vel"]
    ref.add_tracer("T", inp["tom"], tq).C[...] = inp["C"]
    if forcing == "gravity":
        ref.set_gravity(inp["g"])
    else:
        ref.add_boussinesq_coupling("T", 0.02, 0.5, inp["cg"])
    ref.iterate(steps)
    rd = lambda nm, shp: np.fromfile(str(tmp_path / nm), np.float32).reshape(shp)
    nw = (inp["flags"] & 1) == 0
    np.testing.assert_array_equal(rd("f_out.bin", counts + (q,)), ref.f)
    np.testing.assert_array_equal(rd("g_out.bin", counts + (tq,)), ref.tracers["T"].g)
    np.testing.assert_array_equal(rd("rho_out.bin", counts)[nw], ref.rho[nw])
    np.testing.assert_array_equal(rd("vel_out.bin", counts + (len(counts),))[nw], ref.vel[nw])
    np.testing.assert_array_equal(rd("C_out.bin", counts)[nw], ref.tracers["T"].C[nw])
