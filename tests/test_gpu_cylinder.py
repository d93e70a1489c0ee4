"""Flow past a disc in a channel (D2Q9 TRT): the cylinder workload of the reference's
vortex-street example (python/examples/vortex_street/vortex_street.py:100-115) through the
product API, with fixed-velocity inflow, fixed-density outflow, bounce-back walls and disc.

* A reduced channel (256 x 64, disc r = 8) with the full configuration's relaxation rates
  (tau+ = 0.5192 from u = 0.05, r = 64, Re = 1000) is BIT-EXACT with the oracle in FP32
  storage for 300 steps, both with the magic-number TRT (tau- = 1/4 / (tau+ - 1/2) + 1/2,
  the default of core.py:40-62) and with tau- = 0.52 (the example's choice,
  vortex_street.py:100-102).
* The full 4096 x 1024 channel (disc r = 64 at (1024, 512), u = 0.05, Re = 1000 on the
  diameter) is run for 1000 steps in f64, f32, f16s and bf16s storage: every format stays
  finite, its velocity is within a stated relative max-norm of the f64 run and its total
  mass within a stated relative distance of the f64 run's (the channel is open, so mass is
  not conserved, but every format must carry the same in/outflow).
"""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

WALL, FIXED_FLUID, FIXED_DENSITY = 1, 6, 4


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    oracle.build()


def f32(x):
    return float(np.float32(x))


def channel(nx, ny, cx, cy, r, u0):
    """Flags and initial velocity of the channel: inflow column x = 0 holds (u0, 0) and
    rho = 1, the outflow column x = nx-1 holds rho = 1 (and, with the reference's c & 6
    rule, its initial velocity), rows y = 0 and y = ny-1 and the disc
    (x-cx)^2 + (y-cy)^2 < r^2 are walls, written last."""
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    flags = np.zeros((nx, ny), np.int32)
    flags[0, :] = FIXED_FLUID
    flags[-1, :] = FIXED_DENSITY
    flags[:, 0] = WALL
    flags[:, -1] = WALL
    flags[(X - cx) ** 2 + (Y - cy) ** 2 < r * r] = WALL
    vel = np.zeros((nx, ny, 2), np.float32)
    vel[..., 0] = u0
    vel[flags == WALL] = 0.0
    return flags, vel


def omegas(u0, r, re, tau_minus):
    nu = u0 * 2 * r / re
    tau_p = 3 * nu + 0.5
    tau_m = 0.25 / (tau_p - 0.5) + 0.5 if tau_minus == "magic" else tau_minus
    return f32(1 / tau_p), f32(1 / tau_m)


@pytest.mark.parametrize("tau_minus", ["magic", 0.52])
def test_reduced_cylinder_bit_exact(tau_minus):
    import boltzmann_amd
    from boltzmann_amd import core
    nx, ny, r, u0 = 256, 64, 8, 0.05
    omp, omn = omegas(u0, 64, 1000.0, tau_minus)  # the 4096 x 1024 channel's rates
    flags, vel = channel(nx, ny, 64, 32, r, u0)
    sim = boltzmann_amd.Simulation("gpu", [nx, ny], core.Omega.TRT(omp, omn))
    ref = oracle.OracleSimulation((nx, ny), ("trt", omp, omn), q=9)
    sim.cells.flags[...] = flags
    sim.fluid.vel[...] = vel
    ref.flags[...] = flags
    ref.vel[...] = vel
    nw = flags != WALL
    for _ in range(3):
        sim.iterate(100)
        ref.iterate(100)
        f = sim.read_field("f")
        assert np.isfinite(f).all()
        np.testing.assert_array_equal(f, ref.f)
        np.testing.assert_array_equal(sim.fluid.rho[nw], ref.rho[nw])
        np.testing.assert_array_equal(sim.fluid.vel[nw], ref.vel[nw])
    # the flow has developed: the disc's wake differs from the uniform start
    assert np.abs(sim.fluid.vel[..., 1]).max() > 1e-3


# relative max-norm of u against the f64 run after 1000 steps, and relative distance of the
# total mass, per tau- and storage format.  Measured with tau- = 0.52: f32 ~3e-5, f16s ~3e-3,
# bf16s ~3e-2 in u; the magic number's slowly relaxed odd moments (tau- = 13.5) carry the
# rounding further: f16s ~1.5e-2, bf16s ~2e-1.
BOUNDS = {
    0.52: {"f32": (1e-4, 1e-6), "f16s": (1e-2, 1e-4), "bf16s": (6e-2, 1e-3)},
    "magic": {"f32": (1e-4, 1e-6), "f16s": (3e-2, 1e-4), "bf16s": (3e-1, 1e-3)},
}


@pytest.mark.parametrize("tau_minus", ["magic", 0.52])
def test_full_cylinder_formats_against_f64(tau_minus):
    import boltzmann_amd
    from boltzmann_amd import core
    nx, ny, r, u0 = 4096, 1024, 64, 0.05
    omp, omn = omegas(u0, r, 1000.0, tau_minus)
    flags, vel = channel(nx, ny, 1024, 512, r, u0)
    res = {}
    for st in ("f64", "f32", "f16s", "bf16s"):
        sim = boltzmann_amd.Simulation("gpu", [nx, ny], core.Omega.TRT(omp, omn), storage=st)
        sim.cells.flags[...] = flags
        sim.fluid.vel[...] = vel
        sim.iterate(2)
        sim.step(998)
        assert sim.count_nonfinite("vel") == 0, st
        res[st] = (sim.read_field("vel", dtype="f64"), sim.reduce("mass"))
        del sim
    u64, m64 = res["f64"]
    assert np.isfinite(u64).all()
    assert np.abs(u64[..., 1]).max() > 1e-3  # the flow round the disc has developed
    errs = {st: (float(np.abs(res[st][0] - u64).max() / np.abs(u64).max()),
                 abs(res[st][1] - m64) / m64) for st in BOUNDS[tau_minus]}
    print(f"tau- {tau_minus}: u error and mass distance vs f64 after 1000 steps: {errs}")
    for st, (ub, mb) in BOUNDS[tau_minus].items():
        assert errs[st][0] < ub, (st, errs)
        assert errs[st][1] < mb, (st, errs)
