"""The wide fluid step (2 or 4 cells of the contiguous axis per lane, aligned vector accesses,
one-lane rotations across the wavefront for the streaming shift) against the one-cell-per-
lane step and the oracle.

Both kernels evaluate the same per-cell arithmetic, so every storage format must give
BIT-IDENTICAL populations and macroscopic fields; float storage is also compared bit-exactly
with the oracle.  The shapes put one, several or a fraction of a wavefront on a row (and
rows of a single lane), with walls, fixed cells and gravity, and unfused tracers (whose
coupling runs in the wide kernel's Boussinesq path).
"""

import numpy as np
import pytest

import oracle

from test_gpu_parity import make_pair, f32, omega_obj

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    oracle.build()


SHAPES_2D = [(7, 256), (5, 200), (3, 520), (4, 4), (3, 132), (2, 1024)]
SHAPES_3D = [(3, 4, 256), (2, 3, 68), (2, 2, 520), (3, 2, 8)]
OMEGAS = [("bgk", f32(1 / 0.6)), ("trt", f32(1 / 0.55), f32(1 / (0.25 / 0.05 + 0.5)))]


def _pair_sims(counts, om, storage, seed, gravity, vw):
    """Two Simulations with identical inputs: one cell per lane and `vw` cells per lane."""
    import boltzmann_amd
    rng = np.random.default_rng(seed)
    d = len(counts)
    flags = np.zeros(counts, np.int32)
    w = rng.random(counts) < 0.12
    flags[w] = 1
    fx = (rng.random(counts) < 0.05) & ~w
    flags[fx] = rng.choice([2, 4, 6], size=int(fx.sum()))
    rho = (1 + 0.02 * rng.standard_normal(counts)).astype(np.float32)
    vel = (0.04 * rng.standard_normal(counts + (d,))).astype(np.float32)
    sims = []
    for width in (1, vw):
        s = boltzmann_amd.Simulation("gpu", list(counts), omega_obj(om), storage=storage)
        s.set_vector_width(width)
        s.cells.flags[...] = flags
        s.fluid.rho[...] = rho
        s.fluid.vel[...] = vel
        if gravity:
            s.set_gravity(np.array([1e-5, -2e-5, 3e-6][:d], np.float32))
        sims.append(s)
    return sims, flags


def _same(a, b, fields=("f", "rho", "vel")):
    for name in fields:
        x, y = a.read_field(name, dtype="f64"), b.read_field(name, dtype="f64")
        assert np.array_equal(x, y, equal_nan=True), (name, np.nanmax(np.abs(x - y)))


@pytest.mark.parametrize("counts", SHAPES_2D + SHAPES_3D)
@pytest.mark.parametrize("storage", ["f32", "f32f", "f16s", "bf16s", "f64"])
@pytest.mark.parametrize("coll", [0, 1])
def test_wide_equals_one_cell_per_lane(counts, storage, coll):
    widths = [2] if storage == "f64" else [2, 4]
    for vw in widths:
        if counts[-1] % vw:
            continue
        (one, wide), flags = _pair_sims(counts, OMEGAS[coll], storage, seed=vw + coll, gravity=coll == 1,
                                        vw=vw)
        for n in (2, 4):
            one.iterate(n)
            wide.iterate(n)
            _same(one, wide)


@pytest.mark.parametrize("counts", [(5, 200), (3, 520), (2, 3, 68), (3, 4, 256)])
@pytest.mark.parametrize("om", OMEGAS)
def test_wide_f32_bit_exact_against_oracle(counts, om):
    q = 9 if len(counts) == 2 else 27
    for vw in (2, 4):
        sim, ref, flags = make_pair(counts, om, q, seed=3, gravity=[1e-5, -1e-5, 2e-6][:len(counts)])
        sim.set_vector_width(vw)
        nw = (flags & 1) == 0
        for _ in range(2):
            sim.iterate(4)
            ref.iterate(4)
            np.testing.assert_array_equal(sim.read_field("f"), ref.f)
            np.testing.assert_array_equal(sim.fluid.rho[nw], ref.rho[nw])
            np.testing.assert_array_equal(sim.fluid.vel[nw], ref.vel[nw])


@pytest.mark.parametrize("storage", ["f16s", "bf16s", "f32f"])
def test_wide_with_unfused_tracer_and_coupling(storage):
    # the shifted formats run their tracer in a separate kernel, so the fluid step is the
    # wide one, with the buoyancy read from the tracer's previous value
    from boltzmann_amd import core
    counts = (6, 256)
    (one, wide), flags = _pair_sims(counts, OMEGAS[1], storage, seed=11, gravity=False, vw=4)
    rng = np.random.default_rng(12)
    C = rng.random(counts).astype(np.float32)
    for s in (one, wide):
        t = s.add_tracer("T", core.bgk(0.7))
        t.val[...] = C
        s.add_boussinesq_coupling(t, 0.01, 0.5, np.array([0.0, -1e-3], np.float32))
    for _ in range(3):
        one.iterate(4)
        wide.iterate(4)
        _same(one, wide)
        np.testing.assert_array_equal(one.read_field("g", "T"), wide.read_field("g", "T"))


def test_vector_width_is_validated():
    import boltzmann_amd
    from boltzmann_amd.core import bgk
    s = boltzmann_amd.Simulation("gpu", [4, 30], bgk(0.6))
    with pytest.raises(boltzmann_amd.BoltError, match="multiple"):
        s.set_vector_width(4)
    with pytest.raises(boltzmann_amd.BoltError):
        s.set_vector_width(3)
    d = boltzmann_amd.Simulation("gpu", [4, 32], bgk(0.6), storage="f64")
    with pytest.raises(boltzmann_amd.BoltError, match="double"):
        d.set_vector_width(4)
    s.set_vector_width(2)
    s.set_vector_width(0)


@pytest.mark.parametrize("counts", [(7, 256), (5, 200), (3, 4, 256), (2, 3, 68)])
@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
def test_default_widths_equal_one_cell_per_lane(counts, storage):
    # the built-in choice mixes kernels (odd steps of the 2-byte formats at 2 cells per lane)
    (one, auto), flags = _pair_sims(counts, OMEGAS[1], storage, seed=21, gravity=True, vw=2)
    auto.set_vector_width(0)
    for n in (2, 4, 6):
        one.iterate(n)
        auto.iterate(n)
        _same(one, auto)
