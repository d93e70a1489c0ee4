"""The row-staged even step (one work-group per row of the contiguous axis, every population
access an aligned row, the shift along the row taken through LDS) against the even step that
addresses each neighbour's slot directly, and against the oracle.

Both evaluate the same per-cell arithmetic, so every storage format must give BIT-IDENTICAL
populations and macroscopic fields; float storage is also bit-exact with the oracle.  The
shapes cover rows shorter than a wavefront, rows that are not a multiple of 64, rows of one
cell (every shift wraps onto the cell itself), the longest rows the kernel takes (1024
cells) and longer ones, where the even step falls back to direct addressing -- with walls,
fixed cells, gravity and an unfused tracer with Boussinesq coupling.
"""

import numpy as np
import pytest

import oracle

from test_gpu_parity import make_pair, f32, omega_obj

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    oracle.build()


SHAPES_2D = [(7, 256), (5, 200), (3, 1024), (4, 4), (3, 1), (6, 37), (2, 1100)]
SHAPES_3D = [(3, 4, 256), (2, 3, 68), (2, 2, 1), (3, 2, 8), (2, 2, 520), (3, 3, 100)]
OMEGAS = [("bgk", f32(1 / 0.6)), ("trt", f32(1 / 0.55), f32(1 / (0.25 / 0.05 + 0.5)))]


def _pair(counts, om, storage, seed, gravity):
    """Two Simulations with identical inputs: the even step row-staged and direct."""
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
    for staged in (True, False):
        s = boltzmann_amd.Simulation("gpu", list(counts), omega_obj(om), storage=storage)
        s.set_row_staging(staged)
        s.cells.flags[...] = flags
        s.fluid.rho[...] = rho
        s.fluid.vel[...] = vel
        if gravity:
            s.set_gravity(np.array([1e-5, -2e-5, 3e-6][:d], np.float32))
        sims.append(s)
    return sims


def _same(a, b, fields=("f", "rho", "vel")):
    for name in fields:
        x, y = a.read_field(name, dtype="f64"), b.read_field(name, dtype="f64")
        assert np.array_equal(x, y, equal_nan=True), (name, np.nanmax(np.abs(x - y)))


@pytest.mark.parametrize("counts", SHAPES_2D + SHAPES_3D)
@pytest.mark.parametrize("storage", ["f32", "f32f", "f16s", "bf16s", "f64"])
@pytest.mark.parametrize("coll", [0, 1])
def test_row_staged_equals_direct(counts, storage, coll):
    staged, direct = _pair(counts, OMEGAS[coll], storage, seed=7 + coll, gravity=coll == 1)
    for n in (2, 4, 6):
        staged.iterate(n)
        direct.iterate(n)
        _same(staged, direct)


@pytest.mark.parametrize("counts", [(5, 200), (3, 1024), (6, 37), (2, 3, 68), (3, 4, 256), (2, 2, 1)])
@pytest.mark.parametrize("om", OMEGAS)
def test_row_staged_f32_bit_exact_against_oracle(counts, om):
    q = 9 if len(counts) == 2 else 27
    sim, ref, flags = make_pair(counts, om, q, seed=5, gravity=[1e-5, -1e-5, 2e-6][:len(counts)])
    sim.set_row_staging(True)
    nw = (flags & 1) == 0
    for _ in range(3):
        sim.iterate(4)
        ref.iterate(4)
        np.testing.assert_array_equal(sim.read_field("f"), ref.f)
        np.testing.assert_array_equal(sim.fluid.rho[nw], ref.rho[nw])
        np.testing.assert_array_equal(sim.fluid.vel[nw], ref.vel[nw])


@pytest.mark.parametrize("storage", ["f16s", "bf16s", "f32f"])
def test_row_staged_with_unfused_tracer_and_coupling(storage):
    # these formats run their tracer in a separate kernel, so the fluid's even step is the
    # row-staged one, with the buoyancy read from the tracer's previous value
    from boltzmann_amd import core
    counts = (6, 256)
    staged, direct = _pair(counts, OMEGAS[1], storage, seed=13, gravity=False)
    C = np.random.default_rng(14).random(counts).astype(np.float32)
    for s in (staged, direct):
        t = s.add_tracer("T", core.bgk(0.7))
        t.val[...] = C
        s.add_boussinesq_coupling(t, 0.01, 0.5, np.array([0.0, -1e-3], np.float32))
    for _ in range(3):
        staged.iterate(4)
        direct.iterate(4)
        _same(staged, direct)
        np.testing.assert_array_equal(staged.read_field("g", "T"), direct.read_field("g", "T"))


@pytest.mark.parametrize("storage", ["f16s", "f32"])
def test_row_staged_with_every_width(storage):
    # the row-staged even step beside the wide odd steps (the built-in mix of the 2-byte
    # formats) and with the width forced to one cell per lane: same bits as direct
    (staged, direct) = _pair((4, 3, 128), OMEGAS[1], storage, seed=17, gravity=True)
    for vw in (0, 1):
        staged.set_vector_width(vw)
        staged.iterate(4)
        direct.iterate(4)
        _same(staged, direct)
