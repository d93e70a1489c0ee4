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


@pytest.mark.parametrize("counts", [(6, 256), (5, 200), (3, 1024), (4, 37), (2, 3, 68), (3, 2, 8)])
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_row_staged_fused_tracer(counts, storage):
    # float and double storage fuse the first tracer into the fluid step; its planes are
    # staged with the fluid's, with fixed-scalar cells and a Boussinesq coupling
    import boltzmann_amd
    from boltzmann_amd import core
    d = len(counts)
    rng = np.random.default_rng(19)
    flags = np.zeros(counts, np.int32)
    flags[rng.random(counts) < 0.1] = 1
    flags[(rng.random(counts) < 0.05) & (flags == 0)] = 6
    flags[(rng.random(counts) < 0.08) & (flags != 1)] |= 8
    vel = (0.03 * rng.standard_normal(counts + (d,))).astype(np.float32)
    C = rng.random(counts).astype(np.float32)
    sims = []
    for staged in (True, False):
        s = boltzmann_amd.Simulation("gpu", list(counts), omega_obj(OMEGAS[1]), storage=storage)
        s.set_row_staging(staged)
        s.cells.flags[...] = flags
        s.fluid.vel[...] = vel
        t = s.add_tracer("T", core.bgk(0.7))
        t.val[...] = C
        s.add_boussinesq_coupling(t, 0.02, 0.5, np.array([0.0, -1e-3, 0.0][:d], np.float32))
        sims.append(s)
    a, b = sims
    for n in (2, 4):
        a.iterate(n)
        b.iterate(n)
        _same(a, b)
        np.testing.assert_array_equal(a.read_field("g", "T", dtype="f64"), b.read_field("g", "T", dtype="f64"))
        np.testing.assert_array_equal(a.read_field("C", "T", dtype="f64"), b.read_field("C", "T", dtype="f64"))


@pytest.mark.parametrize("counts", [(6, 200), (3, 4, 100)])
def test_row_staged_fused_tracer_f32_bit_exact_against_oracle(counts):
    q = 9 if len(counts) == 2 else 27
    d = len(counts)
    sim, ref, flags = make_pair(counts, OMEGAS[1], q, seed=23, tracers=(("T", f32(1 / 0.7)),),
                                couplings=(("T", f32(0.02), f32(0.5), [0.0, -1e-3, 0.0][:d]),), scalar=0.06)
    sim.set_row_staging(True)
    nw = (flags & 1) == 0
    for _ in range(3):
        sim.iterate(4)
        ref.iterate(4)
        np.testing.assert_array_equal(sim.read_field("f"), ref.f)
        np.testing.assert_array_equal(sim.read_field("g", "T"), ref.tracers["T"].g)
        np.testing.assert_array_equal(sim.get_tracer("T").val[nw], ref.tracers["T"].C[nw])


@pytest.mark.parametrize("coll", ["bgk", "trt"])
@pytest.mark.parametrize("shape", [(64, 0), (128, 400)])
def test_batch_row_staged_equals_direct(coll, shape):
    # uneven instances (rows of 17 to 130 cells in one work-group size), direct launches and
    # replayed graphs, and an instance switched off and on again
    from test_gpu_batch import specs, build_batch
    sp = specs(coll)
    a, b = build_batch(sp, coll), build_batch(sp, coll)
    a.set_row_staging(True)
    b.set_row_staging(False)
    for x in (a, b):
        x.set_shape(*shape)
        x.step(24)
        active = np.ones(len(sp), bool)
        active[2] = False
        x.set_active(active)
        x.step(10)
        x.set_active(np.ones(len(sp), bool))
        x.step(444)
    for k in range(len(sp)):
        for field in ("f", "g", "C", "vel", "rho"):
            np.testing.assert_array_equal(a.read(k, field), b.read(k, field))


def test_batch_with_a_row_too_long_for_staging():
    # one instance's rows exceed 1024 cells: the batch's even steps stay direct; the other
    # shapes exercise one launch per row-length group when they are on their own
    import test_gpu_batch as tb
    shapes = tb.SHAPES
    try:
        for sh in ([(4, 1030), (5, 60)], [(3, 64), (4, 65), (2, 1024), (3, 200)]):
            tb.SHAPES = sh
            sp = tb.specs("trt")
            a, b = tb.build_batch(sp, "trt"), tb.build_batch(sp, "trt")
            a.set_row_staging(True)
            b.set_row_staging(False)
            for x in (a, b):
                x.step(30)
            for k, s in enumerate(sp):
                np.testing.assert_array_equal(a.read(k, "f"), b.read(k, "f"))
                np.testing.assert_array_equal(a.read(k, "g"), b.read(k, "g"))
                np.testing.assert_array_equal(a.read(k, "f"), tb.oracle_run(s, "trt", 30).f)
    finally:
        tb.SHAPES = shapes


@pytest.mark.parametrize("counts", [(5, 256), (5, 200), (2, 3, 1024), (2, 2, 64)])
@pytest.mark.parametrize("storage", ["f32", "f16s"])
def test_row_staging_modes_agree(counts, storage):
    # "auto" (the default: row-staged where it measured faster), on and off give the same bits
    import boltzmann_amd
    from boltzmann_amd import core
    d = len(counts)
    vel = (0.03 * np.random.default_rng(29).standard_normal(counts + (d,))).astype(np.float32)
    sims = []
    for mode in ("auto", True, False):
        s = boltzmann_amd.Simulation("gpu", list(counts), core.trt(0.6), storage=storage)
        s.set_row_staging(mode)
        s.fluid.vel[...] = vel
        sims.append(s)
    for s in sims:
        s.iterate(6)
    _same(sims[0], sims[1])
    _same(sims[0], sims[2])
    with pytest.raises(boltzmann_amd.BoltError):
        sims[0].set_row_staging("sometimes")


@pytest.mark.parametrize("ny", [48, 64])
def test_fluid_only_batch_row_staged_matches_oracle(ny):
    # the warm start's batches (D2Q9 TRT channels with walls, obstacles and fixed in/outflow
    # columns, no tracer): row-staged even steps against direct ones and the oracle
    from boltzmann_amd import warmstart as ws
    cases = [ws.make_case(91 + s, nx=40 + 8 * s, ny=ny, size_range=(4, 8)) for s in range(3)]
    runs = []
    for mode in (True, False):
        b = ws._batch(cases)
        b.set_row_staging(mode)
        for k, c in enumerate(cases):
            ws.load_state(b, k, c, *c.cold_state())
        b.equilibrate()
        b.step(30)
        runs.append(b)
    for k, c in enumerate(cases):
        wp, wn = (float(np.float32(w)) for w in c.omega)
        ref = oracle.OracleSimulation((c.nx, c.ny), ("trt", wp, wn))
        ref.flags[...] = c.flags()
        ref.rho[...], ref.vel[...] = c.cold_state()
        ref.iterate(30)
        for b in runs:
            np.testing.assert_array_equal(b.read(k, "f"), ref.f)
        nw = ~c.mask
        np.testing.assert_array_equal(runs[0].read(k, "vel")[nw], ref.vel[nw])
