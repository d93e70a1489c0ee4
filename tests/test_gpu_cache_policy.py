"""The two population load forms -- streaming (non-temporal) loads and ordinary loads, which
keep the lines in the Infinity Cache for the next step -- against each other and the oracle.

The load form changes no arithmetic, so every kernel (one cell per lane, the wide steps, the
row-staged even step, the fused and the separate tracer kernels, the batch kernels) must
give BIT-IDENTICAL populations and macroscopic fields under "stream", "keep" and "auto";
float storage is also bit-exact with the oracle under "keep".
"""

import numpy as np
import pytest

import oracle

from test_gpu_parity import make_pair, f32, omega_obj

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    oracle.build()


OMEGAS = [("bgk", f32(1 / 0.6)), ("trt", f32(1 / 0.55), f32(1 / (0.25 / 0.05 + 0.5)))]
POLICIES = ("keep", "stream")


def _sims(counts, om, storage, seed, gravity, policies=POLICIES):
    """One Simulation per load policy, with identical inputs (walls, fixed cells)."""
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
    for policy in policies:
        s = boltzmann_amd.Simulation("gpu", list(counts), omega_obj(om), storage=storage)
        s.set_cache_policy(policy)
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


# rows staged (whole cache lines), rows direct (200 and 37 cells), one-cell rows, and 3D
SHAPES = [(5, 256), (4, 200), (6, 37), (3, 1), (3, 4, 256), (2, 3, 68), (3, 3, 100)]


@pytest.mark.parametrize("counts", SHAPES)
@pytest.mark.parametrize("storage", ["f32", "f32f", "f16s", "bf16s", "f64"])
@pytest.mark.parametrize("coll", [0, 1])
def test_cached_loads_equal_streaming(counts, storage, coll):
    keep, stream = _sims(counts, OMEGAS[coll], storage, seed=31 + coll, gravity=coll == 1)
    for n in (2, 4, 6):
        keep.iterate(n)
        stream.iterate(n)
        _same(keep, stream)


@pytest.mark.parametrize("storage", ["f32", "f16s", "f64"])
@pytest.mark.parametrize("vw", [1, 2, 4])
@pytest.mark.parametrize("staged", [True, False])
def test_cached_loads_every_kernel_shape(storage, vw, staged):
    # every cells-per-lane width (wide even and odd steps) with the even step row-staged or
    # direct; double storage takes 2 cells per lane at most
    if storage == "f64" and vw == 4:
        pytest.skip("double storage runs 1 or 2 cells per lane")
    keep, stream = _sims((4, 3, 128), OMEGAS[1], storage, seed=37, gravity=True)
    for s in (keep, stream):
        s.set_vector_width(vw)
        s.set_row_staging(staged)
    for _ in range(3):
        keep.iterate(4)
        stream.iterate(4)
        _same(keep, stream)


@pytest.mark.parametrize("counts", [(5, 256), (6, 37), (2, 3, 68), (3, 4, 256)])
@pytest.mark.parametrize("om", OMEGAS)
@pytest.mark.parametrize("staged", [True, False])
def test_cached_loads_f32_bit_exact_against_oracle(counts, om, staged):
    q = 9 if len(counts) == 2 else 27
    sim, ref, flags = make_pair(counts, om, q, seed=41, gravity=[1e-5, -1e-5, 2e-6][:len(counts)])
    sim.set_cache_policy("keep")
    sim.set_row_staging(staged)
    nw = (flags & 1) == 0
    for _ in range(3):
        sim.iterate(4)
        ref.iterate(4)
        np.testing.assert_array_equal(sim.read_field("f"), ref.f)
        np.testing.assert_array_equal(sim.fluid.rho[nw], ref.rho[nw])
        np.testing.assert_array_equal(sim.fluid.vel[nw], ref.vel[nw])


@pytest.mark.parametrize("counts", [(6, 256), (2, 3, 68)])
def test_cached_loads_fused_tracer_f32_bit_exact_against_oracle(counts):
    q = 9 if len(counts) == 2 else 27
    d = len(counts)
    sim, ref, flags = make_pair(counts, OMEGAS[1], q, seed=43, tracers=(("T", f32(1 / 0.7)),),
                                couplings=(("T", f32(0.02), f32(0.5), [0.0, -1e-3, 0.0][:d]),), scalar=0.06)
    sim.set_cache_policy("keep")
    nw = (flags & 1) == 0
    for _ in range(3):
        sim.iterate(4)
        ref.iterate(4)
        np.testing.assert_array_equal(sim.read_field("f"), ref.f)
        np.testing.assert_array_equal(sim.read_field("g", "T"), ref.tracers["T"].g)
        np.testing.assert_array_equal(sim.get_tracer("T").val[nw], ref.tracers["T"].C[nw])


@pytest.mark.parametrize("storage", ["f32", "f64", "f16s", "bf16s"])
def test_cached_loads_with_tracers(storage):
    # float and double storage fuse the first tracer into the fluid step and run the second
    # in a kernel of its own; the 2-byte formats run both on their own
    from boltzmann_amd import core
    counts = (6, 256)
    keep, stream = _sims(counts, OMEGAS[1], storage, seed=47, gravity=False)
    rng = np.random.default_rng(48)
    C1, C2 = rng.random(counts).astype(np.float32), rng.random(counts).astype(np.float32)
    for s in (keep, stream):
        t = s.add_tracer("T", core.bgk(0.7))
        t.val[...] = C1
        u = s.add_tracer("S", core.bgk(0.8))
        u.val[...] = C2
        s.add_boussinesq_coupling(t, 0.01, 0.5, np.array([0.0, -1e-3], np.float32))
    for _ in range(3):
        keep.iterate(4)
        stream.iterate(4)
        _same(keep, stream)
        for name in ("T", "S"):
            np.testing.assert_array_equal(keep.read_field("g", name), stream.read_field("g", name))
            np.testing.assert_array_equal(keep.read_field("C", name), stream.read_field("C", name))


@pytest.mark.parametrize("counts", [(5, 256), (2, 3, 68)])
@pytest.mark.parametrize("storage", ["f32", "f16s"])
def test_cache_policies_agree(counts, storage):
    # "auto" (the default: by the step's working set) gives the same bits as both forms
    import boltzmann_amd
    sims = _sims(counts, OMEGAS[1], storage, seed=53, gravity=True, policies=("auto", "keep", "stream"))
    for s in sims:
        s.iterate(6)
    _same(sims[0], sims[1])
    _same(sims[0], sims[2])
    with pytest.raises(boltzmann_amd.BoltError):
        sims[0].set_cache_policy("sometimes")


@pytest.mark.parametrize("coll", ["bgk", "trt"])
@pytest.mark.parametrize("shape", [(64, 0), (128, 400)])
@pytest.mark.parametrize("staged", [True, False])
def test_batch_cached_loads_equal_streaming(coll, shape, staged):
    # uneven instances, direct launches and replayed graphs, an instance switched off and on
    from test_gpu_batch import specs, build_batch
    sp = specs(coll)
    a, b = build_batch(sp, coll), build_batch(sp, coll)
    a.set_cache_policy("keep")
    b.set_cache_policy("stream")
    for x in (a, b):
        x.set_row_staging(staged)
        x.set_shape(*shape)
        x.step(24)
        active = np.ones(len(sp), bool)
        active[1] = False
        x.set_active(active)
        x.step(10)
        x.set_active(np.ones(len(sp), bool))
        x.step(444)
    for k in range(len(sp)):
        for field in ("f", "g", "C", "vel", "rho"):
            np.testing.assert_array_equal(a.read(k, field), b.read(k, field))


def test_batch_policy_switch_between_graph_replays_matches_oracle():
    # the captured graph is rebuilt when the load form changes; each instance still equals
    # the oracle's run of it alone
    import boltzmann_amd
    from test_gpu_batch import specs, build_batch, oracle_run
    sp = specs("trt")
    b = build_batch(sp, "trt")
    b.set_shape(64, 100)
    b.set_cache_policy("stream")
    b.step(204)
    b.set_cache_policy("keep")
    b.step(204)
    b.set_cache_policy("auto")
    b.step(204)
    for k, s in enumerate(sp):
        ref = oracle_run(s, "trt", 612)
        np.testing.assert_array_equal(b.read(k, "f"), ref.f)
        np.testing.assert_array_equal(b.read(k, "g"), ref.tracers["T"].g)
    with pytest.raises(boltzmann_amd.BoltError):
        b.set_cache_policy("never")
