"""Batched small lattices: every instance of a batch must evolve exactly as it does alone.

Compared bit-for-bit against the oracle (the reference's step order: Boussinesq
acceleration from the previous tracer value, fluid update, tracer update) and against a
standalone Simulation, through both launch paths (direct launches and replayed HIP graphs),
before and after the autotuner, on uneven instance sizes that leave partial work-groups.
"""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_built():
    oracle.build()


def f32(x):
    return float(np.float32(x))


SHAPES = [(20, 17), (33, 70), (16, 40), (9, 130)]


def specs(coll):
    rng = np.random.default_rng(11)
    out = []
    for k, (nx, ny) in enumerate(SHAPES):
        tau = 0.55 + 0.1 * k
        om = (f32(1 / tau), f32(1 / (0.25 / (tau - 0.5) + 0.5)) if coll == "trt" else f32(1 / tau))
        flags = np.zeros((nx, ny), np.int32)
        flags[:, 0] = flags[:, -1] = 1
        flags[:, 1] |= 8
        flags[:, -2] |= 8
        flags[rng.random((nx, ny)) < 0.03] = 1
        T = rng.random((nx, ny)).astype(np.float32)
        vel = (0.02 * rng.standard_normal((nx, ny, 2))).astype(np.float32)
        out.append(dict(nx=nx, ny=ny, om=om, tom=f32(1 / (tau + 0.05)), alpha=f32(-1e-3 * (k + 1)),
                        c0=0.5, g=(0.0, -1.0), flags=flags, T=T, vel=vel))
    return out


def build_batch(sp, coll):
    import boltzmann_amd
    b = boltzmann_amd.Batch("gpu", [(s["nx"], s["ny"]) for s in sp], coll,
                            [(s["om"][0], s["om"][-1]) for s in sp], [s["tom"] for s in sp],
                            [s["alpha"] for s in sp], [s["c0"] for s in sp],
                            [s["g"] for s in sp])
    for k, s in enumerate(sp):
        b.write(k, "flags", s["flags"])
        b.write(k, "vel", s["vel"])
        b.write(k, "C", s["T"])
    b.equilibrate()
    return b


def oracle_run(s, coll, steps):
    om = ("bgk", s["om"][0]) if coll == "bgk" else ("trt", s["om"][0], s["om"][1])
    ref = oracle.OracleSimulation((s["nx"], s["ny"]), om)
    ref.flags[...] = s["flags"]
    ref.vel[...] = s["vel"]
    ref.add_tracer("T", s["tom"]).C[...] = s["T"]
    ref.add_boussinesq_coupling("T", s["alpha"], s["c0"], np.array(s["g"], np.float32))
    ref.iterate(steps)
    return ref


@pytest.mark.parametrize("coll", ["bgk", "trt"])
@pytest.mark.parametrize("shape", [(64, 4), (128, 2), (256, 400)])
def test_batch_matches_oracle_bit_exactly(coll, shape):
    sp = specs(coll)
    b = build_batch(sp, coll)
    b.set_shape(*shape)
    b.step(24)
    for k, s in enumerate(sp):
        ref = oracle_run(s, coll, 24)
        np.testing.assert_array_equal(b.read(k, "f"), ref.f)
        np.testing.assert_array_equal(b.read(k, "g"), ref.tracers["T"].g)
        np.testing.assert_array_equal(b.read(k, "C"), ref.tracers["T"].C)
        nw = (s["flags"] & 1) == 0
        np.testing.assert_array_equal(b.read(k, "vel")[nw], ref.vel[nw])


def test_batch_matches_standalone_simulation_and_tuner_keeps_state():
    import boltzmann_amd
    coll = "trt"
    sp = specs(coll)
    b = build_batch(sp, coll)
    report = b.tune(8)
    assert "chosen" in report
    b.step(30)
    for k, s in enumerate(sp):
        sim = boltzmann_amd.Simulation("gpu", [s["nx"], s["ny"]],
                                       boltzmann_amd.Omega.TRT(s["om"][0], s["om"][1]))
        sim.cells.flags[...] = s["flags"]
        sim.fluid.vel[...] = s["vel"]
        t = sim.add_tracer("T", boltzmann_amd.Omega.BGK(s["tom"]))
        t.val[...] = s["T"]
        sim.add_boussinesq_coupling(t, s["alpha"], s["c0"], np.array(s["g"], np.float32))
        sim.iterate(30)
        np.testing.assert_array_equal(b.read(k, "f"), sim.read_field("f"))
        np.testing.assert_array_equal(b.read(k, "C"), t.val)


def test_rb_sweep_small_runs_and_conserves():
    from boltzmann_amd import rb
    sp = rb.sweep(n=6, seed=4)
    b = rb.make_batch(sp)
    m0 = [b.reduce(k, "mass") for k in range(len(sp))]
    b.step(200)
    for k, s in enumerate(sp):
        # walls and periodic sides: fluid mass is conserved (fixed-scalar rows do not touch it)
        assert abs(b.reduce(k, "mass") - m0[k]) / m0[k] < 1e-6
        T = b.read(k, "C")
        assert np.isfinite(T).all()
        np.testing.assert_array_equal(T[:, 1], 1.0)
        np.testing.assert_array_equal(T[:, -2], 0.0)


def test_inactive_instances_keep_their_state_and_the_others_run_on():
    coll = "trt"
    sp = specs(coll)
    b = build_batch(sp, coll)
    b.step(10)
    frozen = b.read(1, "f").copy()
    active = np.ones(len(sp), bool)
    active[1] = False
    b.set_active(active)
    b.step(14)
    np.testing.assert_array_equal(b.read(1, "f"), frozen)            # not advanced
    for k in (0, 2):
        np.testing.assert_array_equal(b.read(k, "f"), oracle_run(sp[k], coll, 24).f)
    b.set_active(np.zeros(len(sp), bool))
    b.step(4)                                                        # nothing to advance
    np.testing.assert_array_equal(b.read(1, "f"), frozen)
    b.set_active(np.ones(len(sp), bool))
    b.step(14)                                                       # instance 1 resumes
    np.testing.assert_array_equal(b.read(1, "f"), oracle_run(sp[1], coll, 24).f)


def test_rayleigh_benard_onset_and_nusselt():
    # aspect-2 cells (one roll pair of wavelength 2H, near the critical wavenumber 3.12):
    # below the rigid-wall critical Rayleigh number 1708 the noise decays and heat is only
    # conducted (Nu = 1); above it convection sets in and carries heat, more at higher Ra
    from boltzmann_amd import rb
    specs = [rb.spec(256, 130, ra, seed=2000 + k) for k, ra in enumerate((1000.0, 1e4, 1e5))]
    b = rb.make_batch(specs)
    nus = []
    for steps in (300_000, 100_000):
        b.step(steps)
        nus.append([rb.nusselt(b.read(k, "C"), b.read(k, "vel")[..., 1], s)
                    for k, s in enumerate(specs)])
    (sub, lo, hi), (sub2, lo2, hi2) = nus
    assert abs(sub2 - 1.0) < 1e-3, nus
    assert abs(lo2 / lo - 1) < 1e-2 and abs(hi2 / hi - 1) < 1e-2, nus   # steady rolls
    assert 2.0 < lo2 < 3.5 and 4.0 < hi2 < 6.5 and hi2 > lo2, nus


def test_fluid_only_batch_matches_oracle_bit_exactly():
    # the warm start's batches: D2Q9 TRT channels with walls, obstacles and fixed in/outflow
    # columns, no tracer (the batch kernel's fluid-only instantiation)
    from boltzmann_amd import warmstart as ws
    cases = [ws.make_case(77 + s, nx=64, ny=48, size_range=(4, 8)) for s in range(3)]
    b = ws._batch(cases)
    for k, c in enumerate(cases):
        ws.load_state(b, k, c, *c.cold_state())
    b.equilibrate()
    b.step(30)
    for k, c in enumerate(cases):
        wp, wn = (float(np.float32(w)) for w in c.omega)
        ref = oracle.OracleSimulation((c.nx, c.ny), ("trt", wp, wn))
        ref.flags[...] = c.flags()
        ref.rho[...], ref.vel[...] = c.cold_state()
        ref.iterate(30)
        np.testing.assert_array_equal(b.read(k, "f"), ref.f)
        nw = ~c.mask
        np.testing.assert_array_equal(b.read(k, "vel")[nw], ref.vel[nw])


def test_batched_residuals_equal_per_instance_residuals():
    # one launch for every active instance against bolt_batch_residual one by one, on two
    # identical batches; maxima do not depend on the order of the reduction, so the values
    # are equal, and an inactive instance reports NaN
    sp = specs("trt")
    a, b = build_batch(sp, "trt"), build_batch(sp, "trt")
    for x in (a, b):
        x.step(20)
    ra = a.residuals(0)
    rb = [b.residual(k, 0) for k in range(len(sp))]
    assert np.isinf(ra).all() and np.isinf(rb).all()
    for x in (a, b):
        x.step(40)
    ra = a.residuals(40)
    rb = np.array([b.residual(k, 40) for k in range(len(sp))])
    assert np.isfinite(ra).all() and (ra > 0).all()
    np.testing.assert_array_equal(ra, rb)
    active = np.array([True, False, True, True])
    a.set_active(active)
    b.set_active(active)
    for x in (a, b):
        x.step(20)
    ra = a.residuals(20)
    assert np.isnan(ra[1])
    rb = np.array([b.residual(k, 20) for k in (0, 2, 3)])
    np.testing.assert_array_equal(ra[[0, 2, 3]], rb)
