"""What the trained network gains over the plain method (the learned warm start).

A reduced version of the steady-channel workload (128 x 64 channels, small obstacles, low
Reynolds numbers so every case is steady): the solver generates training states, the
network (depth-5 U-Net) is trained on the GPU with this library's kernels, and on held-out
geometries the plain LBM, which decides convergence, is run to the same FP64 residual from the cold
start and from the network's prediction (averaged with its mirror image).
The warm start must need fewer steps; a rejected prediction falls back to the cold start.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW = dict(nx=128, ny=64, re_range=(10.0, 40.0), size_range=(4, 10))


def test_residual_definition_on_a_known_decay():
    import boltzmann_amd
    from boltzmann_amd.core import bgk
    sim = boltzmann_amd.Simulation("gpu", [64, 32], bgk(0.8))
    sim.fluid.vel[..., 0] = 0.02 * np.sin(2 * np.pi * np.arange(32) / 32)[None, :]
    sim.iterate(2)
    assert sim.residual(100) == np.inf          # first call only stores the snapshot
    u0 = sim.read_field("vel", dtype="f64")
    sim.iterate(100)
    r = sim.residual(100)
    u1 = sim.read_field("vel", dtype="f64")
    want = np.sqrt(((u1 - u0) ** 2).sum(-1)).max() / (100 * np.sqrt((u1 ** 2).sum(-1)).max())
    assert r == pytest.approx(want, rel=1e-5)


def test_warm_start_needs_fewer_steps_than_cold_start():
    from boltzmann_amd import warmstart as ws
    cases, fields, steps = ws.generate(600, seed0=0, batch=300, max_steps=40000, check_every=1000,
                                       **KW)
    assert len(cases) >= 500, len(cases)
    model, losses = ws.train(cases, fields, epochs=60, batch=8, width=32, depth=5, lr=2e-3, seed=0)
    assert np.isfinite(losses).all() and losses[-1] < 0.05 * losses[0], losses
    held_out = [ws.make_case(50_000 + s, **KW) for s in range(32)]
    cmp = ws.compare(model, held_out, check_every=200, max_steps=40000, mirror=True, judge_at=400)
    s = cmp.summary()
    print("warm start", s, flush=True)
    assert s["cases"] >= 24
    # measured about 0.78 (0.68 with ten times the training data); a network that predicted
    # nothing useful scores about 1.0, and a rejected prediction falls back to its cold twin
    assert s["median_step_ratio"] < 0.85, s
    # the fallback is decided in the warm run itself: a dropped start's twin carries on
    assert (cmp.warm_steps[cmp.cold_steps > 0] > 0).all()


def test_coarse_prior_warm_start():
    # the network given the steady state of a 2x coarser lattice learns the correction to it
    from boltzmann_amd import warmstart as ws
    cases, fields, _ = ws.generate(400, seed0=0, batch=200, max_steps=40000, check_every=1000, **KW)
    model, losses = ws.train(cases, fields, epochs=40, batch=8, width=32, depth=5, lr=2e-3, seed=0,
                             prior_factor=2)
    assert model.net.features == 11 and np.isfinite(losses).all()
    held_out = [ws.make_case(50_000 + s, **KW) for s in range(32)]
    s = ws.compare(model, held_out, check_every=200, max_steps=40000, judge_at=400).summary()
    print("coarse-prior warm start", s, flush=True)
    assert s["cases"] >= 24
    assert s["median_step_ratio"] < 0.85, s
