"""Host-side parts of the warm start (no GPU): the consistent initial populations and the
case generator.

populations() is pinned by its moments: mass and momentum of the equilibrium, and the
first-order Chapman-Enskog stress Pi^(1) = -2 tau rho cs^2 S of a field with a known
constant strain rate S (Krueger et al., The Lattice Boltzmann Method, 2017, sec. 4.1).
"""

import numpy as np
import pytest

ws = pytest.importorskip("boltzmann_amd.warmstart")


def _linear_field(nx, ny, rho0=1.01, a=1e-3, b=5e-4, c=2e-4):
    rho = np.full((nx, ny), rho0, np.float32)
    vel = np.zeros((nx, ny, 2), np.float32)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    vel[..., 0] = 0.02 + a * Y + c * X          # du/dy = a, du/dx = c
    vel[..., 1] = b * X - c * Y                  # dv/dx = b, dv/dy = -c
    return rho, vel


def _case(nx=48, ny=32):
    c = ws.make_case(3, nx=nx, ny=ny, size_range=(4, 6))
    c.mask[:] = False
    return c


def test_populations_keep_mass_and_momentum():
    c = _case()
    rho, vel = _linear_field(c.nx, c.ny)
    f = ws.populations(c, rho, vel).astype(np.float64)
    assert f.shape == (c.nx, c.ny, 9)
    np.testing.assert_allclose(f.sum(-1), rho, rtol=0, atol=1e-6)
    np.testing.assert_allclose(f @ ws._C9, rho[..., None] * vel, rtol=0, atol=1e-7)


def test_populations_carry_the_chapman_enskog_stress():
    c = _case()
    a, b, cc, rho0 = 1e-3, 5e-4, 2e-4, 1.01
    rho, vel = _linear_field(c.nx, c.ny, rho0, a, b, cc)
    f = ws.populations(c, rho, vel).astype(np.float64)
    u = vel.astype(np.float64)
    cu = u @ ws._C9.T
    feq = ws._W9 * rho0 * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(-1)[..., None])
    fn = (f - feq)[5:-5, 5:-5]                   # interior: central differences are exact
    cx, cy = ws._C9[:, 0], ws._C9[:, 1]
    S = {"xx": cc, "yy": -cc, "xy": 0.5 * (a + b)}
    for name, (wx, wy) in {"xx": (cx * cx, 1), "yy": (cy * cy, 1), "xy": (cx, cy)}.items():
        pi = (fn * wx * wy).sum(-1)
        want = -2.0 * c.tau * rho0 / 3.0 * S[name]
        # populations are stored as float32 (about 3e-8 absolute on values near 0.4)
        np.testing.assert_allclose(pi, want, rtol=1e-5, atol=5e-8)


def test_populations_without_strain_are_the_equilibrium():
    c = _case()
    rho = np.full((c.nx, c.ny), 0.99, np.float32)
    vel = np.zeros((c.nx, c.ny, 2), np.float32)
    vel[..., 0] = 0.03
    f = ws.populations(c, rho, vel).astype(np.float64)
    u = np.array([0.03, 0.0])
    cu = ws._C9 @ u
    feq = ws._W9 * 0.99 * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * u @ u)
    np.testing.assert_allclose(f, np.broadcast_to(feq, f.shape), rtol=1e-6)


def test_cases_are_seeded_and_well_formed():
    a = ws.make_case(7, nx=128, ny=64, size_range=(4, 10))
    b = ws.make_case(7, nx=128, ny=64, size_range=(4, 10))
    assert a.u_in == b.u_in and a.tau == b.tau and (a.mask == b.mask).all()
    assert a.mask[:, 0].all() and a.mask[:, -1].all()
    assert not a.mask[0, 1:-1].any() and not a.mask[-1, 1:-1].any()   # open in/outflow
    assert 0.5 < a.tau and 0.02 <= a.u_in <= 0.05
    flags = a.flags()
    assert (flags[0, 1:-1] == 6).all() and (flags[-1, 1:-1] == 4).all()
    x = a.features()
    assert x.shape == (4, 128, 64) and np.isfinite(x).all()
    t = ws.to_targets(a, *a.cold_state())
    r, v = ws.from_targets(a, t)
    np.testing.assert_allclose(r, 1.0, atol=1e-6)
    np.testing.assert_allclose(v, a.cold_state()[1], atol=1e-7)
