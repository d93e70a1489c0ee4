"""Physics closed forms the update must reproduce, written once for two back ends: the
oracle (CPU) and the library's Simulation (GPU, any storage format).

* shear wave u_x = u0 sin(k y): amplitude decays as exp(-nu k^2 t), nu = (tau+ - 1/2) / 3;
* Taylor-Green vortex (3D, low Re): kinetic energy decays as exp(-6 nu k^2 t);
* Gaussian tracer at rest: each axis' variance grows by 2 kappa per step,
  kappa = cs^2 (tau - 1/2) with cs^2 = 1/3 (D2Q5) or 19/60 (D3Q7, weights 1/20, 19/120);
* Poiseuille flow between bounce-back walls at x = 1/2 and x = nx - 3/2 under a body force:
  u(x) = g'/(2 nu) (x - 1/2)(nx - 3/2 - x).  The reference shifts the equilibrium velocity
  by a / omega+ (lib.cl:253-256, 682-685) while TRT relaxes momentum at omega-, so the
  force that acts is g' = g omega- / omega+ (g itself for BGK).

Each function returns (measured, expected) or an error norm; the tolerances live in the
tests.
"""

import numpy as np


def f32(x):
    return float(np.float32(x))


def omega_of(coll, tau, magic=0.25):
    """("bgk", w) or ("trt", w+, w-) with tau- = magic / (tau+ - 1/2) + 1/2 (core.py:40-62)."""
    if coll == "bgk":
        return ("bgk", f32(1 / tau))
    return ("trt", f32(1 / tau), f32(1 / (magic / (tau - 0.5) + 0.5)))


class OracleBackend:
    def __init__(self, counts, om):
        import oracle
        self.sim = oracle.OracleSimulation(tuple(counts), om)

    def flags(self):
        return self.sim.flags

    def vel(self):
        return self.sim.vel

    def rho(self):
        return self.sim.rho

    def set_gravity(self, g):
        self.sim.set_gravity(np.asarray(g, np.float32))

    def add_tracer(self, tau):
        return self.sim.add_tracer("T", f32(1 / tau))

    def tracer_values(self, t):
        return t.C

    def iterate(self, n):
        self.sim.iterate(n)


class GpuBackend:
    def __init__(self, counts, om, storage="f32"):
        import boltzmann_amd
        from boltzmann_amd import Omega
        o = Omega.BGK(om[1]) if om[0] == "bgk" else Omega.TRT(om[1], om[2])
        self.sim = boltzmann_amd.Simulation("gpu", list(counts), o, storage=storage)

    def flags(self):
        return self.sim.cells.flags

    def vel(self):
        return self.sim.fluid.vel

    def rho(self):
        return self.sim.fluid.rho

    def set_gravity(self, g):
        self.sim.set_gravity(np.asarray(g, np.float32))

    def add_tracer(self, tau):
        from boltzmann_amd.core import bgk
        return self.sim.add_tracer("T", bgk(tau))

    def tracer_values(self, t):
        return t.val

    def iterate(self, n):
        self.sim.iterate(n)


def shear_wave(make, coll, tau=0.8, n=64, u0=0.01, t1=100, dt=1000):
    """Decay rate of a D2Q9 shear wave on an 8 x n periodic lattice."""
    om = omega_of(coll, tau)
    b = make((8, n), om)
    k = 2 * np.pi / n
    b.vel()[..., 0] = u0 * np.sin(k * np.arange(n))[None, :]
    b.iterate(t1)
    a1 = np.abs(b.vel()[..., 0].astype(np.float64)).max()
    b.iterate(dt)
    a2 = np.abs(b.vel()[..., 0].astype(np.float64)).max()
    nu = (1 / om[1] - 0.5) / 3
    return np.log(a1 / a2) / dt, nu * k * k


def taylor_green(make, coll, tau=0.8, n=48, u0=0.005, t1=50, dt=300):
    """Kinetic-energy decay rate of the 3D Taylor-Green vortex on an n^3 periodic lattice."""
    om = omega_of(coll, tau)
    b = make((n, n, n), om)
    k = 2 * np.pi / n
    i = np.arange(n)
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    v = b.vel()
    v[..., 0] = u0 * np.sin(k * X) * np.cos(k * Y) * np.cos(k * Z)
    v[..., 1] = -u0 * np.cos(k * X) * np.sin(k * Y) * np.cos(k * Z)
    b.rho()[...] = 1 + 3 * (u0 ** 2 / 16) * (np.cos(2 * k * X) + np.cos(2 * k * Y)) * (
        np.cos(2 * k * Z) + 2)
    b.iterate(t1)
    e1 = (b.vel().astype(np.float64) ** 2).sum()
    b.iterate(dt)
    e2 = (b.vel().astype(np.float64) ** 2).sum()
    nu = (1 / om[1] - 0.5) / 3
    return np.log(e1 / e2) / dt, 6 * nu * k * k


def gaussian_diffusion(make, d, tau_g=0.8, n=None, var0=4.0, t1=None, dt=None):
    """Per-axis variance growth per step of a Gaussian tracer in a fluid at rest."""
    n = n or (96 if d == 2 else 64)
    t1 = t1 or (200 if d == 2 else 50)
    dt = dt or (400 if d == 2 else 100)
    b = make((n,) * d, ("bgk", 1.0))
    t = b.add_tracer(tau_g)
    axes = np.meshgrid(*([np.arange(n) - n / 2] * d), indexing="ij")
    b.tracer_values(t)[...] = np.exp(-sum(a ** 2 for a in axes) / (2 * var0))

    def var():
        C = b.tracer_values(t).astype(np.float64)
        m = C.sum()
        return np.array([(C * a ** 2).sum() / m for a in axes])

    b.iterate(t1)
    v1 = var()
    b.iterate(dt)
    v2 = var()
    cs2 = 1 / 3 if d == 2 else 19 / 60
    return (v2 - v1) / dt, 2 * cs2 * (tau_g - 0.5)


def poiseuille(make, coll, tau=0.8, magic=0.25, nx=34, g=1e-5, steps=3000):
    """Max relative error of the steady profile between walls at rows 0 and nx-1."""
    om = omega_of(coll, tau, magic)
    b = make((nx, 2), om)
    b.flags()[0, :] = 1
    b.flags()[-1, :] = 1
    b.set_gravity([0.0, g])
    wp = om[1]
    wm = om[2] if om[0] == "trt" else om[1]
    nu = (1 / wp - 0.5) / 3
    x = np.arange(nx)
    ua = (g * wm / wp) / (2 * nu) * (x - 0.5) * (nx - 1.5 - x)
    b.vel()[1:-1, :, 1] = ua[1:-1, None]
    b.iterate(steps)
    u = b.vel()[1:-1, :, 1].astype(np.float64)
    return np.abs(u - ua[1:-1, None]).max() / ua.max()
