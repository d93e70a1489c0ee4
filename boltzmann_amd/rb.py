"""Rayleigh-Benard parameter sweeps run as one batch.

A sweep is many independent 2D Rayleigh-Benard cells of uneven sizes: D2Q9 fluid plus a
D2Q5 temperature tracer coupled by Boussinesq buoyancy, the setup of the reference's
python/examples/rayleigh_benard/rayleigh_benard.py:50-72 (walls on the first and last row of
the contiguous axis, fixed-temperature rows next to them, gravity (0, -1) with a negative
coupling constant so warm fluid rises).  All instances advance together in one launch per
step (see ``boltzmann_amd.Batch``).

Parameters (lattice units): free-fall velocity u_ff = sqrt(g beta dT H) = 0.05 with
H = ny - 2 and dT = 1, so g beta = u_ff^2 / H; Pr = 1 gives nu = kappa = u_ff H sqrt(Pr / Ra)
and tau = 3 nu + 1/2 for both populations.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import Batch
from .core import CellFlags


@dataclass
class RBSpec:
    nx: int
    ny: int
    ra: float
    tau: float
    alpha: float
    seed: int

    @property
    def height(self) -> int:
        return self.ny - 2

    @property
    def kappa(self) -> float:
        return (self.tau - 0.5) / 3.0


def sweep(n: int = 64, seed: int = 4, u_ff: float = 0.05, pr: float = 1.0) -> list[RBSpec]:
    """n instances: nx in {128, 192, ..., 512}, ny in {128, 160, ..., 256}, Ra log-uniform
    in [1e4, 1e6], temperature noise seeds 1000 + k."""
    rng = np.random.default_rng(seed)
    nxs = np.arange(128, 513, 64)
    nys = np.arange(128, 257, 32)
    out = []
    for k in range(n):
        nx = int(rng.choice(nxs))
        ny = int(rng.choice(nys))
        ra = float(10 ** rng.uniform(4, 6))
        out.append(spec(nx, ny, ra, 1000 + k, u_ff, pr))
    return out


def spec(nx: int, ny: int, ra: float, seed: int = 0, u_ff: float = 0.05,
         pr: float = 1.0) -> RBSpec:
    """One cell of the sweep's parametrisation at a chosen size and Rayleigh number."""
    h = ny - 2
    nu = u_ff * h * np.sqrt(pr / ra)
    return RBSpec(nx, ny, ra, 3 * nu + 0.5, -(u_ff ** 2 / h), seed)


def initial_state(spec: RBSpec):
    """flags, temperature: walls at iy = 0 and ny-1, T = 1 fixed at iy = 1, T = 0 fixed at
    iy = ny-2, linear profile in between plus N(0, 1e-3) noise."""
    nx, ny = spec.nx, spec.ny
    flags = np.zeros((nx, ny), np.int32)
    flags[:, 0] = CellFlags.WALL
    flags[:, -1] = CellFlags.WALL
    flags[:, 1] = CellFlags.FIXED_SCALAR_VALUE
    flags[:, -2] = CellFlags.FIXED_SCALAR_VALUE
    y = np.arange(ny, dtype=np.float64)
    prof = np.clip(1.0 - (y - 1) / (ny - 3), 0.0, 1.0)
    rng = np.random.default_rng(spec.seed)
    T = prof[None, :] + rng.normal(0.0, 1e-3, (nx, ny))
    T[:, 1] = 1.0
    T[:, -2] = 0.0
    return flags, T.astype(np.float32)


def make_batch(specs: list[RBSpec], dev: str = "gpu", collision: str = "bgk") -> Batch:
    """Builds the batch, writes every instance's initial state and sets the populations to
    equilibrium."""
    b = Batch(dev, [(s.nx, s.ny) for s in specs], collision,
              [(1 / s.tau, 1 / s.tau) for s in specs], [1 / s.tau for s in specs],
              [s.alpha for s in specs], [0.5] * len(specs), [(0.0, -1.0)] * len(specs))
    for k, s in enumerate(specs):
        flags, T = initial_state(s)
        b.write(k, "flags", flags)
        b.write(k, "C", T)
    b.equilibrate()
    return b


def nusselt(T: np.ndarray, uy: np.ndarray, spec: RBSpec) -> float:
    """Nu = 1 + <u_y T> H / (kappa dT) over the fluid interior (dT = 1); 1 for conduction."""
    inner = (slice(None), slice(2, spec.ny - 2))
    return float(1.0 + np.mean(uy[inner] * T[inner]) * spec.height / spec.kappa)
