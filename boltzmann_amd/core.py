"""Collision constructors and cell flags, as the reference's ``boltzmann.core``
(python/boltzmann/core.py:40-81), re-exporting the native classes."""

from __future__ import annotations

from typing import Literal

from . import BoltError, Cells, Fluid, Omega, Scalar, Simulation  # noqa: F401

__all__ = ["bgk", "trt", "Device", "CellFlags", "Simulation", "Omega", "Fluid", "Scalar",
           "Cells", "BoltError"]


def bgk(tau: float) -> Omega:
    """BGK collision with relaxation time ``tau`` (core.py:40-47): Omega.BGK(1 / tau)."""
    return Omega.BGK(1.0 / tau)


def trt(tau_pos: float, tau_neg: float | None = None, magic_no: float = 0.25) -> Omega:
    """TRT collision (core.py:50-62).  ``tau_neg`` defaults to
    ``magic_no / (tau_pos - 0.5) + 0.5`` (magic number 1/4)."""
    tau_neg = tau_neg or (magic_no / (tau_pos - 0.5) + 0.5)
    return Omega.TRT(1.0 / tau_pos, 1.0 / tau_neg)


Device = Literal["cpu", "gpu"]


class CellFlags:
    """Cell flag bits (core.py:68-81, simulation.rs:80-89).  The update tests
    ``flag & 1`` (wall), ``flag & 6`` (fixed fluid: either bit fixes both density and
    velocity, as the reference's kernels do) and ``flag & 8`` (fixed scalar)."""

    FLUID = 0
    WALL = 1
    FIXED_FLUID_VELOCITY = 2
    FIXED_FLUID_DENSITY = 4
    FIXED_SCALAR_VALUE = 8
    FIXED_FLUID = FIXED_FLUID_VELOCITY | FIXED_FLUID_DENSITY
