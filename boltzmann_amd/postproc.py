"""Post-processing of a simulation's velocity on the GPU.

Same names and results as the reference's ``boltzmann.utils.postproc``
(python/boltzmann/utils/postproc.py:12-127), computed from the velocity the device holds
after ``iterate`` instead of from a host array.  ``scales`` is any object with the
reference's ``Scales`` interface (``velocity.to_physical_units`` and
``converter(...).to_physical_units``); without it the results are in lattice units.
"""

from __future__ import annotations

import numpy as np

__all__ = ["calc_vmag", "calc_curl", "calc_stream_func"]


def calc_vmag(sim, scales=None) -> np.ndarray:
    """|u| per cell (postproc.py:12-32)."""
    out = sim.postproc("vmag")
    if scales is not None:
        out[:] = scales.velocity.to_physical_units(out)
    return out


def calc_curl(sim, scales=None) -> np.ndarray:
    """du/dy - dv/dx per cell in 2D, the three components in 3D (postproc.py:35-85)."""
    out = sim.postproc("curl")
    if scales is not None:
        out[:] = scales.converter(T=-1).to_physical_units(out)
    return out


def calc_stream_func(sim, scales=None, origin=None) -> np.ndarray:
    """2D stream function, zero at ``origin`` (cell indices, or distances with ``scales``;
    default (0, 0)) (postproc.py:88-127)."""
    origin = origin or (0, 0)
    if scales is not None:
        origin = (scales.distance.to_lattice_units(origin[0]),
                  scales.distance.to_lattice_units(origin[1]))
    out = sim.postproc("stream", origin=(int(origin[0]), int(origin[1])))
    if scales is not None:
        out[:] = scales.converter(L=2, T=-1).to_physical_units(out)
    return out
