"""boltzmann_amd: the lattice-Boltzmann update of djbarker/boltzmann on AMD MI355X.

The classes mirror the reference's ``boltzmann.boltzmann`` extension (src/lib.rs:22-411):
``Simulation``, ``Fluid``, ``Scalar``, ``Cells`` and ``Omega``; ``boltzmann_amd.core``
mirrors ``boltzmann.core`` (``bgk``, ``trt``, ``CellFlags``).  The update runs in HIP
kernels for gfx950 behind the C ABI of ``include/bolt.h`` (``libbolt_hip.so``).

Importing fails loudly if the native module has not been built
(``python boltzmann_amd/build.py``); there is no fallback implementation.
"""

try:
    from ._boltzmann import (  # noqa: F401
        Batch,
        BoltError,
        Cells,
        Fluid,
        Network,
        Omega,
        Scalar,
        Simulation,
        gemm_check,
        version,
    )
except ImportError as exc:  # pragma: no cover - exercised only on unbuilt trees
    raise ImportError(
        "boltzmann_amd's native module is not built; run `python boltzmann_amd/build.py` "
        f"(HIP/gfx950 toolchain required): {exc}"
    ) from exc

__all__ = ["Batch", "BoltError", "Cells", "Fluid", "Network", "Omega", "Scalar", "Simulation", "version"]
