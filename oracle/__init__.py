"""TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.

The oracle: a CPU restatement of the reference's lattice update (djbarker/boltzmann,
``src/lib.cl``) and of its host step loop (``src/simulation.rs:278-446``).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
package, and only as the checker.  ``boltzmann_amd`` never imports it.

``bolt_oracle.cpp`` holds the per-cell kernels (built by ``oracle/Makefile`` into
``libbolt_oracle.so``); :class:`OracleSimulation` below restates the host loop:

* ``iterate(n)``: ``n`` must be even (simulation.rs:279-281); on the first call the
  fluid populations are set to equilibrium from rho/vel and every tracer's from its
  value and the fluid velocity, in f32 (simulation.rs:283-286, fields.rs:106-114,185-193);
  the acceleration buffer is allocated (zeros) if gravity or a coupling exists
  (simulation.rs:289-295); gravity is written once before the loop
  (simulation.rs:326-342); per step: every Boussinesq coupling overwrites the
  acceleration from its tracer's current value (simulation.rs:353-373), then the fluid
  update (even = 1 - step % 2, simulation.rs:350,376-408), then every tracer update
  reading the fluid velocity just written (simulation.rs:410-432).

Every function here is pinned by tests/test_oracle_*.py (see bolt_oracle.cpp's header for
the list of checks).  Parity of the checkpoint byte format is not covered here.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbolt_oracle.so")
_lib = None


def build() -> None:
    """Compile the oracle library in place (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32, f32, f64, i64 = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_int64
        L.orc_update_ns.argtypes = [i32, P, i32, i32, i32, f32, f32, P, P, P, P, i32, P]
        L.orc_update_ad.argtypes = [i32, P, i32, i32, f32, P, P, P, P]
        L.orc_update_ns_f64.argtypes = [i32, P, i32, i32, i32, f64, f64, P, P, P, P, i32, P]
        L.orc_update_ad_f64.argtypes = [i32, P, i32, i32, f64, P, P, P, P]
        L.orc_set_constant_acc.argtypes = [i32, P, P, P]
        L.orc_set_boussinesq_acc.argtypes = [i32, P, P, P, f32, f32, P]
        L.orc_equilibrate.argtypes = [i32, i32, i64, P, P, P]
        L.orc_velocity_set.argtypes = [i32, i32, P, P, P]
        for name in ("orc_update_ns", "orc_update_ad", "orc_update_ns_f64",
                     "orc_update_ad_f64", "orc_equilibrate", "orc_velocity_set",
                     "orc_num_threads"):
            getattr(L, name).restype = ctypes.c_int
        L.orc_set_constant_acc.restype = None
        L.orc_set_boussinesq_acc.restype = None
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def velocity_set(d: int, q: int):
    """(c[q,3] int32, w[q] float32, opp[q] int32) as held by the oracle."""
    c = np.zeros((q, 3), np.int32)
    w = np.zeros(q, np.float32)
    opp = np.zeros(q, np.int32)
    rc = lib().orc_velocity_set(d, q, _p(c), _p(w), _p(opp))
    if rc != 0:
        raise ValueError(f"Unknown model: D{d}Q{q}")
    return c[:, :d].copy(), w, opp


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---- single kernels --------------------------------------------------------------------

def update_ns(counts, q, omega, even, f, rho, vel, flags, acc=None):
    """One AA update of the fluid, in place.  ``omega`` = ("bgk", w) or ("trt", w+, w-)."""
    d = len(counts)
    s = np.asarray(counts, np.int32)
    coll = 0 if omega[0] == "bgk" else 1
    wp = omega[1]
    wn = omega[2] if coll else omega[1]
    if f.dtype == np.float64:
        rc = lib().orc_update_ns_f64(d, _p(s), q, coll, int(even), wp, wn, _p(f), _p(rho),
                                     _p(vel), _p(acc), int(acc is not None), _p(flags))
    else:
        rc = lib().orc_update_ns(d, _p(s), q, coll, int(even), wp, wn, _p(f), _p(rho),
                                 _p(vel), _p(acc), int(acc is not None), _p(flags))
    if rc != 0:
        raise ValueError(f"Kernel not found: D{d}Q{q} NS")


def update_ad(counts, q, omega, even, g, val, vel, flags):
    d = len(counts)
    s = np.asarray(counts, np.int32)
    if g.dtype == np.float64:
        rc = lib().orc_update_ad_f64(d, _p(s), q, int(even), omega, _p(g), _p(val), _p(vel),
                                     _p(flags))
    else:
        rc = lib().orc_update_ad(d, _p(s), q, int(even), omega, _p(g), _p(val), _p(vel),
                                 _p(flags))
    if rc != 0:
        raise ValueError(f"Kernel not found: D{d}Q{q} AD")


def equilibrate(q, rho, vel) -> np.ndarray:
    """Host f32 equilibrium of velocities.rs:43-52 for every cell."""
    d = vel.shape[-1]
    ncell = rho.size
    f = np.zeros(rho.shape + (q,), np.float32)
    rho = np.ascontiguousarray(rho, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    rc = lib().orc_equilibrate(d, q, ncell, _p(rho), _p(vel), _p(f))
    if rc != 0:
        raise ValueError(f"Unknown model: D{d}Q{q}")
    return f


# ---- host loop ---------------------------------------------------------------------------

@dataclass
class _Tracer:
    q: int
    omega: float
    g: np.ndarray
    C: np.ndarray


@dataclass
class _Coupling:
    tracer: str
    alpha: float
    c0: float
    gravity: np.ndarray


@dataclass
class OracleSimulation:
    """Restatement of ``Simulation`` (simulation.rs:130-446) on host arrays.

    ``omega`` is ("bgk", w) or ("trt", w+, w-).  ``dtype`` float32 follows the reference's
    mixed-precision kernels; float64 runs the same update in double throughout.
    """

    counts: tuple
    omega: tuple
    q: int | None = None
    dtype: type = np.float32
    flags: np.ndarray = field(init=False)
    f: np.ndarray = field(init=False)
    rho: np.ndarray = field(init=False)
    vel: np.ndarray = field(init=False)
    acc: np.ndarray | None = field(init=False, default=None)
    tracers: dict = field(init=False, default_factory=dict)
    couplings: list = field(init=False, default_factory=list)
    gravity: np.ndarray | None = field(init=False, default=None)
    iteration: int = field(init=False, default=0)

    def __post_init__(self):
        self.counts = tuple(int(c) for c in self.counts)
        d = len(self.counts)
        if self.q is None:
            self.q = {2: 9, 3: 27}[d]
        self.flags = np.zeros(self.counts, np.int32)
        self.f = np.zeros(self.counts + (self.q,), self.dtype)
        self.rho = np.ones(self.counts, self.dtype)
        self.vel = np.zeros(self.counts + (d,), self.dtype)

    @property
    def d(self) -> int:
        return len(self.counts)

    def add_tracer(self, name: str, omega: float, q: int | None = None) -> _Tracer:
        if name not in self.tracers:
            q = q or {2: 5, 3: 7}[self.d]
            self.tracers[name] = _Tracer(q, omega, np.zeros(self.counts + (q,), self.dtype),
                                         np.zeros(self.counts, self.dtype))
        return self.tracers[name]

    def set_gravity(self, g) -> None:
        self.gravity = np.asarray(g, np.float32).copy()

    def add_boussinesq_coupling(self, tracer: str, alpha: float, c0: float, g) -> None:
        assert tracer in self.tracers
        self.couplings.append(_Coupling(tracer, float(alpha), float(c0),
                                        np.asarray(g, np.float32).copy()))

    def _equilibrate(self) -> None:
        if self.dtype == np.float32:
            self.f[...] = equilibrate(self.q, self.rho, self.vel)
            for t in self.tracers.values():
                t.g[...] = equilibrate(t.q, t.C, self.vel)
        else:
            self.f[...] = _feq_f64(self.q, self.rho, self.vel)
            for t in self.tracers.values():
                t.g[...] = _feq_f64(t.q, t.C, self.vel)

    def iterate(self, n: int) -> None:
        if n % 2 != 0:
            raise ValueError("iters must be even")
        if self.iteration == 0:
            self._equilibrate()
        d = self.d
        if (self.gravity is not None or self.couplings) and self.acc is None:
            self.acc = np.zeros(self.counts + (d,), self.dtype)
        if self.gravity is not None:
            self.acc[...] = self.gravity.astype(self.dtype)
        f32 = self.dtype == np.float32
        for it in range(n):
            even = 1 - it % 2
            for cp in self.couplings:
                C = self.tracers[cp.tracer].C
                if f32:
                    s = np.asarray(self.counts, np.int32)
                    lib().orc_set_boussinesq_acc(d, _p(s), _p(self.acc), _p(C), cp.alpha,
                                                 cp.c0, _p(cp.gravity))
                else:
                    self.acc[...] = cp.gravity.astype(np.float64) * (
                        cp.alpha * (C[..., None] - cp.c0))
            update_ns(self.counts, self.q, self.omega, even, self.f, self.rho, self.vel,
                      self.flags, self.acc)
            for t in self.tracers.values():
                update_ad(self.counts, t.q, t.omega, even, t.g, t.C, self.vel, self.flags)
        self.iteration += n


def _feq_f64(q, rho, vel):
    d = vel.shape[-1]
    c, w, _ = velocity_set(d, q)
    wd = _weights_f64(d, q)
    cu = np.einsum("...d,qd->...q", vel.astype(np.float64), c.astype(np.float64))
    uu = np.sum(vel.astype(np.float64) ** 2, axis=-1)[..., None]
    return wd * rho[..., None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu)


def _weights_f64(d, q):
    from fractions import Fraction
    table = {
        (2, 9): ["4/9"] + ["1/9"] * 4 + ["1/36"] * 4,
        (2, 5): ["1/3"] + ["1/6"] * 4,
        (3, 7): ["1/20"] + ["19/120"] * 6,
        (3, 27): ["8/27"] + ["2/27"] * 6 + ["1/54"] * 12 + ["1/216"] * 8,
    }
    return np.array([float(Fraction(x)) for x in table[(d, q)]])
