"""Steady channel flows past obstacles, solved in batches, and the learned warm start.

Case (lattice units): D2Q9 TRT (magic number 1/4) channel of nx x ny cells (x along the
channel, y across, y contiguous); walls on y = 0 and y = ny-1; the first column holds the
inflow (fixed fluid: rho = 1, u = (u_in, 0)); the last column the outflow (fixed density
rho = 1; like every fixed cell of the reference's kernels it also keeps its initial
velocity, lib.cl:257-262); 1 to 8 circles or rectangles of radius / half-size 8..32 as
walls.  Re = u_in D / nu on D = twice the largest obstacle size sets tau = 3 nu + 1/2.

The plain solver decides convergence: a run has converged when the FP64 residual
max|u^n - u^(n-100)| / (100 max|u^n|) drops below the tolerance (1e-6).
"""

from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import Batch
from .core import CellFlags


@dataclass
class Case:
    seed: int
    nx: int
    ny: int
    u_in: float
    re: float
    tau: float
    mask: np.ndarray = field(repr=False)   # bool [nx, ny]: obstacle or wall

    @property
    def omega(self):
        wp = 1.0 / self.tau
        wn = 1.0 / (0.25 / (self.tau - 0.5) + 0.5)
        return (wp, wn)

    def flags(self) -> np.ndarray:
        f = np.zeros((self.nx, self.ny), np.int32)
        f[0, :] = CellFlags.FIXED_FLUID
        f[-1, :] = CellFlags.FIXED_FLUID_DENSITY
        f[self.mask] = CellFlags.WALL
        return f

    def cold_state(self):
        """rho = 1, u = (u_in, 0) outside walls (the usual plug-flow start)."""
        rho = np.ones((self.nx, self.ny), np.float32)
        vel = np.zeros((self.nx, self.ny, 2), np.float32)
        vel[..., 0] = self.u_in
        vel[self.mask] = 0.0
        return rho, vel

    def features(self) -> np.ndarray:
        """Network input [8, nx, ny] (the channels of features())."""
        return features([self])[0].transpose(2, 0, 1)


def make_case(seed: int, nx: int = 512, ny: int = 256, re_range=(10.0, 100.0),
              size_range=(8, 32)) -> Case:
    rng = np.random.default_rng(seed)
    u_in = float(rng.uniform(0.02, 0.05))
    nobs = int(rng.integers(1, 9))
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    mask = np.zeros((nx, ny), bool)
    mask[:, 0] = mask[:, -1] = True
    lo, hi = size_range
    lo = max(2, min(lo, ny // 6))
    hi = max(lo + 1, min(hi, ny // 4))
    biggest = 0
    for _ in range(nobs):
        r = int(rng.integers(lo, hi + 1))
        biggest = max(biggest, r)
        cx = int(rng.integers(nx // 8 + r, int(nx * 0.6) - r))
        cy = int(rng.integers(r + 4, ny - r - 4))
        if rng.random() < 0.5:
            mask |= (X - cx) ** 2 + (Y - cy) ** 2 < r * r
        else:
            b = int(rng.integers(lo, hi + 1))
            mask |= (np.abs(X - cx) < r) & (np.abs(Y - cy) < b)
    re = float(rng.uniform(*re_range))
    nu = u_in * 2 * biggest / re
    return Case(seed, nx, ny, u_in, re, 3 * nu + 0.5, mask)


def _batch(cases, dev="gpu") -> Batch:
    return Batch(dev, [(c.nx, c.ny) for c in cases], "trt", [c.omega for c in cases])


def load_state(b: Batch, k: int, case: Case, rho: np.ndarray, vel: np.ndarray):
    """Writes flags and an initial (rho, u); fixed cells keep their prescribed values.
    Returns the (rho, u) written."""
    cold_rho, cold_vel = case.cold_state()
    flags = case.flags()
    fixed = (flags & 6) != 0
    rho = np.where(fixed, cold_rho, rho).astype(np.float32)
    vel = np.where(fixed[..., None], cold_vel, vel).astype(np.float32)
    vel[case.mask] = 0.0
    b.write(k, "flags", flags)
    b.write(k, "rho", rho)
    b.write(k, "vel", vel)
    return rho, vel


# D2Q9 velocities, weights and opposites in the reference's slot order (velocities.rs:66-83)
_C9 = np.array([(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1)],
               np.float64)
_W9 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
_OPP9 = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7])


def _derivative(u: np.ndarray, wall: np.ndarray, axis: int) -> np.ndarray:
    """du/dx_axis by central differences; next to a bounce-back wall (u = 0 half a cell
    away) from the quadratic through the far neighbour, the cell and the wall."""
    up, um = np.roll(u, -1, axis), np.roll(u, 1, axis)
    wp, wm = np.roll(wall, -1, axis), np.roll(wall, 1, axis)
    d = 0.5 * (up - um)
    d = np.where(wp & ~wm, -u - um / 3.0, d)
    d = np.where(wm & ~wp, u + up / 3.0, d)
    return np.where(wp & wm, 0.0, d)


def populations(case: Case, rho: np.ndarray, vel: np.ndarray) -> np.ndarray:
    """D2Q9 populations [nx, ny, 9] for a given (rho, u), as the lattice stores them after
    an even number of AA steps: post-collision and not yet streamed.

    Fluid cells: equilibrium plus the first-order Chapman-Enskog non-equilibrium part
    after collision, (1 - omega+) f^(1) with f^(1)_i = -tau+ w_i rho Q_i:S / cs^2,
    Q_i = c_i c_i - cs^2 I and S the strain rate (Skordos 1993; Mei et al. 2006).  Fixed
    cells collide with omega = 1, so they hold the equilibrium.  A wall slot i holds the
    population opp(i) that the fluid neighbour at x + c_i sent into the wall (full-way
    bounce-back).  A state started this way carries its viscous stress instead of relaxing
    to it, which would launch sound waves that the fixed inflow and outflow columns reflect
    back and forth."""
    wall = case.mask
    fixed = (case.flags() & 6) != 0
    r = rho.astype(np.float64)
    u = vel.astype(np.float64).copy()
    u[wall] = 0.0
    cu = u @ _C9.T
    uu = (u ** 2).sum(-1)[..., None]
    f = _W9 * r[..., None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * uu)
    dux_dx, dux_dy = (_derivative(u[..., 0], wall, a) for a in (0, 1))
    duy_dx, duy_dy = (_derivative(u[..., 1], wall, a) for a in (0, 1))
    sxy = 0.5 * (dux_dy + duy_dx)
    cx, cy = _C9[:, 0], _C9[:, 1]
    qs = ((cx ** 2 - 1 / 3) * dux_dx[..., None] + (cy ** 2 - 1 / 3) * duy_dy[..., None]
          + 2 * cx * cy * sxy[..., None])
    f1 = -3.0 * case.tau * _W9 * r[..., None] * qs
    f1[fixed | wall] = 0.0
    f += (1.0 - 1.0 / case.tau) * f1
    for i in range(1, 9):
        shift = (-int(cx[i]), -int(cy[i]))
        src = np.roll(f[..., _OPP9[i]], shift, axis=(0, 1))
        take = wall & ~np.roll(wall, shift, axis=(0, 1))
        f[..., i] = np.where(take, src, f[..., i])
    return f.astype(np.float32)


# ---- the coarse-lattice prior ------------------------------------------------------------
def coarse_case(case: Case, factor: int = 2) -> Case:
    """The same channel on a lattice `factor` times coarser: a coarse cell is a wall when at
    least half of its fine cells are, the channel walls stay, u_in and Re are kept, so the
    viscosity in coarse lattice units is nu / factor."""
    nxc, nyc = case.nx // factor, case.ny // factor
    m = case.mask[:nxc * factor, :nyc * factor].reshape(nxc, factor, nyc, factor).mean(axis=(1, 3)) >= 0.5
    m[:, 0] = m[:, -1] = True
    nu = (case.tau - 0.5) / 3.0 / factor
    return Case(case.seed, nxc, nyc, case.u_in, case.re, 3 * nu + 0.5, m)


def upsample(case: Case, factor: int, rho: np.ndarray, vel: np.ndarray):
    """Bilinear interpolation of a coarse (rho, u) to the fine cell centres of `case`
    (clamped at the lattice edges); u = 0 in the fine walls.  Non-finite coarse values (a
    coarse run that blew up) give the cold start instead."""
    if not (np.isfinite(rho).all() and np.isfinite(vel).all()):
        return case.cold_state()
    nxc, nyc = rho.shape

    def axis(n_f, n_c):
        t = (np.arange(n_f) + 0.5) / factor - 0.5
        i0 = np.clip(np.floor(t).astype(np.int64), 0, n_c - 2)
        return i0, np.clip(t - i0, 0.0, 1.0)

    x0, tx = axis(case.nx, nxc)
    y0, ty = axis(case.ny, nyc)

    def up(a):
        a = a.astype(np.float64)
        TX, TY = tx[:, None], ty[None, :]
        if a.ndim == 3:
            TX, TY = TX[..., None], TY[..., None]
        return ((1 - TX) * (1 - TY) * a[x0][:, y0] + TX * (1 - TY) * a[x0 + 1][:, y0]
                + (1 - TX) * TY * a[x0][:, y0 + 1] + TX * TY * a[x0 + 1][:, y0 + 1])

    r = up(rho).astype(np.float32)
    v = up(vel).astype(np.float32)
    v[case.mask] = 0.0
    return r, v


def coarse_prior(cases, factor: int = 2, tol: float = 1e-5, check_every: int = 500,
                 max_steps: int = 40000, dev: str = "gpu"):
    """Steady states of the cases on a `factor` times coarser lattice, upsampled to the
    cases' lattice: [(rho, vel)] per case.  The plain solver runs them in one batch to tol
    (a case that does not get there within max_steps, or stalls, gives its last state)."""
    if not cases:
        return []
    cc = [coarse_case(c, factor) for c in cases]
    _, fields = solve(cc, tol=tol, check_every=check_every, max_steps=max_steps, dev=dev,
                      drop_stalled=8)
    return _each(lambda k: upsample(cases[k], factor, *fields[k]), range(len(cases)))


@dataclass
class SolveResult:
    steps: np.ndarray        # steps to reach tol (-1: not reached)
    residual: np.ndarray     # last residual per case
    seconds: float
    first_residual: np.ndarray
    history: np.ndarray = None   # [checks, cases] residual at every check
    check_every: int = 0
    steps_run: int = 0           # steps the batch ran
    case_steps: int = 0          # steps summed over the cases while they were in the batch

    @property
    def seconds_per_case_step(self) -> float:
        """Wall time of one step of one case in this batch, set-up and residual checks
        included."""
        return self.seconds / max(1, self.case_steps)


def solve(cases, init=None, tol=1e-6, check_every=1000, max_steps=100000, dev="gpu",
          batch: Batch | None = None, stop_when_done=True, keep_fields=True, neq=False,
          drop_stalled: int = 0, judge_twins_at: int = 0):
    """Runs every case to the residual tolerance in one batch.

    init: None for the cold start, else a list of (rho, vel) per case (None entries start
    cold); with neq the initial populations of the given states carry the non-equilibrium
    part of populations(), else they are at equilibrium.  SolveResult.seconds covers the
    set-up (state upload, initial populations) and the stepping.  With stop_when_done a
    converged case leaves the batch (its state is kept); with drop_stalled > 0 so does a
    case whose residual has not halved over that many checks (an unsteady flow; its steps
    stay -1).

    judge_twins_at > 0: the batch holds n cases twice, instances k (given start) and
    n + k (its twin).  At the first check at or after that many steps, of each pair the
    one with the larger residual leaves the batch (a converged one always stays);
    SolveResult.rejected[k] is True where instance k lost to its twin and judged_at is the
    step of that check.  Every step both twins ran is counted in case_steps.

    Returns the SolveResult and, with keep_fields, the final (rho, vel) of every case."""
    t0 = time.perf_counter()
    b = batch or _batch(cases, dev)
    written = []
    for k, c in enumerate(cases):
        if init is None or init[k] is None:
            written.append(load_state(b, k, c, *c.cold_state()))
        else:
            written.append(load_state(b, k, c, *init[k]))
    b.equilibrate()
    if neq and init is not None:
        for k, c in enumerate(cases):
            if init[k] is not None:
                b.write(k, "f", populations(c, *written[k]))
    n = len(cases)
    steps = np.full(n, -1, np.int64)
    last = np.full(n, np.inf)
    first = np.full(n, np.nan)
    hist = []
    done = 0
    active = np.ones(n, bool)
    best = np.full(n, np.inf)          # lowest residual so far, and the check it was halved at
    best_at = np.zeros(n, np.int64)
    checks = 0
    case_steps = 0
    rejected = np.zeros(n, bool)
    judged_at = 0
    npair = n // 2 if judge_twins_at > 0 else 0
    timings = {"set_up": time.perf_counter() - t0, "step": 0.0, "residual": 0.0, "set_active": 0.0}
    while done < max_steps and active.any():
        case_steps += int(active.sum()) * check_every
        t1 = time.perf_counter()
        b.step(check_every - 100)
        t2 = time.perf_counter()
        b.residuals(0)                # snapshot u^(n-100) of every active case
        t3 = time.perf_counter()
        b.step(100)
        done += check_every
        checks += 1
        t4 = time.perf_counter()
        rs = b.residuals(100)
        timings["step"] += (t2 - t1) + (t4 - t3)
        timings["residual"] += (t3 - t2) + (time.perf_counter() - t4)
        for k in np.flatnonzero(active):
            r = rs[k]
            last[k] = r
            if np.isnan(first[k]):
                first[k] = r
            if steps[k] < 0 and r <= tol:
                steps[k] = done
            if r < 0.5 * best[k]:
                best[k], best_at[k] = r, checks
        hist.append(last.copy())
        leave = np.zeros(n, bool)
        if npair and not judged_at and done >= judge_twins_at:
            judged_at = done
            for k in range(npair):
                w, c = k, npair + k
                if steps[w] >= 0:
                    lose = c
                elif steps[c] >= 0:
                    lose = w
                else:
                    lose = w if last[w] > last[c] else c
                rejected[lose] = True
                leave[lose] = True
        if stop_when_done:
            leave |= steps >= 0
        if drop_stalled > 0:
            leave |= (steps < 0) & (checks - best_at >= drop_stalled)
        if (active & leave).any():
            active &= ~leave
            t1 = time.perf_counter()
            b.set_active(active)
            timings["set_active"] += time.perf_counter() - t1
    secs = time.perf_counter() - t0
    res = SolveResult(steps, last, secs, first, np.array(hist), check_every, done, case_steps)
    res.rejected = rejected
    res.judged_at = judged_at
    res.timings = timings  # wall seconds: set-up, stepping (b.step: the launches, synchronous
    #                       only when a later call waits), residual checks, active-set changes
    if not keep_fields:
        return res, None
    fields = [(b.read(k, "rho"), b.read(k, "vel")) for k in range(n)]
    return res, fields


# ---- learned warm start ----------------------------------------------------------------
def to_targets(case: Case, rho: np.ndarray, vel: np.ndarray) -> np.ndarray:
    """Network targets [nx, ny, 3]: u / u_in (two components) and (rho - 1) / u_in^2."""
    t = np.empty((case.nx, case.ny, 3), np.float32)
    t[..., 0] = vel[..., 0] / case.u_in
    t[..., 1] = vel[..., 1] / case.u_in
    t[..., 2] = (rho - 1.0) / case.u_in ** 2
    return t


def from_targets(case: Case, t: np.ndarray):
    rho = (1.0 + t[..., 2] * case.u_in ** 2).astype(np.float32)
    vel = (t[..., :2] * case.u_in).astype(np.float32)
    vel[case.mask] = 0.0
    return rho, vel


def _each(fn, items) -> list:
    """fn over items on a thread pool (the numpy copies and scipy's distance transform release
    the GIL); results in order."""
    items = list(items)
    if len(items) < 2:
        return [fn(i) for i in items]
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1, len(items))) as ex:
        return list(ex.map(fn, items))


def features(cases, prior=None) -> np.ndarray:
    """[n, nx, ny, 8] channels-last network inputs, per cell:
    0 wall mask, 1 u_in / 0.05, 2 (tau - 1/2) * 2, 3 Re / 100,
    4 x / nx (along the channel), 5 y / ny (across),
    6 chessboard distance to the nearest wall cell / 16, 7 the laminar channel profile
    6 s (1 - s), s = (y - 1/2) / (ny - 2), of mean 1 between the channel walls, times
    (ny - 2) / (open cells of the column): a mass-conserving guess of u / u_in.
    With prior (a (rho, vel) per case, e.g. coarse_prior()), three more channels
    [n, nx, ny, 11]: 8, 9 the prior's u / u_in and 10 its (rho - 1) / u_in^2."""
    from scipy.ndimage import distance_transform_cdt
    nx, ny = cases[0].nx, cases[0].ny
    nf = 8 if prior is None else 11
    out = np.empty((len(cases), nx, ny, nf), np.float32)
    xs = (np.arange(nx, dtype=np.float32) / nx)[:, None]
    ys = (np.arange(ny, dtype=np.float32) / ny)[None, :]
    sy = np.clip((np.arange(ny) - 0.5) / (ny - 2), 0.0, 1.0)
    prof = (6 * sy * (1 - sy)).astype(np.float32)

    def fill(k):
        # channel planes first, then one interleaving copy (3.5x faster than eight strided
        # channel writes)
        c = cases[k]
        assert (c.nx, c.ny) == (nx, ny), "every case of one call has the same size"
        planes = np.empty((nf, nx, ny), np.float32)
        planes[0] = c.mask
        planes[1] = c.u_in / 0.05
        planes[2] = (c.tau - 0.5) * 2.0
        planes[3] = c.re / 100.0
        planes[4] = xs
        planes[5] = ys
        planes[6] = distance_transform_cdt(~c.mask, metric="chessboard") / 16.0
        open_cells = np.maximum((~c.mask).sum(1), 1).astype(np.float32)
        lam = prof[None, :] * ((ny - 2) / open_cells)[:, None]
        planes[7] = np.where(c.mask, 0.0, lam)
        if prior is not None:
            t = to_targets(c, *prior[k])
            planes[8:11] = t.transpose(2, 0, 1)
        np.copyto(out[k], planes.transpose(1, 2, 0))

    _each(fill, range(len(cases)))
    return out


def mirrored_features(x: np.ndarray) -> np.ndarray:
    """features() of the reflected cases from those of the cases: every channel reflects
    with the geometry except y / ny (channel 5), which belongs to the lattice, and the
    prior's u_y (channel 9), which also changes sign."""
    m = np.ascontiguousarray(x[:, :, ::-1])
    m[..., 5] = x[..., 5]
    if x.shape[-1] > 9:
        m[..., 9] *= -1
    return m


def generate(n: int, seed0: int = 0, batch: int = 64, max_steps: int = 60000,
             check_every: int = 2000, tol: float = 1e-6, log=None, drop_stalled: int = 8,
             **case_kw):
    """Solves n cases (seeds seed0 ..) to the tolerance in batches; returns the converged
    cases, their steady (rho, vel) and the steps they took.  Converged cases leave their
    batch, and so do cases whose residual stalls for drop_stalled checks (unsteady flows,
    which would be dropped anyway)."""
    cases, fields, steps = [], [], []
    for b0 in range(0, n, batch):
        cs = [make_case(seed0 + s, **case_kw) for s in range(b0, min(n, b0 + batch))]
        res, fl = solve(cs, tol=tol, check_every=check_every, max_steps=max_steps,
                        drop_stalled=drop_stalled)
        for c, f, st in zip(cs, fl, res.steps):
            if st >= 0 and np.isfinite(f[1]).all():
                cases.append(c)
                fields.append(f)
                steps.append(int(st))
        if log:
            log(f"generated {b0 + len(cs)}/{n} cases, {len(cases)} converged, "
                f"{res.seconds:.1f} s for this batch")
    return cases, fields, np.array(steps)


class Predictor:
    """The trained network and the per-channel normalisation of its targets: the network
    learns (t - mean) / std of each target channel of to_targets(), or, with a coarse
    prior (prior_factor > 0), of the difference between the targets and the prior's
    (the network then corrects the upsampled coarse solution it is given as input)."""

    def __init__(self, net, mean: np.ndarray, std: np.ndarray, prior_factor: int = 0,
                 prior_tol: float = 1e-5):
        self.net = net
        self.mean = np.asarray(mean, np.float32)
        self.std = np.asarray(std, np.float32)
        self.prior_factor = prior_factor
        self.prior_tol = prior_tol

    def predict(self, x: np.ndarray) -> np.ndarray:
        """Targets [n, nx, ny, 3] (as to_targets) for inputs [n, nx, ny, 8 or 11]."""
        t = self.net.predict(x) * self.std + self.mean
        if self.prior_factor:
            t = t + x[..., 8:11]
        return t

    def predict_cases(self, cases, mirror: bool = False) -> np.ndarray:
        """Targets for cases (with a prior model, the coarse solves run first); with
        mirror, the mean of the prediction and the reflected prediction of the reflected
        case (exact symmetry, so only the network's error differs between the two)."""
        prior = coarse_prior(cases, self.prior_factor, self.prior_tol) if self.prior_factor else None
        x = features(cases, prior)
        t = self.predict(x)
        if not mirror:
            return t
        tm = self.predict(mirrored_features(x))[:, :, ::-1].copy()
        tm[..., 1] *= -1
        return 0.5 * (t + tm)


def mirrored(case: Case, rho: np.ndarray, vel: np.ndarray):
    """The case reflected across the channel (y -> ny - 1 - y) and its steady state: D2Q9,
    the walls and the fixed columns are symmetric under the reflection, so the reflected
    field (u_y negated) is the reflected case's steady state."""
    c = Case(case.seed, case.nx, case.ny, case.u_in, case.re, case.tau,
             np.ascontiguousarray(case.mask[:, ::-1]))
    v = np.ascontiguousarray(vel[:, ::-1])
    v[..., 1] *= -1
    return c, (np.ascontiguousarray(rho[:, ::-1]), v)


def train(cases, fields, epochs: int = 30, batch: int = 8, width: int = 32, lr: float = 1e-3,
          seed: int = 0, dev: str = "gpu", log=None, log_every: int = 1, depth: int = 5,
          mirror: bool = True, rho_weight: float = 1 / 9, time_budget: float | None = None,
          prior_factor: int = 0, prior_tol: float = 1e-5):
    """Trains the network on solver-generated steady states with the library's kernels.
    Returns the Predictor and the per-epoch mean losses (on normalised targets).

    mirror adds each case reflected across the channel (an exact extra sample).  rho_weight
    scales the density channel's share of the loss (its errors cost the warm start far
    fewer steps than velocity errors).  time_budget (seconds) caps the run: after the first
    epoch the epoch count, and with it the cosine schedule, shrinks to what fits.  With log,
    training runs log_every epochs at a time (same schedule as one call) and reports each
    chunk's last epoch loss.  prior_factor > 0: every case's steady state is also solved on
    a lattice that many times coarser (coarse_prior, to prior_tol), the upsampled result is
    an input and the network learns the correction to it."""
    from . import Network
    c0 = cases[0]
    prior = coarse_prior(cases, prior_factor, prior_tol, dev=dev) if prior_factor else None
    net = Network(dev, c0.nx, c0.ny, batch, width=width, seed=seed, depth=depth,
                  features=8 if prior is None else 11)
    if mirror:
        extra = [mirrored(c, *f) for c, f in zip(cases, fields)]
        if prior is not None:
            prior = list(prior) + [mirrored(c, *pf)[1] for c, pf in zip(cases, prior)]
        cases = list(cases) + [c for c, _ in extra]
        fields = list(fields) + [f for _, f in extra]
    x = features(cases, prior)
    y = np.stack([to_targets(c, *f) for c, f in zip(cases, fields)])
    if prior is not None:
        y = y - x[..., 8:11]
    mean = y.reshape(-1, 3).mean(0)
    std = y.reshape(-1, 3).std(0) + 1e-6
    std[2] /= np.sqrt(rho_weight)
    y = ((y - mean) / std).astype(np.float32)
    net.set_data(x, y)
    del x, y
    model = Predictor(net, mean, std, prior_factor, prior_tol)
    if log is None and time_budget is None:
        return model, net.train(epochs, lr=lr, seed=seed)
    losses = []
    t0 = time.perf_counter()
    e0 = 0
    while e0 < epochs:
        k = 1 if e0 == 0 else min(log_every, epochs - e0)
        losses.extend(net.train(k, lr=lr, seed=seed, first=e0, total=epochs).tolist())
        e0 += k
        spent = time.perf_counter() - t0
        if e0 == 1 and time_budget is not None:
            epochs = max(1, min(epochs, int(time_budget / spent)))
        if log:
            log(f"epoch {e0}/{epochs}: loss {losses[-1]:.5f} ({spent:.1f} s)")
    return model, np.array(losses, np.float32)


@dataclass
class Comparison:
    cold_steps: np.ndarray
    warm_steps: np.ndarray       # steps of the start that was kept (its cold twin's after a fallback)
    fallback: np.ndarray         # warm start dropped (its residual was worse at the judging check)
    cold_seconds: float          # the cold batch's wall time
    warm_seconds: float          # inference + the warm batch (both twins until the judging check)
    inference_seconds: float
    cold_case_seconds: np.ndarray = None   # per case: steps to tol x time per case-step
    warm_case_seconds: np.ndarray = None   # the same warm (both twins' steps) plus its share of inference

    def summary(self) -> dict:
        """Over the cases the cold run brings to the tolerance.  The batch wall times run
        until every case has converged or max_steps, so a case that never converges (an
        unsteady flow) sets both; the per-case times do not depend on the other cases."""
        ok = self.cold_steps > 0
        ratio = self.warm_steps[ok] / self.cold_steps[ok]
        out = {"cases": int(ok.sum()), "median_step_ratio": float(np.median(ratio)),
               "mean_step_ratio": float(np.mean(ratio)), "fallbacks": int(self.fallback.sum()),
               "cold_seconds": round(self.cold_seconds, 3),
               "warm_seconds": round(self.warm_seconds, 3),
               "batch_wall_ratio": round(self.warm_seconds / self.cold_seconds, 3),
               "inference_seconds": round(self.inference_seconds, 3)}
        if getattr(self, "twin_steps", None) is not None:
            charged = (self.warm_steps[ok] + self.twin_steps[ok]) / self.cold_steps[ok]
            out["median_charged_step_ratio"] = float(np.median(charged))
        if self.cold_case_seconds is not None:
            w, c = self.warm_case_seconds[ok], self.cold_case_seconds[ok]
            out["time_to_residual_ratio"] = round(float(w.sum() / c.sum()), 4)
            out["median_case_time_ratio"] = round(float(np.median(w / c)), 4)
        return out


def compare(net, cases, tol: float = 1e-6, check_every: int = 500, max_steps: int = 100000,
            judge_at: int = 1000, dev: str = "gpu", neq: bool = False,
            mirror: bool = False) -> Comparison:
    """Cold start vs learned warm start on the same cases, both judged by the plain solver.

    The warm start is the network's (rho, u) (with mirror, averaged with the reflected
    prediction of the reflected case: about 1.5 % fewer steps for twice the inference, a net
    loss in time) and, with neq, the non-equilibrium part of populations().  neq pays when
    the prediction is very close (a converged state restarts in a third of the steps); at
    the network's present accuracy it changes the step count by under 2 % and costs host
    time, so it is off by default.

    The warm run decides its own fallback, as a deployment would: each case runs from the
    prediction and, beside it in the same batch, from the cold start; at the first check
    at or after judge_at steps the start with the larger residual is dropped (the
    north_star's "cold start if the initial residual is worse").  Every step of both twins
    is charged to the warm start.  A separate, complete cold run gives the reference
    times.  Both runs are timed from the set-up of their batch; the warm run also counts
    feature construction and inference."""
    n = len(cases)
    cold, _ = solve(cases, tol=tol, check_every=check_every, max_steps=max_steps, dev=dev,
                    keep_fields=False)
    t0 = time.perf_counter()
    pred = net.predict_cases(cases, mirror) if hasattr(net, "predict_cases") else \
        net.predict(features(cases))
    init = _each(lambda k: from_targets(cases[k], pred[k]), range(n))
    t_inf = time.perf_counter() - t0
    warm, _ = solve(list(cases) + list(cases), init=list(init) + [None] * n, tol=tol,
                    check_every=check_every, max_steps=max_steps, dev=dev, keep_fields=False,
                    neq=neq, judge_twins_at=judge_at)
    fallback = warm.rejected[:n]
    steps = np.where(fallback, warm.steps[n:], warm.steps[:n])
    # the dropped twin's steps: until the judging check, or its own convergence before it
    lost = np.where(fallback, warm.steps[:n], warm.steps[n:])
    judged = warm.judged_at if warm.judged_at else warm.steps_run
    extra = np.where((lost > 0) & (lost < judged), lost, judged)
    # a warm run that misses the tolerance where the cold run reached it counts max_steps
    steps = np.where((steps < 0) & (cold.steps > 0), max_steps, steps)
    # per-case time to the tolerance: its steps (both twins') at the measured cost of one
    # case-step, plus its share of the inference for the warm start
    per_cold = cold.seconds_per_case_step
    per_warm = warm.seconds_per_case_step
    cold_case = np.where(cold.steps > 0, cold.steps * per_cold, np.nan)
    warm_case = np.where(steps > 0, (steps + extra) * per_warm + t_inf / n, np.nan)
    cmp = Comparison(cold.steps, steps, fallback, cold.seconds, t_inf + warm.seconds, t_inf,
                     cold_case, warm_case)
    cmp.warm_raw_steps = warm.steps[:n]
    cmp.twin_steps = extra
    return cmp
