"""Throughput of the lattice update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: Taylor-Green vortex, D3Q27, 256^3 cells, periodic, u0 = 0.04, TRT with
tau+ = 0.6 and magic number 1/4, float populations (the reference's format).  A step is
one in-place AA update of every cell (one kernel launch).  The K timed steps run
back-to-back on the lattice's HIP stream with inputs already resident in HBM and are
timed with HIP events on that stream; MLUPS = cells * K / (1e6 * seconds).

The same line reports the other storage formats (FP16- and BF16-shifted, with their
error against a double-storage run of the same kernels), D3Q27 BGK, and a CPU baseline:
the oracle (OpenMP C++ restatement of the reference's kernels) on a bounded sample.

roofline.achieved = 217 B per lattice update (27 float populations read and written
once, 1 flag byte) * cells / average launch duration.  memory_ceilings_gbs reports what
plain copy / read / write / in-place plane kernels moving the same bytes reach on the same
GPU (boltzmann_amd/bw_probe).  The AA update alternates two access mixes: one step reads
and writes each cell's own slots (the probe's inplace27), the next reads and writes the
neighbours' slots (shift27).  frac_of_aa_pair_ceiling divides the achieved rate by their
harmonic mean, the rate of a step pair that ran each mix at its probe ceiling.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
B_ALG = {"f32": 2 * 27 * 4 + 1, "f32f": 2 * 27 * 4 + 1, "f16s": 2 * 27 * 2 + 1,
         "bf16s": 2 * 27 * 2 + 1, "f64": 2 * 27 * 8 + 1}


COUNTERS = os.path.join("profiles", "hot_path_kernels_hbm_counters.md")


def kernel_sources_sha() -> str:
    """sha256 prefix of the headline kernel's code (update.hpp then lattice.hpp) with the
    comments and whitespace taken out, so that a comment edit does not retire counters
    taken on the same code."""
    import hashlib
    import re
    h = hashlib.sha256()
    for name in ("update.hpp", "lattice.hpp"):
        with open(os.path.join(ROOT, "boltzmann_amd", "csrc", name)) as fh:
            text = fh.read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)   # block comments
        text = re.sub(r"//[^\n]*", " ", text)                 # line comments
        h.update(" ".join(text.split()).encode())
    return h.hexdigest()[:16]


def measured_traffic():
    """HBM bytes per launch of the FP32 D3Q27 update from the committed counter summary
    (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE, calibrated on 1 GiB copies),
    with the file it came from; the bytes are None when the kernel's sources changed since
    those counters were taken."""
    vals = {}
    try:
        for line in open(os.path.join(ROOT, COUNTERS)):
            if ":" in line and not line.startswith((" ", "|", "#")):
                k, _, v = line.partition(":")
                vals[k.strip()] = v.strip()
    except OSError:
        return None, {"source": COUNTERS, "status": "missing"}
    info = {"source": COUNTERS, "kernel_sources_sha256_16": vals.get("kernel_sources_sha256_16")}
    if vals.get("kernel_sources_sha256_16") != kernel_sources_sha():
        info["status"] = "stale: the kernel sources changed since these counters were taken"
        return None, info
    info["status"] = "current"
    return int(vals["traffic_bytes_per_launch_d3q27_f32"]), info


def tgv(n: int, u0: float = 0.04):
    k = 2 * np.pi / n
    i = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    sx, cx = np.sin(k * X), np.cos(k * X)
    sy, cy = np.sin(k * Y), np.cos(k * Y)
    cz = np.cos(k * Z)
    vel = np.empty((n, n, n, 3), np.float32)
    vel[..., 0] = u0 * sx * cy * cz
    vel[..., 1] = -u0 * cx * sy * cz
    vel[..., 2] = 0.0
    rho = (1 + 3 * (u0 ** 2 / 16) * (np.cos(2 * k * X) + np.cos(2 * k * Y)) *
           (np.cos(2 * k * Z) + 2)).astype(np.float32)
    return rho, vel


def omega(coll: str, tau: float = 0.6):
    from boltzmann_amd import core
    return core.bgk(tau) if coll == "bgk" else core.trt(tau)


def run_gpu(n, coll, storage, steps, warmup, rho, vel, keep_vel=False, vw=0):
    import boltzmann_amd
    sim = boltzmann_amd.Simulation("gpu", [n, n, n], omega(coll), storage=storage)
    sim.set_vector_width(vw)
    sim.fluid.rho[...] = rho
    sim.fluid.vel[...] = vel
    sim.step(2)  # first-iterate initialisation + upload, then 2 steps
    if warmup > 2:
        sim.step(warmup - warmup % 2)
    m0 = sim.reduce("mass")
    t0 = time.perf_counter()
    ms = sim.step_timed(steps)
    wall = time.perf_counter() - t0
    m1 = sim.reduce("mass")
    cells = n ** 3
    out = {
        "mlups": cells * steps / (1e6 * ms * 1e-3),
        "ms_per_step": ms / steps,
        "wall_mlups": cells * steps / (1e6 * wall),
        "mass_drift": abs(m1 - m0) / m0,
    }
    if keep_vel:
        out["vel"] = sim.read_field("vel", dtype="f64")
    del sim
    return out


def run_sim_style(n, rho, vel, calls: int = 5, per_call: int = 100):
    """The reference's run_sim rate (python/boltzmann/simulation.py:235-246): wall time of
    iterate(per_call), which reads rho and u back to the host, plus the non-finite check,
    which boltzmann_amd.simulation.run_sim makes on the device."""
    import boltzmann_amd
    sim = boltzmann_amd.Simulation("gpu", [n, n, n], omega("trt"), storage="f32")
    sim.fluid.rho[...] = rho
    sim.fluid.vel[...] = vel
    sim.iterate(2)
    t0 = time.perf_counter()
    for _ in range(calls):
        sim.iterate(per_call)
        if sim.count_nonfinite("vel"):
            raise ValueError("Non-finite value detected!")
    secs = time.perf_counter() - t0
    del sim
    return {"mlups": round(n ** 3 * per_call * calls / (1e6 * secs), 1), "steps_per_call": per_call,
            "calls": calls}


def cpu_baseline(budget_s: float = 12.0):
    """Oracle D3Q27 TRT on a 64^3 TGV sample, all host threads, about budget_s seconds."""
    import oracle
    oracle.build()
    n = 64
    rho, vel = tgv(n)
    tau = 0.6
    om = ("trt", float(np.float32(1 / tau)), float(np.float32(1 / (0.25 / (tau - 0.5) + 0.5))))
    sim = oracle.OracleSimulation((n, n, n), om)
    sim.rho[...] = rho
    sim.vel[...] = vel
    sim.iterate(2)  # first call: equilibrium initialisation
    t0 = time.perf_counter()
    sim.iterate(4)
    per = (time.perf_counter() - t0) / 4
    steps = max(2, int(budget_s / max(per, 1e-6)) // 2 * 2)
    t0 = time.perf_counter()
    sim.iterate(steps)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / (1e6 * dt), "unit": "MLUPS", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"oracle (OpenMP C++ restatement of lib.cl) D3Q27 TRT TGV {n}^3, "
                      f"{steps} steps, {dt:.1f} s",
            "cpu": _cpu_model(), "affinity": len(os.sched_getaffinity(0))}


def _cpu_model() -> str:
    """The host CPU's model name (first entry of /proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def rb_batch(steps: int = 50000):
    """64 Rayleigh-Benard cells of uneven sizes (D2Q9 + D2Q5 + Boussinesq), one launch per
    step over all of them, graph-replayed, shape picked by the autotuner."""
    from boltzmann_amd import rb
    specs = rb.sweep(64, seed=4)
    b = rb.make_batch(specs)
    report = b.tune(40)
    b.step(20)
    ms = b.step_timed(steps)
    cells = b.cells
    mlups = cells * steps / (1e6 * ms * 1e-3)
    nus = []
    for k in (0, 1, 2, 3):
        T = b.read(k, "C")
        uy = b.read(k, "vel")[..., 1]
        nus.append(round(rb.nusselt(T, uy, specs[k]), 3))
    # physics sanity: aspect-2 cells run to developed convection (subcritical: Nu = 1)
    check = [rb.spec(256, 130, ra, seed=2000 + k) for k, ra in enumerate((1000.0, 1e4, 1e5))]
    cb = rb.make_batch(check)
    cb.step(400_000)
    nu_check = {f"{s.ra:.0e}": round(rb.nusselt(cb.read(k, "C"), cb.read(k, "vel")[..., 1], s), 4)
                for k, s in enumerate(check)}
    return {"instances": len(specs), "cells": int(cells), "steps": steps,
            "mlups": round(mlups, 1), "ms_per_step": round(ms / steps, 5),
            "hbm_frac": round(mlups * 1e6 * 121 / (PEAK_HBM_GBS * 1e9), 4),
            "shape": report.strip().splitlines()[-1],
            "nusselt_first4": nus, "ra_first4": [round(s.ra) for s in specs[:4]],
            "nusselt_256x130_400k_steps": nu_check}


def reference_published_d2q9(calls: int = 3, per_call: int = 1000):
    """The workload of the reference's only published rate (docs/source/guides/script.rst:
    148-155): D2Q9 BGK on 2.0 M cells (104 MB at 52 B per cell), timed as run_sim does --
    iterate(1000) including the host read-back of rho and u, MLUPS per batch of 1000
    iterations (simulation.py:235-246).  The domain shape is not stated there; 2000 x 1000
    periodic cells with a small shear flow stand in for it.  Also reports the kernel rate."""
    import boltzmann_amd
    from boltzmann_amd import core
    nx, ny = 2000, 1000
    sim = boltzmann_amd.Simulation("gpu", [nx, ny], core.bgk(0.6))
    y = np.arange(ny)
    sim.fluid.vel[..., 0] = (0.05 * np.sin(2 * np.pi * y / ny)).astype(np.float32)[None, :]
    sim.iterate(2)
    rates = []
    for _ in range(calls):
        t0 = time.perf_counter()
        sim.iterate(per_call)
        rates.append(nx * ny * per_call / (1e6 * (time.perf_counter() - t0)))
    ms = sim.step_timed(per_call)
    out = {"cells": nx * ny, "run_sim_mlups_per_batch": [round(r, 1) for r in rates],
           "kernel_mlups": round(nx * ny * per_call / (1e6 * ms * 1e-3), 1),
           "reference_published_mlups": [1311, 1633, 1599, 1628, 1627]}
    del sim
    return out


def cylinder_flow(nx=4096, ny=1024, r=64, u0=0.05):
    """Flags and initial velocity of the channel: fixed-velocity inflow column, fixed-density
    outflow column, walls top and bottom and the disc (x-1024)^2 + (y-512)^2 < r^2 (scaled
    with the channel), written last (vortex_street.py:100-115)."""
    from boltzmann_amd.core import CellFlags
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    flags = np.zeros((nx, ny), np.int32)
    flags[0, :] = CellFlags.FIXED_FLUID
    flags[-1, :] = CellFlags.FIXED_FLUID_DENSITY
    flags[:, 0] = CellFlags.WALL
    flags[:, -1] = CellFlags.WALL
    flags[(X - nx // 4) ** 2 + (Y - ny // 2) ** 2 < r * r] = CellFlags.WALL
    vel = np.zeros((nx, ny, 2), np.float32)
    vel[..., 0] = u0
    vel[flags == CellFlags.WALL] = 0.0
    return flags, vel


def cylinder(steps_marks=(1000, 5000, 20000), tau_minus=0.52):
    """Flow past a disc in a 4096 x 1024 channel, D2Q9 TRT, Re = 1000 on the diameter:
    walls top and bottom, bounce-back disc r = 64 at (1024, 512), fixed-velocity inflow
    u = 0.05, fixed-density outflow.  Runs f64, f32, f16s and bf16s storage and reports the
    relative max-norm of u against the f64 run at each mark, plus MLUPS.

    tau- is 0.52, as in the reference's vortex-street example (vortex_street.py:102).  The
    magic number 1/4 (tau- = 13.52 at tau+ = 0.5192) makes this flow diverge: the record's
    magic_quarter_f64 gives the first check (every 250 steps) at which the f64 run holds a
    non-finite velocity."""
    import boltzmann_amd
    from boltzmann_amd import Omega
    nx, ny, r, u0 = 4096, 1024, 64, 0.05
    nu = u0 * 2 * r / 1000.0
    tau_p = 3 * nu + 0.5
    flags, vel0 = cylinder_flow(nx, ny, r, u0)
    out = {"tau_plus": round(tau_p, 6), "tau_minus": tau_minus}
    ref = {}
    for st in ("f64", "f32", "f16s", "bf16s"):
        sim = boltzmann_amd.Simulation("gpu", [nx, ny], Omega.TRT(1 / tau_p, 1 / tau_minus),
                                       storage=st)
        sim.cells.flags[...] = flags
        sim.fluid.vel[...] = vel0
        done, ms_tot, errs = 0, 0.0, {}
        sim.step(2)
        done = 2
        for mark in steps_marks:
            ms_tot += sim.step_timed(mark - done)
            done = mark
            u = sim.read_field("vel", dtype="f64")
            if st == "f64":
                ref[mark] = u
            else:
                errs[str(mark)] = float(np.abs(u - ref[mark]).max() / np.abs(ref[mark]).max())
        mlups = nx * ny * (done - 2) / (1e6 * ms_tot * 1e-3)
        out[st] = {"mlups": round(mlups, 1), **({"u_err_vs_f64": errs} if errs else {})}
        del sim
    # the configuration's magic-number TRT: find where the f64 run stops being finite
    tau_m = 0.25 / (tau_p - 0.5) + 0.5
    sim = boltzmann_amd.Simulation("gpu", [nx, ny], Omega.TRT(1 / tau_p, 1 / tau_m), storage="f64")
    sim.cells.flags[...] = flags
    sim.fluid.vel[...] = vel0
    sim.iterate(2)
    bad_at = None
    for step in range(250, 20001, 250):
        sim.step(248 if step == 250 else 250)
        if sim.count_nonfinite("vel"):
            bad_at = step
            break
    out["magic_quarter_f64"] = {"tau_minus": round(tau_m, 4), "nonfinite_at_step": bad_at,
                                "checked_to": step}
    del sim
    return out


def warmstart_channel(n_train: int = 2000, n_test: int = 100, epochs: int = 20,
                      width: int = 16, depth: int = 6, max_steps: int = 60000,
                      train_budget: float = 420.0, prior_factor: int = 4):
    """Learned warm start against the plain solver's time to steady state.

    D2Q9 TRT channels of 512 x 256 cells with 1-8 seeded obstacles (radius / half-size
    8-32), u_in in [0.02, 0.05], Re in [10, 100].  n_train cases (seeds 0..) are solved to
    the FP64 residual 1e-6 in batches; those that converge within max_steps train the
    network (with their mirror images across the channel) for up to `epochs` epochs within
    train_budget seconds, given the 4x coarser lattice's steady state as input with
    prior_factor 4 (0: geometry features only); n_test held-out cases (seeds 10^6..) run
    cold and warm to the same residual, the warm run deciding its fallback against a cold
    twin in its own batch."""
    from boltzmann_amd import warmstart as ws

    def log(msg):
        print(f"[bench] warm start: {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    cases, fields, steps = ws.generate(n_train, seed0=0, batch=125, max_steps=max_steps,
                                       check_every=2000, log=log)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    net, losses = ws.train(cases, fields, epochs=epochs, batch=8, width=width, depth=depth,
                           lr=2e-3, seed=0, log=log, time_budget=train_budget,
                           prior_factor=prior_factor)
    t_train = time.perf_counter() - t0
    del fields
    held = [ws.make_case(10 ** 6 + s) for s in range(n_test)]
    cmp = ws.compare(net, held, check_every=500, max_steps=2 * max_steps, judge_at=500)
    s = cmp.summary()
    log(f"held-out: {s}")
    return {"train_cases": len(cases), "train_requested": n_train,
            "train_steps_median": int(np.median(steps)) if len(steps) else None,
            "generate_s": round(t_gen, 1), "train_s": round(t_train, 1),
            "epochs": len(losses), "width": width, "depth": depth, "prior_factor": prior_factor,
            "params": int(net.net.param_count),
            "loss_first_last": [round(float(losses[0]), 5), round(float(losses[-1]), 5)],
            "held_out": s}


def warmstart_bounded(n_train: int = 1000, n_test: int = 48, epochs: int = 30,
                      train_budget: float = 45.0):
    """The learned warm start at the metric's size, bounded to about three minutes: 512 x 256
    channels (1-8 obstacles of size 8-32, Re 10-100, u_in 0.02-0.05), n_train seeded cases
    solved to the 1e-6 residual (those that converge train), a depth-6 width-16 network
    (0.64 M parameters) without and
    with the coarse-lattice prior (the steady state on a 4x coarser lattice, upsampled, as
    input; the network learns the correction), each compared with the cold start on n_test
    held-out cases.  Each warm run decides its own fallback against a cold twin in the same
    batch (both twins' steps charged); time_to_residual_ratio counts features, the coarse
    solve and inference."""
    from boltzmann_amd import warmstart as ws

    def log(msg):
        print(f"[bench] warm start: {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    cases, fields, _ = ws.generate(n_train, seed0=0, batch=250, max_steps=60000, check_every=2000,
                                   log=log)
    out = {"lattice": [512, 256], "train_cases": len(cases), "train_requested": n_train,
           "generate_s": round(time.perf_counter() - t0, 1)}
    held = [ws.make_case(10 ** 6 + s) for s in range(n_test)]
    for name, pf in (("coarse_prior", 4), ("plain", 0)):
        t0 = time.perf_counter()
        model, losses = ws.train(cases, fields, epochs=epochs, batch=8, width=16, depth=6, lr=2e-3,
                                 seed=0, prior_factor=pf, time_budget=train_budget, log=log,
                                 log_every=10)
        t_train = time.perf_counter() - t0
        s = ws.compare(model, held, check_every=500, max_steps=120000, judge_at=500).summary()
        s["train_s"] = round(t_train, 1)
        s["epochs"] = len(losses)
        s["params"] = int(model.net.param_count)
        out[name] = s
        log(f"{name}: {s}")
        del model
    return out


def memory_ceilings():
    """GB/s of copy / read-only / write-only / in-place plane kernels moving the bytes of one
    256^3 D3Q27 step (boltzmann_amd/bw_probe, a child process run before this process uses
    the GPU): what the update's access pattern can reach on this GPU."""
    import subprocess
    exe = os.path.join(ROOT, "boltzmann_amd", "bw_probe")
    if not os.path.exists(exe):
        return {"error": "boltzmann_amd/bw_probe is not built (python boltzmann_amd/build.py)"}
    try:
        res = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    except subprocess.TimeoutExpired:
        return {"error": "bw_probe timed out"}
    if res.returncode != 0:
        return {"error": f"bw_probe exit {res.returncode}: {res.stderr.strip()[-200:]}"}
    return json.loads(res.stdout.strip().splitlines()[-1])


def nn_throughput(nx: int = 512, ny: int = 256, batch: int = 8, width: int = 16, depth: int = 6,
                  samples: int = 64, epochs: int = 2):
    """Training and inference throughput of the warm-start network on random data of the
    512 x 256 channel size (the GEMM kernels' workload), after one warm-up epoch."""
    import boltzmann_amd
    net = boltzmann_amd.Network("gpu", nx, ny, batch, width=width, seed=0, depth=depth)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((samples, nx, ny, 8)).astype(np.float32)
    y = rng.standard_normal((samples, nx, ny, 3)).astype(np.float32)
    net.set_data(x, y)
    net.train(1, lr=1e-3)
    t0 = time.perf_counter()
    net.train(epochs, lr=1e-3)
    train = epochs * samples / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    net.predict(x[:32])
    infer = 32 / (time.perf_counter() - t0)
    return {"lattice": [nx, ny], "depth": depth, "width": width, "batch": batch,
            "params": int(net.param_count), "train_samples_per_s": round(train, 1),
            "predict_samples_per_s": round(infer, 1)}


def profile_kernels(steps: int = 20) -> dict:
    """A short run of every hot-path kernel for a profiler: the 256^3 D3Q27 TRT lattice in
    every storage format, the 4096 x 1024 D2Q9 TRT cylinder channel in f32 and f16s (with the
    ordinary loads its working set gets by default, and f32 once more with streaming loads),
    a 1024 x 1024 D2Q9 lattice with a fused D2Q5 tracer and Boussinesq coupling, and the
    64-cell Rayleigh-Benard batch with direct launches (no graph replay), its even steps
    row-staged and then direct (the autotuner's two choices).  Returns the launches and the
    algorithmic bytes per launch of each workload."""
    import boltzmann_amd
    from boltzmann_amd import Omega, core, rb
    out = {}
    n = 256
    rho, vel = tgv(n)
    for st in ("f32", "f16s", "bf16s", "f64"):
        run_gpu(n, "trt", st, steps, 0, rho, vel)
        out[f"tgv256_d3q27_trt_{st}"] = {"launches": steps + 2, "bytes_per_launch": n ** 3 * B_ALG[st]}
    del rho, vel
    flags, vel0 = cylinder_flow()
    nu = 0.05 * 128 / 1000.0
    for st, b, policy in (("f32", 73, "auto"), ("f16s", 37, "auto"), ("f32", 73, "stream")):
        sim = boltzmann_amd.Simulation("gpu", [4096, 1024], Omega.TRT(1 / (3 * nu + 0.5), 1 / 0.52),
                                       storage=st)
        sim.set_cache_policy(policy)
        sim.cells.flags[...] = flags
        sim.fluid.vel[...] = vel0
        sim.step(2)
        sim.step_timed(10 * steps)
        key = f"cylinder_4096x1024_d2q9_trt_{st}" + ("_streaming" if policy == "stream" else "")
        out[key] = {"launches": 10 * steps + 2, "bytes_per_launch": 4096 * 1024 * b}
        del sim
    m = 1024
    sim = boltzmann_amd.Simulation("gpu", [m, m], Omega.TRT(1 / 0.6, 1 / 0.9))
    rng = np.random.default_rng(0)
    t = sim.add_tracer("T", core.bgk(0.6))
    t.val[...] = rng.random((m, m)).astype(np.float32)
    sim.add_boussinesq_coupling(t, -1e-3, 0.5, np.array([0.0, -1.0], np.float32))
    sim.step(2)
    sim.step_timed(10 * steps)
    out["fused_1024x1024_d2q9_trt_d2q5"] = {"launches": 10 * steps + 2, "bytes_per_launch": m * m * 121}
    del sim
    for staged in (True, False):
        b = rb.make_batch(rb.sweep(64, seed=4))
        b.set_shape(128, 0)
        b.set_row_staging(staged)
        b.step(10 * steps)
        b.step_timed(2)
        out[f"rb_batch_64_rows_{'staged' if staged else 'direct'}"] = {
            "launches": 10 * steps + 2, "bytes_per_launch": int(b.cells) * 121}
        del b
    return out


def copy_calibration(nbytes: int = 1 << 30) -> None:
    """One device-to-device copy of nbytes (hipMemcpy): a kernel of known traffic for
    scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy calibration: 3 x {nbytes} bytes device-to-device", file=sys.stderr,
          flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the memory-ceiling probe (boltzmann_amd/bw_probe)")
    ap.add_argument("--rb-steps", type=int, default=50000)
    ap.add_argument("--warmstart", action="store_true",
                    help="add the learned warm start vs cold start on 512x256 channels "
                         "(about 15 minutes)")
    ap.add_argument("--ws-train", type=int, default=2000)
    ap.add_argument("--ws-test", type=int, default=100)
    ap.add_argument("--ws-epochs", type=int, default=20)
    ap.add_argument("--nn-throughput", action="store_true",
                    help="add the warm-start network's training / inference throughput")
    ap.add_argument("--nn-width", type=int, default=16,
                    help="channels of the first level of the --nn-throughput network (16: the warm "
                         "start's default)")
    ap.add_argument("--no-warmstart-bounded", action="store_true",
                    help="skip the bounded 512x256 warm-start leg of the default run (~3 min)")
    ap.add_argument("--vw-sweep", action="store_true",
                    help="add MLUPS of every storage format at 1, 2 and 4 cells per lane")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only run every hot-path kernel briefly (for rocprofv3) and print "
                         "the algorithmic bytes per launch of each workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the lattice update here is built and measured on a "
              "single MI355X (multi-GPU slab decomposition is a design note only).",
              file=sys.stderr)
        return 2
    steps = args.steps + args.steps % 2
    n = args.n
    if args.profile_kernels:
        if args.copy_calibration:
            copy_calibration()
        print(json.dumps({"profile_kernels": profile_kernels(min(steps, 20))}), flush=True)
        return 0
    ceilings = None if args.no_probe else memory_ceilings()
    if ceilings is not None:
        print(f"[bench] memory ceilings: {ceilings}", file=sys.stderr, flush=True)
    if args.copy_calibration:
        copy_calibration()
    rho, vel = tgv(n)
    print(f"[bench] TGV D3Q27 {n}^3: building inputs done", file=sys.stderr, flush=True)

    head = run_gpu(n, "trt", "f32", steps, args.warmup, rho, vel)
    print(f"[bench] f32 trt: {head['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
    run_sim_rate = run_sim_style(n, rho, vel)
    print(f"[bench] run_sim-style (iterate + read-back): {run_sim_rate['mlups']:.0f} MLUPS",
          file=sys.stderr, flush=True)
    variants = {}
    if not args.no_variants:
        ref64 = run_gpu(n, "trt", "f64", max(2, min(steps, 200)), 0, rho, vel, keep_vel=True)
        print(f"[bench] f64 trt: {ref64['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
        umax = np.abs(ref64["vel"]).max()
        for coll, st in (("trt", "f16s"), ("trt", "bf16s"), ("bgk", "f32"), ("trt", "f32f")):
            r = run_gpu(n, coll, st, steps, args.warmup, rho, vel)
            v = {"mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 4),
                 "hbm_frac": round(r["mlups"] * 1e6 * B_ALG[st] / (PEAK_HBM_GBS * 1e9), 4)}
            if st in ("f16s", "bf16s") and ceilings and "shift27_h" in ceilings:
                gbs = r["mlups"] * 1e6 * B_ALG[st] / 1e9
                v["frac_of_aa_pair_ceiling"] = round(
                    gbs * 0.5 * (1 / ceilings["inplace27_h"] + 1 / ceilings["shift27_h"]), 4)
            if coll == "trt":
                e = run_gpu(n, coll, st, max(2, min(steps, 200)), 0, rho, vel, keep_vel=True)
                v["u_err_vs_f64_200steps"] = float(np.abs(e["vel"] - ref64["vel"]).max() / umax)
            variants[f"{coll}_{st}"] = v
            print(f"[bench] {coll} {st}: {r['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
        variants["trt_f64"] = {"mlups": round(ref64["mlups"], 1)}
        variants["d2q9_bgk_2M_reference_published_workload"] = reference_published_d2q9()
        print(f"[bench] reference's published workload: "
              f"{variants['d2q9_bgk_2M_reference_published_workload']}", file=sys.stderr, flush=True)
        variants["cylinder_d2q9_trt_4096x1024"] = cylinder()
        print("[bench] cylinder done", file=sys.stderr, flush=True)
        variants["rb_batch"] = rb_batch(args.rb_steps)
        print(f"[bench] rb batch: {variants['rb_batch']['mlups']:.0f} MLUPS", file=sys.stderr,
              flush=True)
        variants["f16s_over_f32"] = round(variants["trt_f16s"]["mlups"] / head["mlups"], 3)
        if not args.no_warmstart_bounded:
            variants["warmstart_channel_512x256"] = warmstart_bounded()
    if args.vw_sweep:
        sweep = {}
        for st in ("f32", "f16s", "bf16s", "f32f", "f64"):
            for vw in ((1, 2) if st == "f64" else (1, 2, 4)):
                r = run_gpu(n, "trt", st, steps, args.warmup, rho, vel, vw=vw)
                sweep[f"{st}_vw{vw}"] = round(r["mlups"], 1)
                print(f"[bench] trt {st} {vw} cells per lane: {r['mlups']:.0f} MLUPS",
                      file=sys.stderr, flush=True)
        variants["vector_width_sweep_mlups"] = sweep
    if args.nn_throughput:
        variants["nn_unet_512x256"] = nn_throughput(width=args.nn_width)
        print(f"[bench] network: {variants['nn_unet_512x256']}", file=sys.stderr, flush=True)
    if args.warmstart:
        variants["warmstart_channel_512x256_full"] = warmstart_channel(
            args.ws_train, args.ws_test, args.ws_epochs)

    achieved = head["mlups"] * 1e6 * B_ALG["f32"] / 1e9  # GB/s
    traffic, traffic_info = measured_traffic()
    line = {
        "metric": "MLUPS & % HBM roofline (D3Q27 FP32/FP16); NN warm-start vs plain time-to-steady",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex D3Q27 TRT {n}^3 periodic, u0=0.04, "
                               f"tau+=0.6, magic 1/4, FP32 populations, AA in-place",
                   "cells": n ** 3, "lattice": "D3Q27", "collision": "TRT",
                   "storage": "f32"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic if n == 256 else None,
                     "traffic_source": traffic_info,
                     "bytes_per_lup": B_ALG["f32"]},
        "wall_mlups": round(head["wall_mlups"], 1),
        "run_sim_mlups": run_sim_rate,
        "memory_ceilings_gbs": ceilings,
        "frac_of_inplace_ceiling": (round(achieved / ceilings["inplace27"], 4)
                                    if ceilings and "inplace27" in ceilings else None),
        "frac_of_aa_pair_ceiling": (
            round(achieved * 0.5 * (1 / ceilings["inplace27"] + 1 / ceilings["shift27"]), 4)
            if ceilings and "shift27" in ceilings else None),
        "mass_drift": head["mass_drift"],
        "variants": variants,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
