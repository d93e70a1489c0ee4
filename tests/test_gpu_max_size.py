"""Lattices at the size limit of their storage format, driven through the C ABI.

The kernels address every per-cell array as a base pointer plus a 32-bit byte offset, and
bolt_lattice_create admits a lattice while cells * (widest element) < 2^32 bytes: about
1.07e9 cells in float, half-float and bfloat16 storage (whose rho and u are float) and
5.4e8 in double.  These tests run lattices just under that limit, where the last cells'
offsets use the top bit of the 32-bit offset.

Property checked (size-independent): a periodic lattice made of M x K (x L) copies of a
small periodic lattice evolves into the same copies of the small lattice's result, bit for
bit -- every cell runs the same arithmetic on the same inputs, only the addresses differ.
The small lattice's float run is also compared with the oracle bit for bit, and the large
lattice's mass (a double reduction) with the copies' total.  The large lattices never hold
their populations on the host: they start from the device-side equilibrium
(bolt_lattice_equilibrate) and only rho and u are read back.
"""

import ctypes
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "boltzmann_amd", "libbolt_hip.so")

GPU, F32, F64 = 1, 0, 1
FIELD_FLAGS, FIELD_RHO, FIELD_VEL = 0, 2, 3
DT_F32, DT_F64, DT_I32 = 0, 1, 2
STORE = {"f32": 0, "f16s": 1, "bf16s": 2, "f64": 3}


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libbolt_hip.so is not built (python boltzmann_amd/build.py)")
    L = ctypes.CDLL(LIB)
    vp, P = ctypes.c_void_p, ctypes.POINTER
    L.bolt_last_error.restype = ctypes.c_char_p
    L.bolt_ctx_create.argtypes = [ctypes.c_int, ctypes.c_int, P(vp)]
    L.bolt_ctx_destroy.argtypes = [vp]
    L.bolt_lattice_create.argtypes = [vp, ctypes.c_int, P(ctypes.c_int64), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_int, P(vp)]
    L.bolt_lattice_destroy.argtypes = [vp]
    L.bolt_lattice_set_gravity.argtypes = [vp, P(ctypes.c_float)]
    L.bolt_lattice_write.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t]
    L.bolt_lattice_read.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t]
    L.bolt_lattice_equilibrate.argtypes = [vp]
    L.bolt_lattice_step.argtypes = [vp, ctypes.c_uint64]
    L.bolt_lattice_reduce.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
    return L


def _ok(L, rc, what):
    assert rc == 0, f"{what}: {L.bolt_last_error().decode()}"


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_inputs(counts, seed):
    rng = np.random.default_rng(seed)
    d = len(counts)
    flags = np.zeros(counts, np.int32)
    w = rng.random(counts) < 0.1
    flags[w] = 1
    fx = (rng.random(counts) < 0.04) & ~w
    flags[fx] = rng.choice([2, 4, 6], size=int(fx.sum()))
    rho = (1 + 0.02 * rng.standard_normal(counts)).astype(np.float32)
    vel = (0.04 * rng.standard_normal(counts + (d,))).astype(np.float32)
    return flags, rho, vel


def tiled(small, reps):
    """The periodic lattice of reps[a] copies of `small` along each axis a, written into
    one new array without temporaries."""
    d = len(reps)
    cs = small.shape[:d]
    extra = small.shape[d:]
    big = np.empty(tuple(r * c for r, c in zip(reps, cs)) + extra, small.dtype)
    view = big.reshape(tuple(v for r, c in zip(reps, cs) for v in (r, c)) + extra)
    src = small.reshape(tuple(v for c in cs for v in (1, c)) + extra)
    np.copyto(view, src)
    return big


def equal_to_copies(big, small, reps):
    """big is reps copies of small, compared one slab of the first axis at a time."""
    d = len(reps)
    cs = small.shape[:d]
    extra = small.shape[d:]
    rest = tuple(v for r, c in zip(reps[1:], cs[1:]) for v in (r, c))
    src = small.reshape((cs[0],) + tuple(v for c in cs[1:] for v in (1, c)) + extra)
    for m in range(reps[0]):
        slab = big[m * cs[0]:(m + 1) * cs[0]].reshape((cs[0],) + rest + extra)
        if not np.array_equal(slab, np.broadcast_to(src, slab.shape)):
            return False
    return True


def run(L, counts, q, coll, om, storage, flags, rho, vel, gravity, steps):
    """Creates the lattice, uploads flags, rho and u, initialises f on the device, steps,
    and reads rho and u back into the host arrays it was given (as float64 for double
    storage).  Returns (rho, vel, mass)."""
    d = len(counts)
    ctx, lat = ctypes.c_void_p(), ctypes.c_void_p()
    _ok(L, L.bolt_ctx_create(GPU, 0, ctypes.byref(ctx)), "ctx")
    try:
        n = (ctypes.c_int64 * 3)(*(list(counts) + [1] * (3 - d)))
        _ok(L, L.bolt_lattice_create(ctx, d, n, q, coll, om[0], om[1], STORE[storage], ctypes.byref(lat)),
            "create")
        try:
            if gravity is not None:
                g = (ctypes.c_float * 3)(*(list(gravity) + [0.0] * (3 - d)))
                _ok(L, L.bolt_lattice_set_gravity(lat, g), "gravity")
            for field, dt, a in ((FIELD_FLAGS, DT_I32, flags), (FIELD_RHO, DT_F32, rho), (FIELD_VEL, DT_F32, vel)):
                _ok(L, L.bolt_lattice_write(lat, field, 0, dt, a.ctypes.data, a.nbytes), f"write {field}")
            _ok(L, L.bolt_lattice_equilibrate(lat), "equilibrate")
            _ok(L, L.bolt_lattice_step(lat, steps), "step")
            if storage == "f64":
                rho, vel = rho.astype(np.float64), vel.astype(np.float64)
                dt = DT_F64
            else:
                dt = DT_F32
            _ok(L, L.bolt_lattice_read(lat, FIELD_RHO, 0, dt, rho.ctypes.data, rho.nbytes), "read rho")
            _ok(L, L.bolt_lattice_read(lat, FIELD_VEL, 0, dt, vel.ctypes.data, vel.nbytes), "read vel")
            mass = ctypes.c_double()
            _ok(L, L.bolt_lattice_reduce(lat, 0, 0, ctypes.byref(mass)), "reduce")
            return rho, vel, mass.value
        finally:
            L.bolt_lattice_destroy(lat)
    finally:
        L.bolt_ctx_destroy(ctx)


TRT = (float(np.float32(1 / 0.56)), float(np.float32(1 / (0.25 / 0.06 + 0.5))))
BGK = (float(np.float32(1 / 0.62)),) * 2

# (name, small counts, copies, q, collision, omegas, storage, gravity); every large lattice
# is just under its storage format's limit (cells * widest element < 2^32 bytes)
CASES = [
    # 1024 x 1024 x 1023 = 1,072,693,248 cells: 4,290,772,992-byte float planes (27 of them)
    ("d3q27_trt_f32", (16, 16, 31), (64, 64, 33), 27, 1, TRT, "f32", (1e-5, -2e-5, 1e-5)),
    # 16384 x 32767 = 536,854,528 cells: 4,294,836,224-byte double planes
    ("d2q9_bgk_f64", (64, 151), (256, 217), 9, 0, BGK, "f64", (2e-5, -1e-5)),
    # 32766 x 32768 = 1,073,676,288 cells: 2-byte population planes, float rho and u planes
    # of 4,294,705,152 bytes; odd steps run two cells per lane
    ("d2q9_trt_f16s", (86, 128), (381, 256), 9, 1, TRT, "f16s", None),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_largest_lattice_is_copies_of_the_small_one(lib, case):
    name, cs, reps, q, coll, om, storage, gravity = case
    big_counts = tuple(r * c for r, c in zip(reps, cs))
    ncell = int(np.prod(big_counts, dtype=np.int64))
    widest = 8 if storage == "f64" else 4
    assert ncell * widest < 2 ** 32 and ncell * widest > 2 ** 32 - 2 ** 27, "not at the limit"
    flags, rho, vel = small_inputs(cs, seed=len(name))
    steps = 4
    _say(f"[{name}] small {cs}")
    s_rho, s_vel, s_mass = run(lib, cs, q, coll, om, storage, flags, rho.copy(), vel.copy(), gravity, steps)
    if storage == "f32":
        ref = oracle.OracleSimulation(cs, ("trt", *om) if coll else ("bgk", om[0]), q=q, dtype=np.float32)
        ref.flags[...] = flags
        ref.rho[...] = rho
        ref.vel[...] = vel
        if gravity is not None:
            ref.set_gravity(np.asarray(gravity, np.float32))
        ref.iterate(steps)
        nw = (flags & 1) == 0
        np.testing.assert_array_equal(s_rho[nw], ref.rho[nw])
        np.testing.assert_array_equal(s_vel[nw], ref.vel[nw])
    _say(f"[{name}] large {big_counts} ({ncell} cells)")
    b_flags, b_rho, b_vel = tiled(flags, reps), tiled(rho, reps), tiled(vel, reps)
    b_rho, b_vel, b_mass = run(lib, big_counts, q, coll, om, storage, b_flags, b_rho, b_vel, gravity, steps)
    del b_flags
    _say(f"[{name}] comparing")
    assert equal_to_copies(b_rho, s_rho, reps), "rho differs from the copies of the small lattice"
    assert equal_to_copies(b_vel, s_vel, reps), "u differs from the copies of the small lattice"
    copies = int(np.prod(reps))
    assert abs(b_mass - copies * s_mass) <= 1e-12 * abs(copies * s_mass)
