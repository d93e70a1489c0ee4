"""bolt_lattice_read_noalloc against bolt_lattice_read, through the C ABI (ctypes).

Two lattices are read back in a row, the second smaller than the first: a D3Q27 lattice
whose double-precision population array (16.6 MB) passes through the 16 MiB staging area
in two pieces, then a D2Q9 lattice.  Every field (flags, f, rho, vel, tracer g and C) in
every host type must come back byte-identical from both entry points; wrong sizes and
null arguments are refused the same way.
"""

import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "boltzmann_amd", "libbolt_hip.so")

GPU, TRT = 1, 1
F_FLAGS, F_F, F_RHO, F_VEL, F_G, F_C = 0, 1, 2, 3, 4, 5
DT_F32, DT_F64, DT_I32 = 0, 1, 2
P, VP = ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(LIB)
    L.bolt_last_error.restype = ctypes.c_char_p
    L.bolt_ctx_create.argtypes = [ctypes.c_int, ctypes.c_int, VP]
    L.bolt_ctx_destroy.argtypes = [P]
    L.bolt_lattice_create.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                      ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, VP]
    L.bolt_lattice_destroy.argtypes = [P]
    L.bolt_lattice_add_tracer.argtypes = [P, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_int)]
    rw = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t]
    L.bolt_lattice_write.argtypes = rw
    L.bolt_lattice_read.argtypes = rw
    L.bolt_lattice_read_noalloc.argtypes = rw
    L.bolt_lattice_equilibrate.argtypes = [P]
    L.bolt_lattice_step.argtypes = [P, ctypes.c_uint64]
    return L


def ok(L, rc):
    assert rc == 0, L.bolt_last_error().decode()


def _ptr(a):
    return a.ctypes.data_as(P)


def build(L, ctx, counts, q, tq, seed):
    rng = np.random.default_rng(seed)
    d = len(counts)
    c = (ctypes.c_int64 * d)(*counts)
    lat = P()
    ok(L, L.bolt_lattice_create(ctx, d, c, q, TRT, 1 / 0.6, 1 / 0.9, 0, ctypes.byref(lat)))
    tid = ctypes.c_int()
    ok(L, L.bolt_lattice_add_tracer(lat, tq, 1 / 0.7, ctypes.byref(tid)))
    flags = np.where(rng.random(counts) < 0.1, 1, 0).astype(np.int32)
    rho = (1 + 0.01 * rng.standard_normal(counts)).astype(np.float32)
    vel = (0.03 * rng.standard_normal(counts + (d,))).astype(np.float32)
    C = rng.random(counts).astype(np.float32)
    for fid, arr, dt in ((F_FLAGS, flags, DT_I32), (F_RHO, rho, DT_F32), (F_VEL, vel, DT_F32),
                         (F_C, C, DT_F32)):
        ok(L, L.bolt_lattice_write(lat, fid, tid.value, dt, _ptr(arr), arr.nbytes))
    ok(L, L.bolt_lattice_equilibrate(lat))
    ok(L, L.bolt_lattice_step(lat, 6))
    return lat, tid.value


def compare_all(L, lat, tid, counts, q, tq):
    n = int(np.prod(counts))
    d = len(counts)
    cases = [(F_FLAGS, n, DT_I32, np.int32), (F_F, n * q, DT_F32, np.float32),
             (F_F, n * q, DT_F64, np.float64), (F_RHO, n, DT_F64, np.float64),
             (F_VEL, n * d, DT_F32, np.float32), (F_VEL, n * d, DT_F64, np.float64),
             (F_G, n * tq, DT_F32, np.float32), (F_C, n, DT_F32, np.float32)]
    for fid, count, dt, npt in cases:
        a = np.full(count, 7, npt)
        b = np.full(count, -7, npt)
        ok(L, L.bolt_lattice_read(lat, fid, tid, dt, _ptr(a), a.nbytes))
        ok(L, L.bolt_lattice_read_noalloc(lat, fid, tid, dt, _ptr(b), b.nbytes))
        assert a.tobytes() == b.tobytes(), (fid, dt)
    # the same refusals as bolt_lattice_read
    x = np.zeros(n, np.float32)
    assert L.bolt_lattice_read_noalloc(lat, F_RHO, tid, DT_F32, _ptr(x), x.nbytes - 4) == 1
    assert L.bolt_lattice_read_noalloc(lat, F_FLAGS, tid, DT_F32, _ptr(x), x.nbytes) == 1
    assert L.bolt_lattice_read_noalloc(lat, F_RHO, tid, DT_F32, None, x.nbytes) == 1
    assert L.bolt_lattice_read_noalloc(None, F_RHO, tid, DT_F32, _ptr(x), x.nbytes) == 1


def test_read_noalloc_matches_read(lib):
    L = lib
    ctx = P()
    ok(L, L.bolt_ctx_create(GPU, 0, ctypes.byref(ctx)))
    try:
        big = (40, 30, 64)       # f as double: 76,800 cells * 27 * 8 B = 16.6 MB, two pieces
        lat, tid = build(L, ctx, big, 27, 7, seed=1)
        compare_all(L, lat, tid, big, 27, 7)
        ok(L, L.bolt_lattice_destroy(lat))
        small = (37, 29)         # then a smaller lattice through the same area
        lat, tid = build(L, ctx, small, 9, 5, seed=2)
        compare_all(L, lat, tid, small, 9, 5)
        ok(L, L.bolt_lattice_destroy(lat))
    finally:
        L.bolt_ctx_destroy(ctx)
