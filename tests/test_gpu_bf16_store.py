"""bfloat16-shifted storage rounds f - w_i to bfloat16 to nearest, ties to even.

Populations written through the C ABI as float are stored as bf16(f - w_i) and read back as
float(bf16) + w_i (the storage format of the reduced-precision lattices: f - w_i keeps the
deviation from the rest weight).  The expected values follow the definition of rounding
rather than a bit trick: of the two bfloat16 neighbours of the float deviation, the nearer
one, and on a tie the one with an even last mantissa bit; past the largest finite bfloat16
by half a unit or more, infinity.  The 18.9 M inputs are random float bit patterns over every
exponent (about 290 of them land on an exact tie), infinities, the largest finite floats
(which round to infinity), zeros and NaNs (which must stay NaN).
"""

import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "boltzmann_amd", "libbolt_hip.so")
GPU, BGK, BF16S, FIELD_F, DT_F32 = 1, 0, 2, 1, 0
W9 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, np.float32)  # D2Q9 rest weights


def nearest_even_bf16(d):
    """bfloat16 bits of float32 array d (no NaNs), by the definition of round-to-nearest-even."""
    u = d.view(np.uint32).astype(np.int64)
    sign = u & 0x80000000
    mag = u & 0x7FFFFFFF
    lo = mag >> 16                      # bfloat16 magnitude below or at |d| (truncation)
    lo_val = (lo << 16).astype(np.uint32).view(np.float32).astype(np.float64)
    hi = lo + 1                         # the next bfloat16 magnitude up
    # the value of that next magnitude, also past the largest finite one (2^128 for 0x7f80)
    hi_val = np.where(hi >= 0x7F80, 2.0 ** 128, (np.minimum(hi, 0x7F7F) << 16).astype(np.uint32)
                      .view(np.float32).astype(np.float64))
    a = np.abs(d.astype(np.float64))
    with np.errstate(invalid="ignore"):  # infinities: replaced below
        up = (hi_val - a < a - lo_val) | ((hi_val - a == a - lo_val) & ((lo & 1) == 1))
    exact = (mag & 0xFFFF) == 0
    r = np.where(exact | ~up, lo, hi)
    r = np.where(mag >= 0x7F800000, mag >> 16, r)  # infinities stay infinite
    return ((r | (sign >> 16)) & 0xFFFF).astype(np.uint16)


def test_definition_restatement_on_known_values():
    # the restatement itself: ties to even both ways, just past a tie, sign, zeros, and the
    # top of the range (the largest finite bfloat16, a tie half a unit above it, which goes to
    # infinity, the float just below that tie, and the largest finite float)
    bits = [0x3F808000, 0x3F818000, 0x3F808008, 0xBF808000, 0x00000000, 0x80000000,
            0x7F7F0000, 0x7F7F8000, 0x7F7F7FFF, 0x7F7FFFFF, 0x7F800000, 0xFF800000]
    got = nearest_even_bf16(np.array(bits, np.uint32).view(np.float32))
    assert [int(x) for x in got] == [0x3F80, 0x3F82, 0x3F81, 0xBF80, 0x0000, 0x8000,
                                     0x7F7F, 0x7F80, 0x7F7F, 0x7F80, 0x7F80, 0xFF80]


@pytest.mark.gpu
def test_bf16_store_rounds_to_nearest_even():
    L = ctypes.CDLL(LIB)
    vp, P = ctypes.c_void_p, ctypes.POINTER
    L.bolt_last_error.restype = ctypes.c_char_p
    L.bolt_ctx_create.argtypes = [ctypes.c_int, ctypes.c_int, P(vp)]
    L.bolt_ctx_destroy.argtypes = [vp]
    L.bolt_lattice_create.argtypes = [vp, ctypes.c_int, P(ctypes.c_int64), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_int, P(vp)]
    L.bolt_lattice_destroy.argtypes = [vp]
    L.bolt_lattice_write.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t]
    L.bolt_lattice_read.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t]

    def ok(rc, what):
        assert rc == 0, f"{what}: {L.bolt_last_error().decode()}"

    nx, ny = 4096, 512
    rng = np.random.default_rng(61)
    f = rng.integers(0, 2 ** 32, size=(nx, ny, 9), dtype=np.uint64).astype(np.uint32).view(np.float32)
    special = np.array([np.inf, -np.inf, 3.4028235e38, -3.4028235e38, 0.0, -0.0, np.nan], np.float32)
    f.reshape(-1)[:special.size] = special
    ctx, lat = vp(), vp()
    ok(L.bolt_ctx_create(GPU, 0, ctypes.byref(ctx)), "ctx")
    try:
        counts = (ctypes.c_int64 * 3)(nx, ny, 1)
        ok(L.bolt_lattice_create(ctx, 2, counts, 9, BGK, 1.0, 1.0, BF16S, ctypes.byref(lat)), "create")
        ok(L.bolt_lattice_write(lat, FIELD_F, 0, DT_F32, f.ctypes.data, f.nbytes), "write f")
        back = np.empty_like(f)
        ok(L.bolt_lattice_read(lat, FIELD_F, 0, DT_F32, back.ctypes.data, back.nbytes), "read f")
    finally:
        if lat:
            L.bolt_lattice_destroy(lat)
        L.bolt_ctx_destroy(ctx)
    with np.errstate(invalid="ignore", over="ignore"):
        d = f - W9                      # the float deviation the device stores
    nan = np.isnan(d)
    assert np.isnan(back[nan]).all()
    dn = d[~nan]
    bits = nearest_even_bf16(dn)
    assert ((dn.view(np.uint32) & 0xFFFF) == 0x8000).sum() > 100  # ties were exercised
    with np.errstate(invalid="ignore", over="ignore"):
        want = (bits.astype(np.uint32) << 16).view(np.float32) + np.broadcast_to(W9, d.shape)[~nan]
    np.testing.assert_array_equal(back[~nan].view(np.uint32), want.view(np.uint32))
