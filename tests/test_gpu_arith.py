"""The transport kernel's per-step arithmetic on the GPU.

Bit-exact parity between the kernel and the CPU oracle rests on both executing
the same FP32 operations.  Where the kernel uses a different instruction
sequence than the oracle's plain C (the draw conversion under a toward-zero
rounding mode, the saturating conversion instruction, Spin's quotient without
the division's range scaling, the exit pixel index as a double product), these
tests compare the two forms on the GPU over whole input ranges; the shared
FP32 functions (logarithm, reciprocal root, renormalisation, sine/cosine, the
(0,1] draw) are compared between the GPU and the host build of the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import cudamcfl_amd as M
import _oracle as O

SC_U2F_RZ, SC_F2U_RZ, SC_DIV, SC_PIXEL = 0, 1, 2, 3
ME_LOG, ME_RSQRT, ME_RENORM, ME_SINCOS, ME_DRAW_OC = 0, 1, 2, 3, 4
NONE = 2**64 - 1


def _lib():
    L = M.lib()
    L.mcfl_selfcheck.argtypes = [C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]
    L.mcfl_math_eval.argtypes = [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint64]
    return L


def selfcheck(which, p, lo, hi):
    """(mismatches, first mismatching input) of check `which` over inputs [lo, hi)."""
    bad, first = C.c_uint64(), C.c_uint64()
    rc = _lib().mcfl_selfcheck(which, p, lo, hi, C.byref(bad), C.byref(first))
    assert rc == 0, rc
    return bad.value, first.value


def bits(x):
    return int(np.float32(x).view(np.uint32))


def gpu_eval(fn, u):
    u = np.ascontiguousarray(u, np.uint32)
    out = np.zeros(2 * len(u) if fn == ME_SINCOS else len(u), np.uint32)
    rc = _lib().mcfl_math_eval(fn, u.ctypes.data_as(C.POINTER(C.c_uint32)),
                               out.ctypes.data_as(C.POINTER(C.c_uint32)), len(u))
    assert rc == 0, rc
    return out


@pytest.mark.gpu
def test_draw_conversion_every_input(ctx):
    """u32 -> float toward zero (the conversion instruction with the MODE
    register's FP32 rounding set to toward-zero around it) equals the exact
    conversion of u's leading 24 bits for all 2^32 inputs."""
    assert selfcheck(SC_U2F_RZ, 0, 0, 1 << 32) == (0, NONE)


@pytest.mark.gpu
def test_saturating_conversion_every_input(ctx):
    """float -> u32 toward zero (the conversion instruction: NaN and negatives
    0, 2^32 and above 2^32-1) equals the value read off the bits, for all 2^32
    bit patterns."""
    assert selfcheck(SC_F2U_RZ, 0, 0, 1 << 32) == (0, NONE)


@pytest.mark.gpu
@pytest.mark.parametrize("g", [0.0, 0.3, 0.5, 0.8, 0.9, 0.95, 0.99, 0.9999, -0.5, -0.9])
def test_spin_quotient_every_denominator(ctx, g):
    """Spin's (1 - g^2) / (1 - g + 2 g xi) without the division's range
    scaling equals the IEEE quotient for every float denominator between 1-|g|
    and 1+|g| (and a binade either side)."""
    g = np.float32(g)
    a = np.float32(1.0) - g * g  # as the host builds MatDev.onemg2
    lo = bits(min(np.float32(1) - abs(g), np.float32(1) + abs(g)) / 2)
    hi = bits(2 * (np.float32(1) + abs(g))) + 1
    assert selfcheck(SC_DIV, bits(a), lo, hi) == (0, NONE)


@pytest.mark.gpu
@pytest.mark.parametrize("a", [2.0**-20, 1e-4, 0.0199, 0.36, 1.0, 1.5, 1.9999999])
def test_unscaled_division_wide_denominators(ctx, a):
    """The same quotient for numerators across the range and every float
    denominator in [2^-12, 2^12)."""
    assert selfcheck(SC_DIV, bits(a), bits(2.0**-12), bits(2.0**12)) == (0, NONE)


@pytest.mark.gpu
@pytest.mark.parametrize("d", [0.05, 0.1, 0.025, 1.0 / 3.0, 0.0123, 0.5])
def test_pixel_quotient_every_offset(ctx, d):
    """The exit tally's pixel index quotient q / d as (float)((double)q *
    (1/(double)d)) equals the float division for every float offset q in
    [0, 256 d)."""
    assert selfcheck(SC_PIXEL, bits(d), 0, bits(np.float32(256) * np.float32(d))) == (0, NONE)


@pytest.mark.gpu
def test_shared_functions_equal_the_host_build(ctx):
    """Logarithm, reciprocal square root, renormalisation factor, the (0,1]
    draw and sine/cosine of 2 pi draw: the GPU's results equal the oracle's
    host build bit for bit on 4M inputs each (random plus edges)."""
    rng = np.random.default_rng(11)
    n = 1 << 22
    draws = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    draws[:8] = [0, 1, 0x3FFFFFFF, 0x40000000, 0x80000000, 0xC0000000, 0xFFFFFF00, 0xFFFFFFFF]
    # (0,1] draws as the step length's logarithm sees them
    oc = gpu_eval(ME_DRAW_OC, draws).view(np.float32)
    conv = O.math_eval(3, draws.view(np.float32)).astype(np.float64)  # u32 -> float toward zero
    assert np.array_equal(oc, (1.0 - conv * 2.0**-32).astype(np.float32))
    assert np.array_equal(gpu_eval(ME_LOG, oc.view(np.uint32)).view(np.float32), O.math_eval(0, oc))
    pos = np.concatenate([rng.uniform(6e-8, 1.0, n // 2), 10.0 ** rng.uniform(-30, 30, n // 2)]).astype(np.float32)
    assert np.array_equal(gpu_eval(ME_RSQRT, pos.view(np.uint32)).view(np.float32), O.math_eval(4, pos))
    near1 = (1.0 + rng.uniform(-2e-3, 2e-3, n)).astype(np.float32)
    assert np.array_equal(gpu_eval(ME_RENORM, near1.view(np.uint32)).view(np.float32), O.math_eval(6, near1))
    sc = gpu_eval(ME_SINCOS, draws).view(np.float32)
    t = (conv * 2.0**-32).astype(np.float32)  # the [0,1) draw, exact
    assert np.array_equal(sc[0::2], O.math_eval(1, t))
    assert np.array_equal(sc[1::2], O.math_eval(2, t))
