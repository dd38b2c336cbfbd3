"""gm_div and gm_sqrt (include/guacho_amd_user.h), the FP64 division and
square root of every kernel on the path, against the compiler's IEEE
division and sqrt on the GPU, bit for bit.

gm_div is the correctly rounded division sequence without its operand
scaling, which only acts when |a| < 2^-969, b or a/b is subnormal,
|b| > 2^1021 or exp(a) - exp(b) >= 768; gm_sqrt drops the prescaling below
2^-767.  Random operands over exponents -900..900 outside those ranges,
operands of unit size, operands with one mantissa bit set or cleared
(quotients next to halfway cases), and zeros, infinities and NaNs as either
operand must give the IEEE results exactly (NaNs compare as NaNs).
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HSACO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plugins", "arith.hsaco")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hip():
    hip = C.CDLL("libamdhip64.so")
    for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipModuleLoad", "hipModuleGetFunction",
                 "hipModuleLaunchKernel", "hipDeviceSynchronize", "hipModuleUnload"):
        getattr(hip, name).restype = C.c_int
    return hip


def _run(a, b):
    """(gm_div, a / b, gm_sqrt(a), sqrt(a)) on the GPU, one launch."""
    hip = _hip()
    n = a.size
    da, db, do = C.c_void_p(), C.c_void_p(), C.c_void_p()
    mod, fn = C.c_void_p(), C.c_void_p()
    assert hip.hipModuleLoad(C.byref(mod), HSACO.encode()) == 0
    try:
        assert hip.hipModuleGetFunction(C.byref(fn), mod, b"gm_test_div_sqrt") == 0
        for p, nb in ((da, 8 * n), (db, 8 * n), (do, 32 * n)):
            assert hip.hipMalloc(C.byref(p), C.c_size_t(nb)) == 0
        try:
            assert hip.hipMemcpy(da, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * n), 1) == 0
            assert hip.hipMemcpy(db, b.ctypes.data_as(C.c_void_p), C.c_size_t(8 * n), 1) == 0
            nn = C.c_longlong(n)
            args = (C.c_void_p * 4)(C.cast(C.byref(da), C.c_void_p), C.cast(C.byref(db), C.c_void_p),
                                    C.cast(C.byref(do), C.c_void_p), C.cast(C.byref(nn), C.c_void_p))
            grid = (n + 255) // 256
            assert hip.hipModuleLaunchKernel(fn, grid, 1, 1, 256, 1, 1, 0, None, args, None) == 0
            assert hip.hipDeviceSynchronize() == 0
            out = np.empty(4 * n)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), do, C.c_size_t(32 * n), 2) == 0
        finally:
            for p in (da, db, do):
                hip.hipFree(p)
    finally:
        hip.hipModuleUnload(mod)
    return out.reshape(n, 4).T


def _same(x, y):
    bx, by = x.view(np.uint64), y.view(np.uint64)
    return (bx == by) | (np.isnan(x) & np.isnan(y))


def _operands(rng):
    n = 1 << 20
    m_a = rng.uniform(1.0, 2.0, n)
    m_b = rng.uniform(1.0, 2.0, n)
    e_a = rng.integers(-900, 901, n)
    e_b = rng.integers(-900, 901, n)
    keep = (e_a - e_b > -1000) & (e_a - e_b < 760)
    a = np.ldexp(m_a, e_a)[keep] * rng.choice([-1.0, 1.0], keep.sum())
    b = np.ldexp(m_b, e_b)[keep] * rng.choice([-1.0, 1.0], keep.sum())
    # unit-sized operands, as the path sees them
    u = rng.uniform(-4.0, 4.0, (2, 1 << 18))
    # one mantissa bit flipped: quotients next to exactly representable ones
    q = rng.uniform(1.0, 2.0, 1 << 16)
    d = rng.uniform(1.0, 2.0, 1 << 16)
    p = q * d
    flip = p.view(np.uint64) ^ np.uint64(1)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 3.0, 1e-300, 1e300])
    sa, sb = np.meshgrid(special, special)
    a = np.concatenate([a, u[0], p, flip.view(np.float64), sa.ravel()])
    b = np.concatenate([b, u[1], d, d, sb.ravel()])
    return a, b


def test_div_sqrt_equal_ieee():
    say("case: gm_div / gm_sqrt vs IEEE division and sqrt, ~1.1M operand pairs")
    rng = np.random.default_rng(2026)
    a, b = _operands(rng)
    fast_q, ieee_q, fast_s, ieee_s = _run(a, b)
    ok = _same(fast_q, ieee_q)
    assert ok.all(), (a[~ok][:5], b[~ok][:5], fast_q[~ok][:5], ieee_q[~ok][:5])
    # sqrt: operands from 2^-767 up (and zeros, infinities, NaN, negatives)
    inside = ~(np.abs(a) < 2.0 ** -767) | (a == 0)
    oks = _same(fast_s, ieee_s) | ~inside
    assert oks.all(), (a[~oks][:5], fast_s[~oks][:5], ieee_s[~oks][:5])
    # and the host's IEEE division agrees with the device's
    with np.errstate(all="ignore"):
        host = a / b
    assert _same(ieee_q, host).all()
