"""Device memory exhaustion in gm_create.

A grid that passes the size check but does not fit the GPU's memory fails
with GM_ERR_OOM and a message; the buffers allocated before the failing
one are freed, so free device memory is what it was, and a context of a
grid that fits is created and stepped afterwards.
"""
import ctypes as C
import ctypes.util
import os
import sys

import numpy as np
import pytest

from guacho_amd import _capi
from tests._pairs import random_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hip():
    for name in (ctypes.util.find_library("amdhip64"), "/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        if name and (os.path.isabs(name) is False or os.path.exists(name)):
            try:
                return C.CDLL(name)
            except OSError:
                continue
    pytest.fail("libamdhip64 not found")


def _free_bytes():
    free, total = C.c_size_t(), C.c_size_t()
    assert _hip().hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value, total.value


def _params(n, prec=_capi.GM_PRECISION_FP64):
    return _capi.make_params(n, n, n, 1.0 / n, 1.0 / n, 1.0 / n, precision=prec)


def test_grid_larger_than_the_gpu_fails_cleanly():
    _capi.lib()
    free0, total = _free_bytes()
    # the conserved state alone, (n + 4)^3 cells x 8 variables x 8 bytes,
    # exceeds the device: the first allocation fails
    n = int(round((2.0 * total / 64.0) ** (1.0 / 3.0))) - 4
    say(f"case: gm_create {n}^3 FP64 on a {total / 2**30:.0f} GiB device (state alone {(n + 4)**3 * 64 / 2**30:.0f} GiB)")
    assert (n + 4) ** 3 <= 1 << 36  # passes the size check, so the allocation is what fails
    with pytest.raises(_capi.GmError) as ei:
        _capi.Subdomain(_params(n))
    assert ei.value.code == _capi.GM_ERR_OOM
    assert "allocation failed" in ei.value.message
    free1, _ = _free_bytes()
    assert abs(free1 - free0) < 256 << 20


def test_failure_after_a_partial_allocation_frees_it():
    _capi.lib()
    free0, total = _free_bytes()
    # U (8 x 8 bytes a cell) takes 60% of the free memory and fits; U* does not
    n = int((0.6 * free0 / 64.0) ** (1.0 / 3.0)) - 4
    say(f"case: gm_create {n}^3 FP64: U {(n + 4)**3 * 64 / 2**30:.0f} GiB of {free0 / 2**30:.0f} GiB free, U* cannot follow")
    with pytest.raises(_capi.GmError) as ei:
        _capi.Subdomain(_params(n))
    assert ei.value.code == _capi.GM_ERR_OOM
    free1, _ = _free_bytes()
    assert abs(free1 - free0) < 256 << 20, (free0, free1)
    say("case: a 24^3 context after the failure, two steps")
    u = random_state(np.random.default_rng(5), 24, 24, 24)
    with _capi.Subdomain(_params(24)) as dev:
        dev.upload_u(u)
        for _ in range(2):
            dev.tstep(0.2 * dev.cfl_dtp())
        out = dev.download_u()
    assert np.isfinite(out).all()
