"""The C ABI without a GPU: the library loads, exports every entry point
include/guacho_amd.h declares, lays out gm_params as the header does, and
rejects unsupported configurations before touching the device."""
import ctypes as C
import os
import re

import pytest

from guacho_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "guacho_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"\b(gm_[a-z_0-9]+)\s*\(", src)


def test_header_and_binding_agree():
    names = declared_functions()
    assert set(names) == set(_capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_params_layout():
    # 20 int32 then 6 doubles, no padding
    assert C.sizeof(_capi.GmParams) == 20 * 4 + 6 * 8
    assert _capi.GmParams.dx.offset == 80


def _create(**kw):
    base = dict(nx=8, ny=8, nz=8, dx=0.1, dy=0.1, dz=0.1)
    base.update(kw)
    p = _capi.make_params(base.pop("nx"), base.pop("ny"), base.pop("nz"), base.pop("dx"), base.pop("dy"),
                          base.pop("dz"), **base)
    return p


@pytest.mark.parametrize("kw,needle", [
    (dict(riemann_solver=7), "riemann_solver"),
    (dict(riemann_solver=2, mhd=True), "HLLC"),
    (dict(mhd=False, pmhd=False), "hydro"),
    (dict(slope_limiter=9), "slope_limiter"),
    (dict(bc=(5, 1, 3, 3, 3, 3)), "boundary kind"),
    (dict(bc=(3, 1, 3, 3, 3, 3)), "periodic"),
    (dict(eta=float("nan")), "eta"),
    (dict(nz=1), "at least 2"),
    (dict(dx=0.0), "dx"),
    (dict(precision=7), "PRECISION"),
])
def test_rejects_unsupported_configuration(kw, needle):
    L = _capi.lib()
    p = _create(**kw)
    h = C.c_void_p()
    rc = L.gm_create(C.byref(p), C.byref(h))
    assert rc == _capi.GM_ERR_CONFIG
    assert not h.value
    assert needle.lower() in L.gm_last_error(None).decode().lower()


def test_null_arguments():
    L = _capi.lib()
    assert L.gm_create(None, None) == _capi.GM_ERR_ARG
    assert L.gm_tstep(None, 0.1) == _capi.GM_ERR_ARG
    assert L.gm_destroy(None) == _capi.GM_ERR_ARG


def test_noalloc_refuses_oversized_grid_before_device_work():
    # 200^3 FP64 needs ~1.3 GB of scratch, past the 256 MiB array; refused
    # from the arguments alone, before any device call
    L = _capi.lib()
    p = _create(nx=200, ny=200, nz=200)
    rc = L.gm_tstep_noalloc(C.byref(p), C.c_void_p(1), 1e-3, None, None, None)
    assert rc == _capi.GM_ERR_CONFIG
    assert "too large" in L.gm_last_error(None).decode()
    p = _create(bc=(4, 4, 3, 3, 3, 3))
    assert L.gm_tstep_noalloc(C.byref(p), C.c_void_p(1), 1e-3, None, None, None) == _capi.GM_ERR_CONFIG
    p = _create()
    assert L.gm_tstep_noalloc(C.byref(p), None, 1e-3, None, None, None) == _capi.GM_ERR_ARG


def test_creation_without_a_device_fails_cleanly():
    # valid configurations: without a GPU creation stops at the first HIP
    # call and releases what it built (the sanitizer run covers that path);
    # with one, it succeeds
    L = _capi.lib()
    h = C.c_void_p()
    rc = L.gm_create(C.byref(_create()), C.byref(h))
    if rc == _capi.GM_OK:
        assert L.gm_destroy(h) == _capi.GM_OK
    else:
        assert rc == _capi.GM_ERR_HIP and not h.value and L.gm_last_error(None)
    bp = _capi.GmBatchParams()
    bp.precision, bp.riemann_solver, bp.enable_flux_cd, bp.max_boxes, bp.steps_per_launch = 1, 4, 1, 8, 2
    hb = C.c_void_p()
    rc = L.gm_batch_create(C.byref(bp), C.byref(hb))
    if rc == _capi.GM_OK:
        assert L.gm_batch_destroy(hb) == _capi.GM_OK
    else:
        assert rc == _capi.GM_ERR_HIP and not hb.value and L.gm_batch_last_error(None)
