"""gm_tstep_noalloc (the step on a caller-held device state, with its
buffers in the library's static device array) against gm_tstep through a
context: the interior state after the step and the next CFL minimum equal
bit for bit, on two calls in a row, the second on a smaller grid (the
array then holds the first call's values), in FP64 and in mixed
precision, with and without flux-CD."""
import ctypes as C

import numpy as np
import pytest
import torch

from guacho_amd import _capi, problems
from guacho_amd.host import timestep_clamp

pytestmark = pytest.mark.gpu


def to_device_layout(u):
    """Fortran u(neq, i, j, k) with ghosts -> device SoA [q][k][j][i]."""
    return np.ascontiguousarray(np.transpose(u, (0, 3, 2, 1)))


def from_device_layout(d):
    return np.asfortranarray(np.transpose(d, (0, 3, 2, 1)))


def context_step(prm, u0):
    with _capi.Subdomain(prm) as dev:
        dev.upload_u(u0)
        dtp0 = dev.cfl_dtp()
        dt, _ = timestep_clamp(11, 10, 0.4, dtp0, 0.0, 1e30)
        dev.tstep(dt)
        u1 = dev.download_u()
        dtp1 = dev.cfl_dtp()
    return dt, u1, dtp1


def noalloc_step(prm, u0, dt):
    L = _capi.lib()
    fp64 = prm.precision == _capi.GM_PRECISION_FP64
    dtype = torch.float64 if fp64 else torch.float32
    host = to_device_layout(u0).astype(np.float64 if fp64 else np.float32)
    ud = torch.from_numpy(host).to("cuda:0")
    dtp = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    err = torch.full((1,), -7, dtype=torch.int32, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.gm_tstep_noalloc(C.byref(prm), C.c_void_p(ud.data_ptr()), dt, C.c_void_p(stream),
                            C.c_void_p(dtp.data_ptr()), C.c_void_p(err.data_ptr()))
    assert rc == 0, L.gm_last_error(None).decode()
    torch.cuda.synchronize()
    assert ud.dtype == dtype
    return from_device_layout(ud.double().cpu().numpy()), float(dtp.item()), int(err.item())


@pytest.mark.parametrize("precision", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED])
@pytest.mark.parametrize("flux_cd", [True, False])
def test_noalloc_equals_context_step(precision, flux_cd):
    cases = []
    u, dx, dy, dz = problems.turbulence(40, seed=11)
    cases.append(("turbulence 40^3", u, _capi.make_params(40, 40, 40, dx, dy, dz, precision=precision,
                                                            enable_flux_cd=flux_cd)))
    u, dx, dy, dz = problems.orszag_tang(36, 28, 6)
    cases.append(("Orszag-Tang 36x28x6", u, _capi.make_params(36, 28, 6, dx, dy, dz, precision=precision,
                                                                enable_flux_cd=flux_cd)))
    for name, u0, prm in cases:
        print(f"case: gm_tstep_noalloc vs gm_tstep, {name}, precision {precision}, flux-CD {flux_cd}", flush=True)
        dt, u_ctx, dtp_ctx = context_step(prm, u0)
        u_na, dtp_na, err = noalloc_step(prm, u0, dt)
        assert err == 0
        inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
        np.testing.assert_array_equal(u_na[inner], u_ctx[inner], err_msg=name)
        assert dtp_na == dtp_ctx, (name, dtp_na, dtp_ctx)


def test_noalloc_refuses_a_grid_past_the_limit():
    L = _capi.lib()
    prm = _capi.make_params(160, 160, 160, 0.01, 0.01, 0.01)
    rc = L.gm_tstep_noalloc(C.byref(prm), C.c_void_p(1), 1e-3, None, None, None)
    assert rc == _capi.GM_ERR_CONFIG
    assert "too large" in L.gm_last_error(None).decode()


def test_noalloc_step_captured_in_a_graph():
    """What the entry point is for: captured in a HIP graph (torch's capture
    stream; the call enqueues kernels only), each replay equals the eager
    call bit for bit -- the state and the next CFL minimum -- with eager
    work and synchronisation between the replays."""
    print("case: gm_tstep_noalloc captured in a graph, 3 replays vs 3 eager steps, turbulence 24^3 FP64", flush=True)
    L = _capi.lib()
    u0, dx, dy, dz = problems.turbulence(24, seed=21)
    prm = _capi.make_params(24, 24, 24, dx, dy, dz)
    dt = context_step(prm, u0)[0]
    host = to_device_layout(u0)
    ue = torch.from_numpy(host.copy()).to("cuda:0")
    ug = torch.from_numpy(host.copy()).to("cuda:0")
    dtp_e = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    dtp_g = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    err = torch.zeros(1, dtype=torch.int32, device="cuda:0")

    def step(u, dtp):
        rc = L.gm_tstep_noalloc(C.byref(prm), C.c_void_p(u.data_ptr()), dt,
                                C.c_void_p(torch.cuda.current_stream().cuda_stream), C.c_void_p(dtp.data_ptr()),
                                C.c_void_p(err.data_ptr()))
        assert rc == 0, L.gm_last_error(None).decode()

    scratch = torch.zeros_like(ue)
    step(scratch, dtp_e)            # first call on the device: scratch address resolved outside capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step(ug, dtp_g)
    for rep in range(3):
        g.replay()
        torch.cuda.synchronize()
        step(ue, dtp_e)             # eager work between the replays
        torch.cuda.synchronize()
        assert int(err.item()) == 0
        assert torch.equal(ug, ue), f"replay {rep}"
        assert dtp_g.item() == dtp_e.item(), f"replay {rep}"
