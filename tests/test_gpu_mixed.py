"""Mixed-precision form (FP32 states, reconstruction and fluxes; FP64
pressure recovery, update accumulation and CFL reduction) against the FP64
path, which equals the restatement of the reference bit for bit.

Tolerances (the mixed form's contract): relative L1 of density <= 1e-5
after the stated steps; every other field <= 1e-4; periodic mass and energy
drift <= 1e-6 relative; central-difference div B stays below 1e-5 of
|B|/dx."""
import sys

import numpy as np
import pytest

from guacho_amd import _capi, problems
from guacho_amd.host import timestep_clamp

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(params, u0, nsteps, cfl=0.4):
    """Advance nsteps with the path's own CFL; returns (u, dts)."""
    dts = []
    with _capi.Subdomain(params) as dev:
        dev.upload_u(u0)
        dev.boundary_u()
        t = 0.0
        for it in range(1, nsteps + 1):
            dt, _ = timestep_clamp(it, 10, cfl, dev.cfl_dtp(), t, 1e30)
            dev.tstep(dt)
            t += dt
            dts.append(dt)
        u = dev.download_u()
        divb = dev.divb()
    return u, dts, divb


def rel_l1(a, b):
    a = a[:, 2:-2, 2:-2, 2:-2]
    b = b[:, 2:-2, 2:-2, 2:-2]
    out = []
    for q in range(8):
        den = np.abs(b[q]).sum()
        out.append(np.abs(a[q] - b[q]).sum() / den if den > 0 else float(np.abs(a[q]).sum()))
    return np.array(out)


def both(n, u0, dx, dy, dz, nsteps, **kw):
    p64 = _capi.make_params(n[0], n[1], n[2], dx, dy, dz, **kw)
    p32 = _capi.make_params(n[0], n[1], n[2], dx, dy, dz, precision=_capi.GM_PRECISION_MIXED, **kw)
    u64, dt64, _ = run(p64, u0, nsteps)
    u32, dt32, divb32 = run(p32, u0, nsteps)
    return u64, u32, dt64, dt32, divb32


def test_mixed_turbulence_64_errors():
    say("case: mixed vs FP64, turbulence 64^3 HLLD minmod, 20 steps")
    u0, dx, dy, dz = problems.turbulence(64, seed=5)
    u64, u32, dt64, dt32, divb = both((64, 64, 64), u0, dx, dy, dz, 20)
    e = rel_l1(u32, u64)
    say(f"relative L1 per field: {e}")
    assert e[0] <= 1e-5, e
    assert np.all(e <= 1e-4), e
    np.testing.assert_allclose(dt32, dt64, rtol=1e-5)
    m0 = u0[:, 2:-2, 2:-2, 2:-2].sum(axis=(1, 2, 3))
    m1 = u32[:, 2:-2, 2:-2, 2:-2].sum(axis=(1, 2, 3))
    assert abs(m1[0] - m0[0]) <= 1e-6 * abs(m0[0])
    assert abs(m1[4] - m0[4]) <= 1e-6 * abs(m0[4])
    bscale = np.sqrt((u0[5:8] ** 2).sum(0)).max() / dx
    assert np.abs(divb).max() <= 1e-5 * bscale


def test_mixed_orszag_tang_64_errors():
    say("case: mixed vs FP64, OT 64x64x4, 60 steps")
    u0, dx, dy, dz = problems.orszag_tang(64, 64, 4)
    u64, u32, _, _, _ = both((64, 64, 4), u0, dx, dy, dz, 60)
    e = rel_l1(u32, u64)
    say(f"relative L1 per field: {e}")
    assert e[0] <= 1e-5, e
    # rho*w and Bz of the 2D problem stay exactly zero
    assert np.all(u32[3] == 0) and np.all(u32[7] == 0)


@pytest.mark.parametrize("solver,mhd,pmhd,bc", [(3, 1, 0, (1,) * 6), (1, 1, 0, (2,) * 6), (2, 0, 1, (3,) * 6)])
def test_mixed_other_solvers_close(solver, mhd, pmhd, bc):
    say(f"case: mixed vs FP64, blast 24^3 solver {solver} bc {bc[0]}, 8 steps")
    u0, dx, dy, dz = problems.blast(24, seed=9)
    u64, u32, _, _, _ = both((24, 24, 24), u0, dx, dy, dz, 8, riemann_solver=solver, mhd=bool(mhd),
                             pmhd=bool(pmhd), bc=bc)
    e = rel_l1(u32, u64)
    say(f"relative L1 per field: {e}")
    assert e[0] <= 1e-5, e
    assert np.all(np.isfinite(u32))


@pytest.mark.parametrize("fcd", [False, True], ids=["8wave", "8wave-fluxcd"])
def test_mixed_eight_wave_close(fcd):
    """The mixed form with Powell's terms added by the stage kernels."""
    say(f"case: mixed vs FP64, turbulence 32^3 8-wave (flux-CD {fcd}), 12 steps")
    u0, dx, dy, dz = problems.turbulence(32, seed=13)
    u64, u32, _, _, _ = both((32, 32, 32), u0, dx, dy, dz, 12, eight_wave=True, enable_flux_cd=fcd)
    e = rel_l1(u32, u64)
    say(f"relative L1 per field: {e}")
    assert e[0] <= 1e-5, e
    assert np.all(e <= 1e-4), e


def test_mixed_user_sources_close():
    """The mixed form with a user source plugin (its s added by the stage
    kernels) against FP64 with the same plugin."""
    import os
    plugin = (os.path.join(os.path.dirname(os.path.abspath(__file__)), "plugins", "gravity.hsaco"),
              "gm_test_gravity")
    say("case: mixed vs FP64, blast 24^3 outflow + gravity plugin + 8-wave, 10 steps")
    u0, dx, dy, dz = problems.blast(24, seed=21)
    out = {}
    for prec in (_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED):
        p = _capi.make_params(24, 24, 24, dx, dy, dz, bc=(1,) * 6, eight_wave=True, precision=prec)
        with _capi.Subdomain(p) as dev:
            dev.set_user_sources(*plugin)
            dev.upload_u(u0)
            dev.boundary_u()
            t = 0.0
            for it in range(1, 11):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
                dev.set_time(t)
                dev.tstep(dt)
                t += dt
            out[prec] = dev.download_u()
    e = rel_l1(out[_capi.GM_PRECISION_MIXED], out[_capi.GM_PRECISION_FP64])
    say(f"relative L1 per field: {e}")
    assert e[0] <= 1e-5, e
    assert np.all(e <= 1e-4), e
