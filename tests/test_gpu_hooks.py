"""User boundary hooks (impose_user_bc, src/boundaries.f90:245, 508) on the
device against the C restatement.

The test plugin tests/plugins/jet.hip computes the oracle's hook (an
inflow jet through the left x face whose speed depends on globals::time,
and an interior obstacle reset every call); the restatement's boundaryI /
boundaryII with that hook are pinned bit for bit against the reference's
own boundaries.f90 and the Fortran hook (tests/golden fixtures
ref_hlld_minmod_userbc_outflow, ref_hlle_vanleer_userbc_other).  With a
hook installed the device keeps ghosts in memory and runs boundaryII +
hook between the stages and boundaryI + hook at the end of each step, as
tstep does (src/hydro_solver.f90:169, 216); the whole array, ghosts
included, must equal the restatement's bit for bit after every step.
"""
import os
import sys

import numpy as np
import pytest

from guacho_amd import _capi
from guacho_amd.host import timestep_clamp
from tests._pairs import random_state, run_pair

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
JET = (os.path.join(HERE, "plugins", "jet.hsaco"), "gm_test_jet")
GRAVITY = (os.path.join(HERE, "plugins", "gravity.hsaco"), "gm_test_gravity")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # name, (nx, ny, nz), solver, limiter, bc, flux_cd, eight_wave
    ("hlld-outflow-cd", (16, 12, 10), 4, 1, (1,) * 6, True, False),
    ("hlle-other-left-8wave", (12, 10, 8), 3, 2, (4, 1, 2, 2, 3, 3), True, True),
    ("hll-closed-nocd", (10, 12, 8), 1, 4, (2, 2, 2, 2, 1, 1), False, False),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_user_bc_hook_bitwise(case):
    name, (nx, ny, nz), solver, lim, bc, fcd, eight = case
    say(f"case: user bc hook {name} {nx}x{ny}x{nz}, 6 steps, whole array checked each step")
    rng = np.random.default_rng(sum(map(ord, name)))
    u0 = random_state(rng, nx, ny, nz)
    p = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, riemann_solver=solver, slope_limiter=lim,
                          bc=bc, enable_flux_cd=fcd, eight_wave=eight)
    ud, uo, _ = run_pair(p, u0, 6, check_each=True, user_bc=JET)
    np.testing.assert_array_equal(ud, uo)


def test_user_bc_and_sources_together():
    say("case: jet hook + gravity plugin + 8-wave, 14x12x10 outflow, 5 steps")
    rng = np.random.default_rng(21)
    u0 = random_state(rng, 14, 12, 10)
    p = _capi.make_params(14, 12, 10, 1 / 14, 1 / 12, 1 / 10, bc=(1,) * 6, eight_wave=True)
    ud, uo, _ = run_pair(p, u0, 5, check_each=True, user_bc=JET, user_sources=GRAVITY)
    np.testing.assert_array_equal(ud, uo)


def test_boundary_plugin_registered_first():
    """Registering the source plugin must leave an earlier boundary plugin
    loaded (each setter releases only its own module)."""
    say("case: jet hook registered before the gravity plugin, 14x12x10 outflow, 4 steps")
    rng = np.random.default_rng(22)
    u0 = random_state(rng, 14, 12, 10)
    p = _capi.make_params(14, 12, 10, 1 / 14, 1 / 12, 1 / 10, bc=(4, 1, 1, 1, 1, 1))
    ud, uo, _ = run_pair(p, u0, 4, check_each=True, user_bc=JET, user_sources=GRAVITY, bc_first=True)
    np.testing.assert_array_equal(ud, uo)


def test_hook_mixed_against_fp64():
    say("case: jet hook mixed vs FP64, 24^3 outflow, 8 steps")
    rng = np.random.default_rng(4)
    u0 = random_state(rng, 24, 24, 24, contrast=0.1, vmax=0.5, bmax=0.5)
    res = {}
    for prec in (_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED):
        p = _capi.make_params(24, 24, 24, 1 / 24, 1 / 24, 1 / 24, bc=(1,) * 6, precision=prec)
        with _capi.Subdomain(p) as dev:
            dev.set_user_bc(*JET)
            dev.upload_u(u0)
            dev.set_time(0.0)
            dev.boundary_u()
            t = 0.0
            for it in range(1, 9):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
                dev.set_time(t)
                dev.tstep(dt)
                t += dt
            res[prec] = dev.download_u()[:, 2:-2, 2:-2, 2:-2]
    a, b = res[_capi.GM_PRECISION_FP64], res[_capi.GM_PRECISION_MIXED]
    assert np.all(np.isfinite(b))
    rel = np.abs(a - b).sum(axis=(1, 2, 3)) / np.maximum(np.abs(a).sum(axis=(1, 2, 3)), 1e-300)
    assert rel[0] <= 1e-5 and np.all(rel <= 1e-4), rel


def test_bc_other_needs_a_hook():
    say("case: BC_OTHER face without a hook is refused at the step")
    p = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8, bc=(4, 1, 1, 1, 1, 1))
    rng = np.random.default_rng(1)
    with _capi.Subdomain(p) as dev:
        dev.upload_u(random_state(rng, 8, 8, 8))
        with pytest.raises(_capi.GmError) as ei:
            dev.tstep(1e-3)
        assert ei.value.code == _capi.GM_ERR_CONFIG
        dev.set_user_bc(*JET)
        dev.tstep(1e-3)
        dev.sync()


def test_batch_refuses_bc_other():
    say("case: batches refuse BC_OTHER boxes")
    p = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8, bc=(4, 1, 1, 1, 1, 1))
    rng = np.random.default_rng(2)
    with _capi.Batch(precision=_capi.GM_PRECISION_FP64, max_boxes=2) as B:
        with pytest.raises(_capi.GmError) as ei:
            B.add(p, random_state(rng, 8, 8, 8), max_steps=1)
        assert ei.value.code == _capi.GM_ERR_CONFIG


@pytest.mark.parametrize("prec", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED], ids=["fp64", "mixed"])
def test_downloads_between_steps_change_nothing(prec):
    """Temperature and div B downloads stage their output in device memory;
    the E ghosts a BC_OTHER face reads must not be that memory."""
    say(f"case: jet hook (BC_OTHER x face), downloads between steps, precision {prec}, 12^3, 5 steps")
    rng = np.random.default_rng(25)
    u0 = random_state(rng, 12, 12, 12)
    p = _capi.make_params(12, 12, 12, 1 / 12, 1 / 12, 1 / 12, bc=(4, 1, 1, 1, 1, 1), precision=prec)

    def run(downloads):
        with _capi.Subdomain(p) as dev:
            dev.set_user_bc(*JET)
            dev.upload_u(u0)
            dev.boundary_u()
            t = 0.0
            for it in range(1, 6):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
                dev.set_time(t)
                dev.tstep(dt)
                t += dt
                if downloads and it in (2, 3):
                    dev.download_primit(with_temp=True)
                    dev.divb()
            return dev.download_u()

    np.testing.assert_array_equal(run(True), run(False))


def test_hook_sources_eight_wave_viscosity_inline():
    """Without passive scalars the source terms are added by the stage
    kernels: a boundary hook, the gravity plugin, 8-wave and viscosity in
    one run, bit for bit against the oracle every step."""
    say("case: jet hook + gravity + 8-wave + eta, npas=0, 14x12x10, 5 steps, checked each step")
    rng = np.random.default_rng(27)
    u0 = random_state(rng, 14, 12, 10)
    p = _capi.make_params(14, 12, 10, 1 / 14, 1 / 12, 1 / 10, bc=(4, 1, 1, 1, 2, 2), eight_wave=True, eta=0.02)
    ud, uo, _ = run_pair(p, u0, 5, check_each=True, user_bc=JET, user_sources=GRAVITY)
    np.testing.assert_array_equal(ud, uo)
