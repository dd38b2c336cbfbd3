"""Artificial viscosity, viscous_copy (src/hydro_solver.f90:47-65), on the
device against the C restatement.

u = up + eta (sum of the six face neighbours of up - 6 up) after the
corrector.  up's ghosts at that point hold boundaryII's fill of the
predictor state, so the cells next to each face mix the new state with the
predictor's ghosts; the device reads those through the boundary mapping
(or as stored, with ghosts in memory under a user boundary hook).  The
restatement's viscous_copy is pinned by closed forms
(tests/test_oracle_invariants.py); here the FP64 device path must equal it
bit for bit after every step.
"""
import os
import sys

import numpy as np
import pytest

from guacho_amd import _capi
from guacho_amd.host import timestep_clamp
from oracle.oracle import Oracle
from tests._pairs import oracle_config, random_state, run_pair

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
JET = (os.path.join(HERE, "plugins", "jet.hsaco"), "gm_test_jet")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # name, (nx, ny, nz), solver, limiter, bc, flux_cd, eta
    ("hlld-minmod-periodic", (16, 12, 10), 4, 1, (3,) * 6, True, 0.02),
    ("hlle-vanleer-outflow", (12, 10, 8), 3, 2, (1,) * 6, False, 0.05),
    ("hll-umist-closed", (10, 12, 8), 1, 4, (2,) * 6, True, 0.01),
    ("hlld-woodward-mixed", (12, 8, 10), 4, 5, (2, 2, 1, 1, 3, 3), True, 0.1),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_viscous_copy_bitwise(case):
    name, (nx, ny, nz), solver, lim, bc, fcd, eta = case
    say(f"case: viscous_copy {name} {nx}x{ny}x{nz} eta={eta}, 6 steps, checked each step")
    rng = np.random.default_rng(sum(map(ord, name)))
    u0 = random_state(rng, nx, ny, nz)
    p = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, riemann_solver=solver, slope_limiter=lim,
                          bc=bc, enable_flux_cd=fcd, eta=eta)
    ud, uo, _ = run_pair(p, u0, 6, check_each=True)
    np.testing.assert_array_equal(ud, uo)


def test_viscous_copy_with_user_bc_hook():
    """Ghosts in memory (a user boundary hook): the corrector reads U*'s
    stored ghosts after boundaryII + hook, and so does viscous_copy."""
    say("case: viscous_copy eta=0.03 + jet hook, 14x12x10 outflow, 5 steps, whole array each step")
    rng = np.random.default_rng(41)
    u0 = random_state(rng, 14, 12, 10)
    p = _capi.make_params(14, 12, 10, 1 / 14, 1 / 12, 1 / 10, bc=(4, 1, 1, 1, 2, 2), eta=0.03)
    ud, uo, _ = run_pair(p, u0, 5, check_each=True, user_bc=JET)
    np.testing.assert_array_equal(ud, uo)


def test_viscosity_changes_the_result():
    say("case: eta=0.05 vs eta=0, 12^3 periodic, one step")
    rng = np.random.default_rng(3)
    u0 = random_state(rng, 12, 12, 12)
    outs = []
    for eta in (0.0, 0.05):
        p = _capi.make_params(12, 12, 12, 1 / 12, 1 / 12, 1 / 12, eta=eta)
        with _capi.Subdomain(p) as dev:
            dev.upload_u(u0)
            dev.tstep(1e-3)
            outs.append(dev.download_u()[:, 2:-2, 2:-2, 2:-2])
    d = np.abs(outs[1] - outs[0])
    assert d.max() > 1e-3
    # smoothing: the Laplacian pass lowers the grid-scale variation of rho
    rough = [np.abs(np.diff(o[0], axis=0)).sum() for o in outs]
    assert rough[1] < rough[0]


def test_viscous_mixed_against_fp64():
    say("case: eta=0.02 mixed vs FP64, 24^3 periodic, 8 steps")
    rng = np.random.default_rng(4)
    u0 = random_state(rng, 24, 24, 24, contrast=0.1, vmax=0.5, bmax=0.5)
    res = {}
    for prec in (_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED):
        p = _capi.make_params(24, 24, 24, 1 / 24, 1 / 24, 1 / 24, precision=prec, eta=0.02)
        with _capi.Subdomain(p) as dev:
            dev.upload_u(u0)
            t = 0.0
            for it in range(1, 9):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
                dev.tstep(dt)
                t += dt
            res[prec] = dev.download_u()[:, 2:-2, 2:-2, 2:-2]
    a, b = res[_capi.GM_PRECISION_FP64], res[_capi.GM_PRECISION_MIXED]
    assert np.all(np.isfinite(b))
    rel = np.abs(a - b).sum(axis=(1, 2, 3)) / np.maximum(np.abs(a).sum(axis=(1, 2, 3)), 1e-300)
    assert rel[0] <= 1e-5 and np.all(rel <= 1e-4), rel


def test_viscosity_then_table_cooling():
    """viscous_copy runs before cooling (src/hydro_solver.f90:188, 206-210):
    one step with both against the oracle; every field but E bit for bit, E
    to 1e-14 (the device's exp and log10)."""
    say("case: eta = 0.02 + DMC cooling, 16^3 outflow, 1 step")
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_cooling.npz")
    tab = np.load(fixture)["dmc_table"]
    tempsc, psc, tsc = 1e6, 1e-22, 10.0
    rng = np.random.default_rng(61)
    n = 16
    u0 = random_state(rng, n, n, n)
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, bc=(1,) * 6, eta=0.02)
    O = Oracle(oracle_config(p))
    O.set_cooling(_capi.GM_COOL_DMC, tab, tempsc, psc, tsc)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    with _capi.Subdomain(p) as dev:
        dev.set_cooling(_capi.GM_COOL_DMC, tab, tempsc, psc, tsc)
        dev.upload_u(u0)
        dev.boundary_u()
        dt, _ = timestep_clamp(1, 10, 0.4, O.cfl_dtp(), 0.0, 1e30)
        assert dev.cfl_dtp() == O.cfl_dtp()
        O.tstep(dt)
        dev.tstep(dt)
        ud = dev.download_u()
    uo = O.get_u()
    O.close()
    for q in (0, 1, 2, 3, 5, 6, 7):
        np.testing.assert_array_equal(ud[q], uo[q], err_msg=f"field {q}")
    np.testing.assert_allclose(ud[4], uo[4], rtol=1e-14, atol=0)
