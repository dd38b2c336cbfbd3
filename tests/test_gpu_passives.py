"""Passive scalars (PASSIVES builds: neq = 8 + npas), EOS_H_RATE and
H-rate cooling on the device against the C restatement.

The restatement's passive fluxes (all four solvers), limiter, u2prim with
EOS_H_RATE, sweeps and boundaries equal the reference's own routines bit
for bit (tests/golden fixtures ref_hlld_minmod_pas2_hrate,
ref_hlle_vanleer_pas1_outflow, ref_hll_umist_pas2_closed,
ref_hllc_superbee_pmhd_pas1_hrate), and its coolingh equals the
reference's (ref_cooling_h.npz).  Here the FP64 device step with passive
scalars must equal the restatement bit for bit after every step.  H-rate
cooling runs its rates in FP64 where the reference uses REAL(16) constants,
and the device's exp/pow/log are its own: after one cooled step every field
but E and the neutral density is bit for bit and those two agree to a
relative 1e-10 (the neutral fraction's closed-form solution,
(-b - d (1 + g e)/(1 - g e)) / 2a, cancels and lifts last-bit differences
of the rates to ~1e-11).
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

HR = _capi.GM_EOS_H_RATE
HERE = os.path.dirname(os.path.abspath(__file__))
JET = (os.path.join(HERE, "plugins", "jet.hsaco"), "gm_test_jet")
GRAVITY = (os.path.join(HERE, "plugins", "gravity.hsaco"), "gm_test_gravity")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # name, (nx, ny, nz), solver, limiter, bc, flux_cd, npas, eos
    ("hlld-minmod-periodic-pas2-hrate", (16, 12, 10), 4, 1, (3,) * 6, True, 2, (HR, 1.0e4)),
    ("hlld-minmod-outflow-pas1", (12, 10, 8), 4, 1, (1,) * 6, True, 1, (1, 1.0)),
    ("hlle-vanleer-closed-pas2-nocd", (12, 10, 8), 3, 2, (2, 2, 1, 1, 3, 3), False, 2, (1, 1.0)),
    ("hll-umist-mixed-pas1-hrate", (10, 12, 8), 1, 4, (3, 3, 2, 2, 1, 1), True, 1, (HR, 2.0e3)),
    ("hllc-superbee-pmhd-pas2", (12, 8, 10), 2, 6, (1,) * 6, False, 2, (HR, 1.0e4)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_passives_bitwise(case):
    name, (nx, ny, nz), solver, lim, bc, fcd, npas, eos = case
    say(f"case: passives {name} {nx}x{ny}x{nz}, 6 steps, checked each step")
    rng = np.random.default_rng(sum(map(ord, name)))
    u0 = random_state(rng, nx, ny, nz, npas=npas)
    pmhd = solver == 2
    p = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, riemann_solver=solver, slope_limiter=lim,
                          bc=bc, enable_flux_cd=fcd, mhd=not pmhd, pmhd=pmhd, npas=npas)
    ud, uo, _ = run_pair(p, u0, 6, check_each=True, eos=eos)
    np.testing.assert_array_equal(ud, uo)
    inner = (slice(8, None),) + (slice(2, -2),) * 3
    assert np.abs(ud[inner] - u0[inner]).max() > 1e-3      # the scalars were advected


def test_passives_with_hooks_sources_viscosity():
    """Ghosts in memory (jet hook), a user source plugin, 8-wave terms and
    viscous_copy, with two passive scalars: all of them carry the passives."""
    say("case: passives npas=2 + jet hook + gravity + 8-wave + eta, 14x12x10, 5 steps, whole array")
    rng = np.random.default_rng(23)
    u0 = random_state(rng, 14, 12, 10, npas=2)
    p = _capi.make_params(14, 12, 10, 1 / 14, 1 / 12, 1 / 10, bc=(4, 1, 1, 1, 2, 2), eight_wave=True, eta=0.02,
                          npas=2)
    ud, uo, _ = run_pair(p, u0, 5, check_each=True, user_bc=JET, user_sources=GRAVITY)
    np.testing.assert_array_equal(ud, uo)


def test_temperature_download_h_rate():
    say("case: primitives and temperature under EOS_H_RATE, npas=2, 10x8x6 outflow")
    rng = np.random.default_rng(29)
    u0 = random_state(rng, 10, 8, 6, npas=2)
    p = _capi.make_params(10, 8, 6, 0.1, 0.125, 1 / 6, bc=(1,) * 6, npas=2)
    O = Oracle(oracle_config(p, eos=(HR, 1.0e4)))
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    prim_o, t_o = O.get_primit(), O.get_temp()
    O.close()
    with _capi.Subdomain(p) as dev:
        dev.set_eos(HR, 1.0e4)
        dev.upload_u(u0)
        dev.boundary_u()
        prim_d, t_d = dev.download_primit(with_temp=True)
    np.testing.assert_array_equal(prim_d, prim_o)
    np.testing.assert_array_equal(t_d, t_o)


def _h_state(rng, n, tempsc):
    """Random MHD state with the neutral density as the first passive
    scalar and T = 5e3..2e5 K under EOS_H_RATE."""
    shape = (n + 4,) * 3
    u = random_state(rng, n, n, n, npas=2, vmax=0.3, bmax=0.3)
    rho = u[0]
    t = 10.0 ** rng.uniform(np.log10(5e3), np.log10(2e5), shape)
    pr = (2.0 * rho - u[8]) * t / tempsc
    v = u[1:4] / rho
    u[4] = 0.5 * rho * (v ** 2).sum(0) + 1.5 * pr + 0.5 * (u[5:8] ** 2).sum(0)
    return u


@pytest.mark.parametrize("nsteps", [1, 4])
def test_h_rate_cooling(nsteps):
    say(f"case: H-rate cooling + EOS_H_RATE, npas=2, 12^3 periodic, {nsteps} step(s)")
    tempsc, tsc = 1.0e4, 1.0e13
    n = 12
    rng = np.random.default_rng(31 + nsteps)
    u0 = _h_state(rng, n, tempsc)
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, npas=2)
    O = Oracle(oracle_config(p, eos=(HR, tempsc)))
    O.set_cooling(_capi.GM_COOL_H, None, tempsc, 1.0, tsc)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    with _capi.Subdomain(p) as dev:
        dev.set_eos(HR, tempsc)
        dev.set_cooling(_capi.GM_COOL_H, None, tempsc, 1.0, tsc)
        dev.upload_u(u0)
        dev.boundary_u()
        t = 0.0
        for it in range(1, nsteps + 1):
            dt, _ = timestep_clamp(it, 10, 0.4, O.cfl_dtp(), t, 1e30)
            O.tstep(dt)
            dev.tstep(dt)
            t += dt
        ud = dev.download_u()
        dev.set_cooling(_capi.GM_COOL_NONE)
        dev.upload_u(u0)
        dev.boundary_u()
        dev.tstep(timestep_clamp(1, 10, 0.4, dev.cfl_dtp(), 0.0, 1e30)[0])
        u_off = dev.download_u()
    uo = O.get_u()
    O.close()
    inner = (slice(2, -2),) * 3
    if nsteps == 1:
        for q in (0, 1, 2, 3, 5, 6, 7, 9):
            np.testing.assert_array_equal(ud[q], uo[q], err_msg=f"field {q}")
        for q in (4, 8):
            np.testing.assert_allclose(ud[q], uo[q], rtol=1e-10, atol=0, err_msg=f"field {q}")
        moved = np.abs(ud[8][inner] - u_off[8][inner]) / u_off[8][inner]
        assert moved.max() > 1e-3                   # the neutral fraction moved
    else:
        scale = np.abs(uo).max(axis=(1, 2, 3), keepdims=True)
        assert np.max(np.abs(ud - uo) / scale) <= 1e-9


def test_h_rate_cooling_adiabatic_eos():
    """coolingh reads u2prim's temperature, so under EOS_ADIABATIC it cools
    with T = (p / rho) Tempsc; one step against the oracle."""
    say("case: H-rate cooling + EOS_ADIABATIC, npas=1, 12^3 periodic, 1 step")
    tempsc, tsc = 1.0e4, 1.0e13
    n = 12
    rng = np.random.default_rng(47)
    u0 = _h_state(rng, n, tempsc)[:9].copy(order="F")
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, npas=1)
    O = Oracle(oracle_config(p, eos=(_capi.GM_EOS_ADIABATIC, tempsc)))
    O.set_cooling(_capi.GM_COOL_H, None, tempsc, 1.0, tsc)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    dt, _ = timestep_clamp(1, 10, 0.4, O.cfl_dtp(), 0.0, 1e30)
    O.tstep(dt)
    uo = O.get_u()
    O.close()
    with _capi.Subdomain(p) as dev:
        dev.set_eos(_capi.GM_EOS_ADIABATIC, tempsc)
        dev.set_cooling(_capi.GM_COOL_H, None, tempsc, 1.0, tsc)
        dev.upload_u(u0)
        dev.boundary_u()
        dev.tstep(dt)
        ud = dev.download_u()
    for q in (0, 1, 2, 3, 5, 6, 7):
        np.testing.assert_array_equal(ud[q], uo[q], err_msg=f"field {q}")
    for q in (4, 8):
        np.testing.assert_allclose(ud[q], uo[q], rtol=1e-10, atol=0, err_msg=f"field {q}")
    assert not np.array_equal(ud[8], u0[8])


def test_passives_mixed_against_fp64():
    say("case: passives npas=2 mixed vs FP64, 24^3 periodic, 8 steps")
    rng = np.random.default_rng(37)
    u0 = random_state(rng, 24, 24, 24, contrast=0.1, vmax=0.5, bmax=0.5, npas=2)
    res = {}
    for prec in (_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED):
        p = _capi.make_params(24, 24, 24, 1 / 24, 1 / 24, 1 / 24, precision=prec, npas=2)
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


def test_passive_argument_checks():
    say("case: passive-scalar argument checks")
    p0 = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8)
    with _capi.Subdomain(p0) as dev:
        with pytest.raises(_capi.GmError) as ei:
            dev.set_eos(HR, 1.0e4)                  # no neutral density
        assert ei.value.code == _capi.GM_ERR_CONFIG
        with pytest.raises(_capi.GmError) as ei:
            dev.set_cooling(_capi.GM_COOL_H, None, 1.0e4, 1.0, 1.0)
        assert ei.value.code == _capi.GM_ERR_CONFIG
    with pytest.raises(_capi.GmError) as ei:
        _capi.Subdomain(_capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8, npas=3))
    assert ei.value.code == _capi.GM_ERR_CONFIG
    rng = np.random.default_rng(2)
    p3 = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8, npas=3)
    with _capi.Batch(precision=_capi.GM_PRECISION_FP64, max_boxes=2) as B:
        with pytest.raises(_capi.GmError) as ei:
            B.add(p3, random_state(rng, 8, 8, 8, npas=3), max_steps=1)
        assert ei.value.code == _capi.GM_ERR_CONFIG


def test_h_rate_cooling_with_conduction():
    """H-rate cooling is followed by a full calcprim (src/hydro_solver.f90:
    218-220), so the conduction after it starts from the cooled state's
    temperature, as on the device.  One step of two conduction sub-steps:
    fields neither module touches bit for bit, E and the neutral density to
    a relative 1e-10."""
    say("case: H-rate cooling + isotropic conduction, EOS_H_RATE, npas=2, 12^3 periodic, 1 step")
    tempsc, tsc = 1.0e4, 1.0e13
    n = 12
    rng = np.random.default_rng(43)
    u0 = _h_state(rng, n, tempsc)
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, npas=2)
    O = Oracle(oracle_config(p, eos=(HR, tempsc)))
    O.set_cooling(_capi.GM_COOL_H, None, tempsc, 1.0, tsc)
    O.set_conduction(1, False, tempsc=tempsc, tsc=tsc)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    dt = 3.0 * O.dt_cond() / tsc
    O.tstep(dt)
    uo = O.get_u()
    O.close()
    with _capi.Subdomain(p) as dev:
        dev.set_eos(HR, tempsc)
        dev.set_cooling(_capi.GM_COOL_H, None, tempsc, 1.0, tsc)
        dev.set_conduction(1, False, tempsc=tempsc, tsc=tsc)
        dev.upload_u(u0)
        dev.boundary_u()
        dev.tstep(dt)
        ud = dev.download_u()
        assert dev.conduction_substeps() == 2
    for q in (0, 1, 2, 3, 5, 6, 7, 9):
        np.testing.assert_array_equal(ud[q], uo[q], err_msg=f"field {q}")
    for q in (4, 8):
        np.testing.assert_allclose(ud[q], uo[q], rtol=1e-10, atol=0, err_msg=f"field {q}")
    inner = (slice(2, -2),) * 3
    assert np.abs(ud[4][inner] - u0[4][inner]).max() > 0


def test_table_cooling_with_conduction_is_refused():
    say("case: DMC cooling with conduction refused, either order")
    tab = np.load(os.path.join(HERE, "golden", "ref_cooling.npz"))["dmc_table"]
    p = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8)
    with _capi.Subdomain(p) as dev:
        dev.set_cooling(_capi.GM_COOL_DMC, tab, 1e6, 1e-22, 10.0)
        with pytest.raises(_capi.GmError) as ei:
            dev.set_conduction(1, False)
        assert ei.value.code == _capi.GM_ERR_CONFIG
        dev.set_cooling(_capi.GM_COOL_NONE)
        dev.set_conduction(1, False)
        with pytest.raises(_capi.GmError) as ei:
            dev.set_cooling(_capi.GM_COOL_CHI, tab, 1e6, 1e-22, 10.0)
        assert ei.value.code == _capi.GM_ERR_CONFIG
