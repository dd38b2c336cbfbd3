"""EOS_SINGLE_SPECIE (u2prim, src/hydro_core.f90:94-99) on the device
against the C restatement.

T = max(1, (p / rho) Tempsc) and p = rho T / Tempsc in every primitive
recovery.  The restatement's u2prim, calcprim sweeps, fluxes and timestep
under this EOS equal the reference's own routines bit for bit
(tests/golden fixtures ref_hlld_minmod_single_specie,
ref_hlle_vanleer_single_specie_closed, Tempsc 1e4 and 2.5e3).  Here the FP64
device step must equal the restatement bit for bit every step, with
Tempsc = 1.2 (the 1 K floor active in part of the box) and 1e4 (no floor:
only the p -> T -> p rounding), and the downloaded temperature must be the
restatement's.  With table cooling or conduction the device's exp, log10
and pow set the tolerance, as for the adiabatic EOS.
"""
import math
import os
import sys

import numpy as np
import pytest

from guacho_amd import _capi
from guacho_amd.host import timestep_clamp
from oracle.oracle import Oracle
from tests._pairs import oracle_config, random_state, run_pair

pytestmark = pytest.mark.gpu

SS = _capi.GM_EOS_SINGLE_SPECIE


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # name, (nx, ny, nz), solver, limiter, bc, flux_cd, tempsc
    ("hlld-minmod-periodic-floor", (16, 12, 10), 4, 1, (3,) * 6, True, 1.2),
    ("hlld-minmod-periodic", (12, 12, 8), 4, 1, (3,) * 6, True, 1.0e4),
    ("hlle-vanleer-closed-floor", (12, 10, 8), 3, 2, (2, 2, 1, 1, 3, 3), False, 1.2),
    ("hllc-superbee-pmhd-outflow", (10, 12, 8), 2, 6, (1,) * 6, False, 2.5e3),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_single_specie_bitwise(case):
    name, (nx, ny, nz), solver, lim, bc, fcd, tempsc = case
    say(f"case: EOS_SINGLE_SPECIE {name} {nx}x{ny}x{nz} Tempsc={tempsc}, 6 steps, checked each step")
    rng = np.random.default_rng(sum(map(ord, name)))
    u0 = random_state(rng, nx, ny, nz)
    pmhd = solver == 2
    p = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, riemann_solver=solver, slope_limiter=lim,
                          bc=bc, enable_flux_cd=fcd, mhd=not pmhd, pmhd=pmhd)
    ud, uo, _ = run_pair(p, u0, 6, check_each=True, eos=(SS, tempsc))
    np.testing.assert_array_equal(ud, uo)


def test_single_specie_changes_the_step():
    """With the floor active the step differs from the adiabatic one."""
    say("case: EOS_SINGLE_SPECIE (Tempsc 1.2) vs adiabatic, 12^3 periodic, 2 steps")
    rng = np.random.default_rng(5)
    u0 = random_state(rng, 12, 12, 12)
    p = _capi.make_params(12, 12, 12, 1 / 12, 1 / 12, 1 / 12)
    a, _, _ = run_pair(p, u0, 2)
    b, _, _ = run_pair(p, u0, 2, eos=(SS, 1.2))
    assert not np.array_equal(a, b)


@pytest.mark.parametrize("eos,tempsc", [(_capi.GM_EOS_ADIABATIC, 3.5), (SS, 1.2), (SS, 1.0e4)])
def test_temperature_download(eos, tempsc):
    say(f"case: primitives and temperature download, eos {eos}, Tempsc {tempsc}, 10x8x6 outflow")
    rng = np.random.default_rng(8)
    u0 = random_state(rng, 10, 8, 6)
    p = _capi.make_params(10, 8, 6, 0.1, 0.125, 1 / 6, bc=(1,) * 6)
    O = Oracle(oracle_config(p, eos=(eos, tempsc)))
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    prim_o, t_o = O.get_primit(), O.get_temp()
    if eos != SS:
        t_o = t_o * tempsc          # the restatement's adiabatic T carries Tempsc = 1
    O.close()
    with _capi.Subdomain(p) as dev:
        dev.set_eos(eos, tempsc)
        dev.upload_u(u0)
        dev.boundary_u()
        prim_d, t_d = dev.download_primit(with_temp=True)
    np.testing.assert_array_equal(prim_d, prim_o)
    np.testing.assert_array_equal(t_d, t_o)
    if eos == SS and tempsc < 2:
        assert np.any(t_d == 1.0)    # the floor is active somewhere


def test_single_specie_with_dmc_cooling():
    """The tabulated cooling this EOS is meant for: T of the EOS feeds the
    cooling rate.  One step: every field but E bit for bit, E to 1e-14."""
    say("case: EOS_SINGLE_SPECIE + DMC cooling, 16^3 outflow, 1 step")
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_cooling.npz")
    tab = np.load(fixture)["dmc_table"]
    tempsc, psc, tsc = 1e6, 1e-22, 10.0
    rng = np.random.default_rng(13)
    n = 16
    u0 = random_state(rng, n, n, n)
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, bc=(1,) * 6)
    O = Oracle(oracle_config(p, eos=(SS, tempsc)))
    O.set_cooling(_capi.GM_COOL_DMC, tab, tempsc, psc, tsc)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    with _capi.Subdomain(p) as dev:
        dev.set_eos(SS, tempsc)
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


def test_single_specie_with_conduction():
    """Conduction reads calcprim's Temp, which under this EOS is u2prim's
    floored T.  One step on a static contact: all fields but E bit for bit,
    E to 1e-13, the sub-step count the restatement's."""
    say("case: EOS_SINGLE_SPECIE + isotropic conduction, 12^3 periodic, one step of 3 conduction timesteps")
    tempsc = 1.0 / 8.3145e7
    n = 12
    rng = np.random.default_rng(17)
    shape = (n + 4,) * 3
    rho = rng.uniform(0.5, 2.0, shape)
    u0 = np.zeros((8,) + shape, order="F")
    u0[0] = rho
    u0[4] = 1.5 * (1.0e6 / tempsc) + 0.5 * 0.25
    u0[5] = 0.5
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, bc=(3,) * 6)
    bsc = math.sqrt(4.0 * math.pi)
    O = Oracle(oracle_config(p, eos=(SS, tempsc)))
    O.set_conduction(1, False, tempsc=tempsc, bsc=bsc)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    dt = 3.0 * O.dt_cond()
    O.tstep(dt)
    uo = O.get_u()
    O.close()
    with _capi.Subdomain(p) as dev:
        dev.set_eos(SS, tempsc)
        dev.set_conduction(1, False, tempsc=tempsc, bsc=bsc)
        dev.upload_u(u0)
        dev.boundary_u()
        dev.tstep(dt)
        ud = dev.download_u()
        assert dev.conduction_substeps() == 2
    for q in (0, 1, 2, 3, 5, 6, 7):
        np.testing.assert_array_equal(ud[q], uo[q], err_msg=f"field {q}")
    np.testing.assert_allclose(ud[4], uo[4], rtol=1e-13, atol=0)


def test_eos_rejects_bad_arguments():
    say("case: EOS argument checks")
    p = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8)
    with _capi.Subdomain(p) as dev:
        for eos, ts in ((3, 1.0), (0, 1.0), (SS, 0.0), (SS, float("inf"))):
            with pytest.raises(_capi.GmError) as ei:
                dev.set_eos(eos, ts)
            assert ei.value.code == _capi.GM_ERR_CONFIG
        dev.set_eos(SS, 2.0)
        dev.set_eos(_capi.GM_EOS_ADIABATIC, 1.0)
