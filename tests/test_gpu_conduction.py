"""Thermal conduction after the step (thermal_conduction,
src/thermal_cond.f90:690-768) on the device against the C restatement.

The restatement equals the reference's own thermal_cond.f90 bit for bit
(tests/golden fixtures ref_hlld_minmod_tciso*, ref_hlld_minmod_tcaniso*:
isotropic and field-aligned fluxes, with and without saturation, 1, 2 and
6 super-steps).  The device's pow (Spitzer T^2.5) is within 1 ulp of the
host's, so conducting runs match to a tolerance: after one step every field
the conduction does not touch is bit for bit, E agrees to relative 1e-13,
and the sub-step count is the restatement's.

The state is a static contact -- random density, uniform pressure, v = 0,
B along x -- so the hydro step leaves E unchanged (sM = 0 and vanishing
transverse fields make HLLD's energy flux zero) and only conduction moves
it; dt is then free to span several conduction timesteps.
"""
import math
import sys

import numpy as np
import pytest

from guacho_amd import _capi
from guacho_amd.host import timestep_clamp
from tests._pairs import oracle_config

pytestmark = pytest.mark.gpu

TEMPSC = 1.0 / 8.3145e7     # T = (p / rho) Tempsc: cgs-consistent for mu = 1
BSC = math.sqrt(4.0 * math.pi)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def contact_state(rng, n, cv=1.5, b0=0.5):
    shape = (n + 4,) * 3
    rho = rng.uniform(0.5, 2.0, shape)
    p = 1.0e6 / TEMPSC                      # T ~ 1e6 K / rho
    u = np.zeros((8,) + shape, order="F")
    u[0] = rho
    u[4] = cv * p + 0.5 * b0 * b0
    u[5] = b0
    return u


CASES = [
    # name, kind, saturation, bc, dt factor (in conduction timesteps), expected sub-steps
    ("iso-outflow-1", 1, False, (1,) * 6, 0.5, 1),
    ("iso-periodic-6", 1, False, (3,) * 6, 20.0, 6),
    ("iso-sat-closed-2", 1, True, (2,) * 6, 3.0, 2),
    ("aniso-mixed-2", 2, False, (1, 1, 3, 3, 2, 2), 3.0, 2),
    ("aniso-sat-outflow-6", 2, True, (1,) * 6, 20.0, 6),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_conduction_one_step(case):
    from oracle.oracle import Oracle
    name, kind, sat, bc, fac, nsub = case
    say(f"case: conduction {name}, 12^3, one step of {fac} conduction timesteps")
    n = 12
    rng = np.random.default_rng(sum(map(ord, name)))
    u0 = contact_state(rng, n)
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, bc=bc)
    O = Oracle(oracle_config(p))
    O.set_conduction(kind, sat, tempsc=TEMPSC, bsc=BSC)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    dt = fac * O.dt_cond()          # tsc = 1
    O.tstep(dt)
    uo = O.get_u()
    O.close()
    with _capi.Subdomain(p) as dev:
        dev.set_conduction(kind, sat, tempsc=TEMPSC, bsc=BSC)
        dev.upload_u(u0)
        dev.boundary_u()
        dev.tstep(dt)
        ud = dev.download_u()
        assert dev.conduction_substeps() == nsub
    for q in (0, 1, 2, 3, 5, 6, 7):
        np.testing.assert_array_equal(ud[q], uo[q], err_msg=f"field {q}")
    assert np.all(np.isfinite(ud[4]))
    np.testing.assert_allclose(ud[4], uo[4], rtol=1e-13, atol=0)
    inner = (slice(2, -2),) * 3
    assert np.abs(ud[4][inner] - u0[4][inner]).max() > 1e-6 * np.abs(u0[4]).max()   # conduction acted


def test_conduction_with_hydro_several_steps():
    """A moving, conducting state over 5 CFL steps: all fields to 1e-10."""
    from oracle.oracle import Oracle
    say("case: isotropic conduction + hydro, 12^3 periodic, 5 CFL steps")
    n = 12
    rng = np.random.default_rng(9)
    u0 = contact_state(rng, n)
    u0[1] = u0[0] * 1e6 * rng.uniform(-1, 1, u0[0].shape)       # velocities ~ 0.1 of the sound speed
    u0[4] += 0.5 * u0[1] ** 2 / u0[0]
    p = _capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, bc=(3,) * 6)
    O = Oracle(oracle_config(p))
    O.set_conduction(1, False, tempsc=TEMPSC, bsc=BSC)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    with _capi.Subdomain(p) as dev:
        dev.set_conduction(1, False, tempsc=TEMPSC, bsc=BSC)
        dev.upload_u(u0)
        dev.boundary_u()
        t = 0.0
        for it in range(1, 6):
            dt, _ = timestep_clamp(it, 10, 0.4, O.cfl_dtp(), t, 1e30)
            O.tstep(dt)
            dev.tstep(dt)
            t += dt
        ud = dev.download_u()
    uo = O.get_u()
    O.close()
    scale = np.abs(uo).max(axis=(1, 2, 3), keepdims=True)
    assert np.max(np.abs(ud - uo) / scale) <= 1e-10


def test_conduction_rejects_bad_arguments():
    say("case: conduction argument checks")
    p = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8)
    with _capi.Subdomain(p) as dev:
        with pytest.raises(_capi.GmError) as ei:
            dev.set_conduction(5)
        assert ei.value.code == _capi.GM_ERR_CONFIG
        with pytest.raises(_capi.GmError) as ei:
            dev.set_conduction(1, tempsc=-1.0)
        assert ei.value.code == _capi.GM_ERR_CONFIG
        dev.set_conduction(_capi.GM_TC_OFF)
