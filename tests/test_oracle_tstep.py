"""The oracle's time step against the reference's own tstep.

The step variants of the reference build (oracle/Makefile, flux-CD off)
link the reference's src/hydro_solver.f90, so its own main-loop iteration
(get_timestep, then tstep: first-order predictor step(dt/2), boundaryII,
calcprim(up), limited corrector step(dt) from u, viscous_copy, boundaryI,
calcprim(u), conduction) runs here.  tests/golden/tstep_<variant>.npz hold
its dt sequence, a CRC of u after every step and the final u and primit;
the restatement must reproduce every bit, ghosts included.  The variants
cover the four solvers, five limiters, the three standard boundary kinds, a
user boundary hook, 8-wave plus user sources inside step, eta != 0
(viscous_copy), thermal conduction after the step and passive scalars under
EOS_H_RATE.  When oracle/_ref is built, longer trajectories from fresh
seeds are also compared live.
"""
import glob
import math
import os
import zlib

import numpy as np
import pytest

from oracle.oracle import Config, Oracle, RefLib, ref_variants, timestep_finish

TSTEP_GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "tstep_*.npz")))


def _cfg(z):
    ci, cr = z["config_int"], z["config_real"]
    dx, dy, dz = z["dxyz"]
    sw = z["switches"]
    return Config(nx=int(ci[0]), ny=int(ci[1]), nz=int(ci[2]), dx=dx, dy=dy, dz=dz, cv=float(cr[0]),
                  eta=float(z["eta"][0]), riemann_solver=int(ci[3]), slope_limiter=int(ci[4]),
                  bc=[int(v) for v in ci[5:11]], mhd=bool(ci[11]), pmhd=bool(ci[12]), enable_flux_cd=False,
                  eight_wave=bool(sw[0]), user_src=bool(sw[1]), user_bc=bool(sw[2]), cfl=float(cr[2]),
                  eq_of_state=int(z["eos_config"][0]), tempsc=float(z["eos_config"][1]), npas=int(z["npas"][0]))


def _oracle(cfg, tc_config, time):
    O = Oracle(cfg, nthreads=1)
    kind, sat, tempsc = tc_config
    if int(kind):
        O.set_conduction(int(kind), bool(sat), tempsc=float(tempsc), bsc=math.sqrt(4.0 * math.pi))
    O.set_time(time)
    return O


def test_tstep_fixtures_present():
    assert len(TSTEP_GOLDEN) >= 9


@pytest.mark.parametrize("path", TSTEP_GOLDEN, ids=lambda p: os.path.basename(p)[6:-4])
def test_tstep_matches_reference(path):
    """get_timestep + tstep, main.f90:94-125, against the reference's own."""
    z = np.load(path)
    cfg = _cfg(z)
    t = float(z["time"][0])
    O = _oracle(cfg, z["tc_config"], t)
    O.set_u(z["u_in"])
    O.boundaryI()
    O.calcprim_u()
    for n, it in enumerate(z["iters"]):
        dt, _ = timestep_finish(int(it), 10, cfg.cfl, O.cfl_dtp(), t, 10.0)
        assert dt == z["dt"][n], f"dt of iteration {it}"
        O.tstep(dt)
        assert zlib.crc32(np.ascontiguousarray(O.get_u()).tobytes()) == int(z["u_crc"][n]), f"u after iteration {it}"
    np.testing.assert_array_equal(O.get_u(), z["u_out"])
    np.testing.assert_array_equal(O.get_primit(), z["primit_out"])


STEP_VARIANTS = [v for v in ref_variants() if v.startswith("step_")]


@pytest.mark.skipif(not STEP_VARIANTS, reason="reference step build (oracle/_ref) not present")
@pytest.mark.parametrize("variant", STEP_VARIANTS)
def test_tstep_live_against_reference(variant):
    """Eight main-loop iterations from a fresh seed, through the ramp and
    past it, against the live reference build: dt and u every step."""
    from tests.golden.make_golden import tstep_state
    R = RefLib(variant)
    dx, dy, dz = 0.09, 0.11, 0.1
    R.init(dx, dy, dz)
    cfg = R.config(dx, dy, dz)
    assert not cfg.enable_flux_cd
    t = 0.25 if R.bc_user else 0.0
    R.set_time(t)
    O = _oracle(cfg, (R.th_cond, int(R.tc_saturation), R.tempsc), t)
    rng = np.random.default_rng(zlib.crc32(variant.encode()) + 17)
    u = tstep_state(R, rng)
    R.set_array(0, u)
    R.boundary(1)
    R.calcprim(0)
    O.set_u(u)
    O.boundaryI()
    O.calcprim_u()
    for it in (1, 2, 3, 9, 10, 11, 12, 13):
        dtr, _ = R.get_timestep(it, 10, t, 10.0)
        dto, _ = timestep_finish(it, 10, cfg.cfl, O.cfl_dtp(), t, 10.0)
        assert dtr == dto, f"dt of iteration {it}"
        R.tstep(dtr)
        O.tstep(dto)
        np.testing.assert_array_equal(O.get_u(), R.get_array(0), err_msg=f"u after iteration {it}")
    np.testing.assert_array_equal(O.get_primit(), R.get_array(2))
