"""FP64 device step against the C restatement of the reference: bit-for-bit
(np.array_equal, which also equates +0 and -0) on the whole conserved
array, ghosts included, after every configuration's steps, and on the CFL
reduction every step.

Runs on an MI355X through libguacho_amd.so (marker `gpu`).
"""
import math
import sys

import numpy as np
import pytest

from guacho_amd import _capi, problems
from tests._pairs import random_state, run_pair

pytestmark = pytest.mark.gpu

P, O_, C_ = 3, 1, 2  # periodic, outflow, closed


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def params(nx, ny, nz, dx, dy, dz, **kw):
    return _capi.make_params(nx, ny, nz, dx, dy, dz, **kw)


def test_orszag_tang_32x32x4_hlld_minmod():
    say("case: OT 32x32x4 HLLD minmod periodic, 20 steps")
    u0, dx, dy, dz = problems.orszag_tang(32, 32, 4)
    ud, uo, dts = run_pair(params(32, 32, 4, dx, dy, dz), u0, 20)
    np.testing.assert_array_equal(ud, uo)
    assert all(d > 0 for d in dts)


@pytest.mark.parametrize("limiter", [-1, 0, 1, 2, 3, 4, 5, 6])
def test_limiters_hlld(limiter):
    say(f"case: random 10x9x8 HLLD limiter {limiter}, 3 steps")
    rng = np.random.default_rng(100 + limiter)
    u0 = random_state(rng, 10, 9, 8)
    p = params(10, 9, 8, 0.1, 0.11, 0.125, slope_limiter=limiter)
    ud, uo, _ = run_pair(p, u0, 3, check_each=True)
    np.testing.assert_array_equal(ud, uo)


@pytest.mark.parametrize("solver,mhd,pmhd,limiter", [
    (3, 1, 0, 5),   # HLLE, Woodward
    (1, 1, 0, 4),   # HLL, UMIST
    (2, 0, 1, 6),   # HLLC passive B, superbee
    (3, 0, 1, 1),   # HLLE passive B
    (4, 1, 0, 2),   # HLLD, van Leer
])
@pytest.mark.parametrize("bc", [(P,) * 6, (C_, C_, O_, O_, P, P), (O_,) * 6, (C_,) * 6],
                         ids=["periodic", "closed-outflow-periodic", "outflow", "closed"])
def test_solvers_and_boundaries(solver, mhd, pmhd, limiter, bc):
    say(f"case: random 9x10x7 solver {solver} mhd {mhd} pmhd {pmhd} limiter {limiter} bc {bc}, 4 steps")
    rng = np.random.default_rng(solver * 31 + limiter * 7 + bc[0] * 3 + bc[2])
    u0 = random_state(rng, 9, 10, 7)
    p = params(9, 10, 7, 0.1, 0.1, 0.1, riemann_solver=solver, mhd=bool(mhd), pmhd=bool(pmhd),
               slope_limiter=limiter, bc=bc)
    ud, uo, _ = run_pair(p, u0, 4)
    np.testing.assert_array_equal(ud, uo)


def test_without_flux_cd():
    say("case: random 8x8x8 HLLD, flux-CD off (upwind update of B), 4 steps")
    rng = np.random.default_rng(5)
    u0 = random_state(rng, 8, 8, 8)
    p = params(8, 8, 8, 0.1, 0.1, 0.1, enable_flux_cd=False)
    ud, uo, _ = run_pair(p, u0, 4)
    np.testing.assert_array_equal(ud, uo)


def test_brio_wu_128():
    say("case: Brio-Wu 128x4x4 gamma=2 HLLD, outflow x, 30 steps")
    u0, dx, dy, dz = problems.brio_wu(128, 4, 4)
    p = params(128, 4, 4, dx, dy, dz, cv=1.0, bc=(O_, O_, P, P, P, P))
    ud, uo, _ = run_pair(p, u0, 30)
    np.testing.assert_array_equal(ud, uo)


def test_blast_24_outflow():
    say("case: blast 24^3 HLLD minmod outflow, 12 steps")
    u0, dx, dy, dz = problems.blast(24, seed=3)
    p = params(24, 24, 24, dx, dy, dz, bc=(O_,) * 6)
    ud, uo, _ = run_pair(p, u0, 12)
    np.testing.assert_array_equal(ud, uo)


def test_turbulence_32():
    say("case: turbulence 32^3 HLLD minmod periodic, 10 steps")
    u0, dx, dy, dz = problems.turbulence(32, seed=11)
    ud, uo, _ = run_pair(params(32, 32, 32, dx, dy, dz), u0, 10)
    np.testing.assert_array_equal(ud, uo)


@pytest.mark.parametrize("solver,mhd,pmhd,lim", [(4, True, False, 1), (1, True, False, 4), (2, False, True, 6)],
                         ids=["hlld-minmod", "hll-umist", "hllc-superbee-pmhd"])
@pytest.mark.parametrize("log2a", [-44, 200])
def test_scaled_state_bitwise(log2a, solver, mhd, pmhd, lim):
    """The step is invariant under rho, rho v, E, p -> a rho, ..., B ->
    sqrt(a) B (velocities and wave speeds unchanged) up to the absolute
    constants of the reference (u2prim's floors 1e-15 and 1e-16, the
    limiters' 1e-7).  At a = 2^-44 (the floors still off) and 2^200 the
    divisions and square roots of the solvers see operands far from unit
    exponents; the device must still equal the restatement bit for bit, and
    the first time step, a function of the scaled initial state alone, must
    not move."""
    say(f"case: turbulence 16^3 scaled by 2^{log2a}, solver {solver} limiter {lim}, 4 steps")
    u0, dx, dy, dz = problems.turbulence(16, seed=19)
    a = 2.0 ** log2a
    us = u0.copy()
    us[0:5] *= a
    us[5:8] *= math.sqrt(a)              # exact: a power of 4
    inner = (slice(None),) + (slice(2, -2),) * 3
    assert us[0][inner[1:]].min() > 1e-15 and us[4][inner[1:]].min() > 1e-15
    p = params(16, 16, 16, dx, dy, dz, riemann_solver=solver, slope_limiter=lim, mhd=mhd, pmhd=pmhd)
    ud, uo, dts = run_pair(p, us, 4, check_each=True)
    np.testing.assert_array_equal(ud, uo)
    _, _, dts0 = run_pair(p, u0, 1)
    assert dts[0] == dts0[0]


@pytest.mark.parametrize("shape,bc", [((2, 2, 2), (P,) * 6), ((2, 3, 2), (O_,) * 6), ((3, 2, 5), (C_,) * 6),
                                      ((33, 17, 3), (P, P, O_, O_, C_, C_)), ((47, 31, 18), (O_, O_, P, P, P, P)),
                                      ((64, 48, 33), (C_, C_, P, P, O_, O_))],
                         ids=["2x2x2-per", "2x3x2-out", "3x2x5-closed", "33x17x3", "47x31x18", "64x48x33"])
def test_minimal_and_ragged_sizes_bitwise(shape, bc):
    """Grids of 2-3 cells (the two-layer ghost mirror at its limit) and
    sizes one past the 32 x 16 plane tile and the 16-plane z-pencil."""
    nx, ny, nz = shape
    say(f"case: HLLD minmod flux-CD {nx}x{ny}x{nz} bc {bc}, 4 steps, checked each step")
    rng = np.random.default_rng(nx * 10007 + ny * 101 + nz)
    u0 = random_state(rng, nx, ny, nz)
    p = params(nx, ny, nz, 1 / nx, 1 / ny, 1 / nz, bc=bc)
    ud, uo, _ = run_pair(p, u0, 4, check_each=True)
    np.testing.assert_array_equal(ud, uo)


def test_nonfinite_state_is_reported():
    say("case: NaN in one interior cell -> GM_ERR_NONFINITE from the CFL reduction")
    u0, dx, dy, dz = problems.orszag_tang(16, 16, 4)
    u0[4, 5, 6, 3] = math.nan
    with _capi.Subdomain(params(16, 16, 4, dx, dy, dz)) as dev:
        dev.upload_u(u0)
        with pytest.raises(_capi.GmError) as ei:
            dev.cfl_dtp()
        assert ei.value.code == _capi.GM_ERR_NONFINITE
        # the error is cleared once reported
        u0[4, 5, 6, 3] = 1.0
        dev.upload_u(u0)
        dev.cfl_dtp()


def test_primitives_and_temperature_download():
    say("case: download_primit vs calcprim of the oracle, OT 16x16x4 after 3 steps")
    from oracle.oracle import Oracle
    from tests._pairs import oracle_config
    u0, dx, dy, dz = problems.orszag_tang(16, 16, 4)
    p = params(16, 16, 4, dx, dy, dz)
    ud, uo, dts = run_pair(p, u0, 3)
    O = Oracle(oracle_config(p))
    O.set_u(uo)
    O.calcprim_u()
    with _capi.Subdomain(p) as dev:
        dev.upload_u(ud)
        prim, temp = dev.download_primit(with_temp=True)
        divb = dev.divb()
    np.testing.assert_array_equal(prim, O.get_primit())
    np.testing.assert_array_equal(temp, O.get_temp())
    np.testing.assert_array_equal(divb, O.divb())


def test_full_size_step_128_bitwise():
    say("case: turbulence 128^3, 2 steps, full-array bitwise vs oracle")
    u0, dx, dy, dz = problems.turbulence(128, seed=2024)
    ud, uo, _ = run_pair(params(128, 128, 128, dx, dy, dz), u0, 2)
    np.testing.assert_array_equal(ud, uo)


def test_turbulence_256_invariants():
    """256^3 at full size: periodic conservation of mass, momentum, energy and
    B to round-off, and flux-CD's preservation of the central-difference
    div B of the dump formula, over 20 steps."""
    say("case: turbulence 256^3 HLLD, 20 steps, conservation and div B")
    from guacho_amd.host import timestep_clamp
    n = 256
    u0, dx, dy, dz = problems.turbulence(n, seed=2024)
    with _capi.Subdomain(params(n, n, n, dx, dy, dz)) as dev:
        dev.upload_u(u0)
        dev.boundary_u()
        div0 = np.abs(dev.divb()).max()
        t = 0.0
        for it in range(1, 21):
            dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
            dev.tstep(dt)
            t += dt
        u1 = dev.download_u()
        div1 = np.abs(dev.divb()).max()
    s0 = u0[:, 2:-2, 2:-2, 2:-2].sum(axis=(1, 2, 3))
    s1 = u1[:, 2:-2, 2:-2, 2:-2].sum(axis=(1, 2, 3))
    scale = np.abs(u0[:, 2:-2, 2:-2, 2:-2]).sum(axis=(1, 2, 3))
    assert np.all(np.abs(s1 - s0) <= 1e-11 * scale), (s0, s1)
    bscale = np.sqrt((u0[5:8, 2:-2, 2:-2, 2:-2] ** 2).sum(0)).max() / dx
    assert div0 <= 1e-10 * bscale
    assert div1 <= 1e-10 * bscale, (div0, div1, bscale)
    assert np.all(np.isfinite(u1))
