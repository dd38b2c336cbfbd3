"""Invariants of the restatement's step.  The reference's own tstep, step and
viscous_copy pin it bit for bit with flux-CD off (tests/test_oracle_tstep.py);
flux_cd_module.f90 does not compile serially (its non-MPI branch leaves an
`if` open at :120), so get_efield, flux_cd_update and boundaryI_ef are
pinned by tests/test_oracle_fluxcd.py and by the checks below.  Each is
fixed by the mathematics and fails for a plausible slip (a dropped term, a
wrong sign or index, a transposed operand):

* a periodic box conserves sum(rho), sum(m), sum(E) and sum(B) over a step
  (the flux differences and the E differences of flux_cd_update,
  src/flux_cd_module.f90:299-321, telescope);
* flux-CD keeps the central-difference div B of a field that starts with
  div B = 0 at round-off (the property flux-CD is built for);
* swapping the x and y axes (and the x/y vector components) commutes with
  the step to round-off;
* viscous_copy (src/hydro_solver.f90:47-65): the 7-point stencil's closed
  forms -- a unit spike spreads eta to its six face neighbours and keeps
  1 - 6 eta; a linear field is left unchanged; a value held in up's first
  ghost layer reaches the adjacent interior cell as eta times that value
  (up's ghosts are read as they stand, which after the corrector is
  boundaryII's fill of the predictor state).
"""
import numpy as np
import pytest

from guacho_amd import problems
from oracle.oracle import Config, Oracle
from tests._pairs import random_state


def _advance(cfg, u0, nsteps=1, dt=None):
    O = Oracle(cfg)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    for _ in range(nsteps):
        step = dt if dt is not None else 0.3 * O.cfl_dtp()
        O.tstep(step)
    u = O.get_u()
    d = O.divb()
    O.close()
    return u, d


@pytest.mark.parametrize("solver,limiter,fcd", [(4, 1, True), (3, 2, True), (1, 4, False)])
def test_periodic_step_conserves(solver, limiter, fcd):
    n = 8
    rng = np.random.default_rng(31 + solver)
    u0 = random_state(rng, n, n, n)
    cfg = Config(nx=n, ny=n, nz=n, dx=1 / n, dy=1 / n, dz=1 / n, riemann_solver=solver, slope_limiter=limiter,
                 enable_flux_cd=fcd)
    u1, _ = _advance(cfg, u0, 2)
    inner = (slice(None),) + (slice(2, -2),) * 3
    s0 = u0[inner].sum(axis=(1, 2, 3))
    s1 = u1[inner].sum(axis=(1, 2, 3))
    scale = np.abs(u0[inner]).sum(axis=(1, 2, 3))
    assert np.all(np.abs(s1 - s0) <= 1e-13 * scale), (s1 - s0) / scale
    assert np.abs(u1[inner] - u0[inner]).max() > 1e-3   # the state moved


def test_flux_cd_keeps_div_b():
    n = 12
    u0, dx, dy, dz = problems.turbulence(n, seed=5, kmax=3)
    cfg = Config(nx=n, ny=n, nz=n, dx=dx, dy=dy, dz=dz)
    O = Oracle(cfg)
    O.set_u(u0)
    O.boundaryI()
    d0 = np.abs(O.divb()).max()
    O.close()
    u1, d1 = _advance(cfg, u0, 3)
    bscale = np.abs(u0[5:8]).max() / dx
    assert d0 <= 1e-12 * bscale
    assert np.abs(d1).max() <= 1e-12 * bscale
    # without flux-CD the same steps do change div B
    _, d2 = _advance(Config(nx=n, ny=n, nz=n, dx=dx, dy=dy, dz=dz, enable_flux_cd=False), u0, 3)
    assert np.abs(d2).max() > 1e-6 * bscale


def _swap_xy(u):
    v = np.asfortranarray(np.transpose(u, (0, 2, 1, 3)))
    v = v[[0, 2, 1, 3, 4, 6, 5, 7]]
    return np.asfortranarray(v)


@pytest.mark.parametrize("bc", [(3,) * 6, (2, 2, 1, 1, 3, 3)])
def test_axis_swap_commutes_with_step(bc):
    nx, ny, nz = 8, 6, 5
    rng = np.random.default_rng(77)
    u0 = random_state(rng, nx, ny, nz)
    dx, dy, dz = 1 / nx, 1.3 / ny, 1 / nz
    cfg = Config(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, bc=list(bc))
    cfg_s = Config(nx=ny, ny=nx, nz=nz, dx=dy, dy=dx, dz=dz, bc=[bc[2], bc[3], bc[0], bc[1], bc[4], bc[5]])
    O = Oracle(cfg)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    dt = 0.3 * O.cfl_dtp()
    O.close()
    a, _ = _advance(cfg, u0, 1, dt)
    b, _ = _advance(cfg_s, _swap_xy(u0), 1, dt)
    inner = (slice(None),) + (slice(2, -2),) * 3
    a_s = _swap_xy(a)[inner]
    scale = np.abs(a[inner]).max(axis=(1, 2, 3), keepdims=True)
    assert np.max(np.abs(a_s - b[inner]) / scale) <= 1e-13


def _viscous(eta, up, n=6):
    O = Oracle(Config(nx=n, ny=n, nz=n, dx=1 / n, dy=1 / n, dz=1 / n, eta=eta))
    O.set_up(up)
    O.set_u(np.zeros_like(up))
    O.viscous_copy()
    u = O.get_u()
    O.close()
    return u


def test_viscous_copy_spike():
    n, eta = 6, 0.05
    up = np.zeros((8, n + 4, n + 4, n + 4), order="F")
    up[:, 5, 4, 5] = np.arange(1.0, 9.0)            # interior cell (4, 3, 4)
    u = _viscous(eta, up, n)[:, 2:-2, 2:-2, 2:-2]
    c = (3, 2, 3)                                   # its 0-based interior index
    v = np.arange(1.0, 9.0)
    np.testing.assert_allclose(u[(slice(None),) + c], v * (1.0 - 6.0 * eta), rtol=1e-15)
    nb = [(4, 2, 3), (2, 2, 3), (3, 3, 3), (3, 1, 3), (3, 2, 4), (3, 2, 2)]
    for p in nb:
        np.testing.assert_array_equal(u[(slice(None),) + p], eta * v)
    mask = np.ones((n, n, n), bool)
    mask[c] = False
    for p in nb:
        mask[p] = False
    assert not np.any(u[:, mask])


def test_viscous_copy_keeps_linear_field():
    n = 6
    i, j, k = np.meshgrid(*(np.arange(n + 4.0),) * 3, indexing="ij")
    up = np.asfortranarray(np.stack([q + i + 2 * j - 3 * k for q in range(8)]))
    u = _viscous(0.1, up, n)
    np.testing.assert_array_equal(u[:, 2:-2, 2:-2, 2:-2], up[:, 2:-2, 2:-2, 2:-2])


def test_viscous_copy_reads_first_ghost_layer():
    n, eta = 6, 0.125
    up = np.zeros((8, n + 4, n + 4, n + 4), order="F")
    up[:, 1, :, :] = 7.0      # ghost layer i = 0 (first outside the domain)
    up[:, 0, :, :] = 100.0    # ghost layer i = -1: never read
    up[:, :, -2, :] = 3.0     # ghost layer j = ny + 1
    u = _viscous(eta, up, n)[:, 2:-2, 2:-2, 2:-2]
    assert np.all(u[:, 0, :-1, :] == eta * 7.0)
    assert np.all(u[:, 0, -1, :] == eta * 10.0)
    assert np.all(u[:, 1:, -1, :] == eta * 3.0)
    assert not np.any(u[:, 1:, :-1, :])
