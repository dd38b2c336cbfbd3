"""Pins of the oracle's flux-CD path: get_efield, flux_cd_update, boundaryI_ef.

src/flux_cd_module.f90 does not compile serially (its non-MPI branch of
boundaryI_ef leaves an `if` open at :120), so these routines have no
reference build.  The step composition they sit in (tstep's dt/2
predictor, boundaryII, the corrector from u, viscous_copy, boundaryI) is
pinned bit for bit against the reference's own tstep with flux-CD off
(tests/test_oracle_tstep.py).  Here each flux-CD routine is pinned by what
the mathematics fixes, chosen so that a plausible slip fails:

* flux_cd_update's hydro and passive part (flux_cd_module.f90:299-308) is
  the upwind update of step (hydro_solver.f90:105-107): bit for bit equal
  to the flux-CD-off step on the same fluxes;
* get_efield (flux_cd_module.f90:258-265) on a uniform state is E = -v x B
  (Toth 2000, eq. 31: the interface fluxes of B are the cell's);
* flux_cd_update's field part (:311-321) is Faraday's law with central
  differences of E: on a smooth field its B change agrees with the upwind
  B change to O(dx) (first-order fluxes) and the gap halves with dx;
* the whole flux-CD step is second order in space and time on a smooth
  circularly polarised Alfven wave, an exact nonlinear solution, along
  each axis in turn;
* boundaryI_ef's outflow copies (:194-235) keep a state that is uniform
  along an outflow axis uniform, bit for bit; its closed branch (:142-192)
  reads, wherever the update reads it, as the outflow branch.
"""
import math

import numpy as np
import pytest

from oracle.oracle import (BC_CLOSED, BC_OUTFLOW, BC_PERIODIC, LIMITER_MINMOD, LIMITER_NO_LIMIT, SOLVER_HLLD,
                           SOLVER_HLLE, Config, Oracle, timestep_finish)


def _conserved(rho, v, p, B, cv):
    u = np.zeros((8,) + rho.shape)
    u[0] = rho
    u[1:4] = rho * v
    u[4] = 0.5 * rho * (v ** 2).sum(0) + cv * p + 0.5 * (B ** 2).sum(0)
    u[5:8] = B
    return u


def _padded(field_fn, n, axis, cv, nt=4):
    """A state on an (n along `axis`, nt across) grid with two ghost layers;
    field_fn(x) gives (rho, v, p, B) at cell centres x = (i - 1/2)/n."""
    shape = [nt + 4] * 3
    shape[axis] = n + 4
    x1 = (np.arange(n + 4) - 1.5) / n
    rho, v, p, B = field_fn(x1)
    u1 = _conserved(rho, v, p, B, cv)          # (8, n+4)
    bshape = [1, 1, 1]
    bshape[axis] = n + 4
    u = np.broadcast_to(u1.reshape((8,) + tuple(bshape)), (8,) + tuple(shape))
    return np.asfortranarray(u)


def _cfg(n, axis, flux_cd=True, solver=SOLVER_HLLD, bc=None, nt=4, cv=1.5, limiter=LIMITER_MINMOD):
    dims = [nt, nt, nt]
    dims[axis] = n
    d = 1.0 / n
    return Config(nx=dims[0], ny=dims[1], nz=dims[2], dx=d, dy=d, dz=d, cv=cv, riemann_solver=solver,
                  slope_limiter=limiter, bc=bc or [BC_PERIODIC] * 6, enable_flux_cd=flux_cd)


def _stage(cfg, u, choice, dt):
    """boundaryI + calcprim + sweep + step(dt) on u; returns (u, up, e)."""
    O = Oracle(cfg, nthreads=1)
    O.set_u(u)
    O.boundaryI()
    O.calcprim_u()
    O.fluxes(choice)
    O.step(dt)
    out = O.get_u(), O.get_up(), (O.get_e() if cfg.enable_flux_cd else None)
    O.close()
    return out


def _smooth(x):
    """A smooth periodic MHD state along one axis (all components vary)."""
    s, c = np.sin(2 * np.pi * x), np.cos(2 * np.pi * x)
    rho = 1.0 + 0.2 * s
    v = np.stack([0.3 + 0.1 * c, 0.2 * s, -0.15 * c])
    p = 0.8 + 0.1 * c
    B = np.stack([0.7 + 0 * x, 0.5 * c, 0.4 * s])
    return rho, v, p, B


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("choice", [1, 2])
def test_fluxcd_hydro_part_is_the_upwind_update(axis, choice):
    """flux_cd_update(1:5) and its passives (flux_cd_module.f90:299-308) are
    step's upwind update (hydro_solver.f90:105-107) on the same fluxes."""
    n = 12
    u = _padded(_smooth, n, axis, 1.5)
    dt = 0.01
    _, up_cd, _ = _stage(_cfg(n, axis, True), u, choice, dt)
    _, up_uw, _ = _stage(_cfg(n, axis, False), u, choice, dt)
    np.testing.assert_array_equal(up_cd[0:5], up_uw[0:5])
    # the field part differs from the upwind one (it is flux-CD's own)
    assert not np.array_equal(up_cd[5:8], up_uw[5:8])


@pytest.mark.parametrize("solver", [SOLVER_HLLD, SOLVER_HLLE])
def test_efield_uniform_state_is_minus_v_cross_b(solver):
    """E = -v x B on a uniform state (flux_cd_module.f90:258-265), and its
    central differences leave B unchanged."""
    rng = np.random.default_rng(3)
    for trial in range(6):
        rho = rng.uniform(0.5, 2.0)
        v = rng.uniform(-1.0, 1.0, 3)
        p = rng.uniform(0.3, 2.0)
        B = rng.uniform(-1.0, 1.0, 3)
        cfg = _cfg(5, 0, True, solver=solver, nt=5)
        shp = (cfg.nx + 4, cfg.ny + 4, cfg.nz + 4)
        u = _conserved(np.full(shp, rho), v[:, None, None, None] * np.ones(shp), np.full(shp, p),
                       B[:, None, None, None] * np.ones(shp), cfg.cv)
        u0, up, e = _stage(cfg, np.asfortranarray(u), 1 + trial % 2, 0.02)
        want = -np.cross(v, B)
        inner = e[:, 2:-2, 2:-2, 2:-2].reshape(3, -1)
        scale = np.abs(v).max() * np.abs(B).max()
        np.testing.assert_allclose(inner, want[:, None] * np.ones_like(inner), rtol=0, atol=1e-13 * scale)
        np.testing.assert_array_equal(up[5:8, 2:-2, 2:-2, 2:-2], u0[5:8, 2:-2, 2:-2, 2:-2])


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fluxcd_field_update_converges_to_upwind(axis):
    """On a smooth field the flux-CD B change (central differences of the
    cell-centred E, flux_cd_module.f90:311-321) and the upwind B change
    (differences of the interface fluxes) agree to O(dx): the relative gap
    halves as dx halves.  A wrong sign or factor in E or in the update
    leaves an O(1) gap."""
    gaps = []
    for n in (16, 32, 64, 128):
        u = _padded(_smooth, n, axis, 1.5)
        dt = 0.2 / n
        u0, up_cd, _ = _stage(_cfg(n, axis, True), u, 1, dt)
        _, up_uw, _ = _stage(_cfg(n, axis, False), u, 1, dt)
        d_cd = (up_cd - u0)[5:8, 2:-2, 2:-2, 2:-2]
        d_uw = (up_uw - u0)[5:8, 2:-2, 2:-2, 2:-2]
        gaps.append(np.abs(d_cd - d_uw).sum() / np.abs(d_uw).sum())
    gaps = np.array(gaps)
    assert gaps[-1] < 0.05, gaps
    assert np.all(gaps[:-1] / gaps[1:] > 1.7), gaps


def _alfven(axis, b0=1.0, amp=0.1, rho0=1.0, p0=0.1):
    """Circularly polarised Alfven wave along `axis`: |B_perp| and the total
    pressure are uniform, v_perp = -B_perp / sqrt(rho) (Walen relation for
    the wave travelling along +B0); it moves at v_A = b0 / sqrt(rho0)
    without changing shape."""
    t1, t2 = [a for a in range(3) if a != axis]

    def fn(x):
        rho = rho0 + 0 * x
        B = np.zeros((3,) + x.shape)
        B[axis] = b0
        B[t1] = amp * np.sin(2 * np.pi * x)
        B[t2] = amp * np.cos(2 * np.pi * x)
        v = -B / math.sqrt(rho0)
        v[axis] = 0.0
        return rho, v, p0 + 0 * x, B
    return fn


def _run(cfg, u, t_end):
    O = Oracle(cfg, nthreads=1)
    O.set_u(u)
    O.boundaryI()
    O.calcprim_u()
    t, it = 0.0, 11          # past the CFL ramp
    while t < t_end:
        dt, _ = timestep_finish(it, 10, 0.4, O.cfl_dtp(), t, t_end)
        O.tstep(dt)
        t += dt
    out = O.get_u()
    O.close()
    return out


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("flux_cd", [True, False])
def test_alfven_wave_second_order(axis, flux_cd):
    """One period (t = 1 at v_A = 1) of the Alfven wave at 16, 32, 64 cells
    with the unlimited reconstruction (minmod clips the wave's extrema and
    lowers the order there): the L1 error of the transverse field falls ~4x
    per halving (second order; measured 2.07, 2.02 with flux-CD).  A
    predictor over dt instead of dt/2 makes it first order; a corrector
    that starts from U* moves the wave 1.5x too far."""
    errs = []
    for n in (16, 32, 64):
        fn = _alfven(axis)
        u = _padded(fn, n, axis, 1.5)
        out = _run(_cfg(n, axis, flux_cd, limiter=LIMITER_NO_LIMIT), u, 1.0)
        t1, t2 = [a for a in range(3) if a != axis]
        sl = (slice(None),) + tuple(slice(2, -2) for _ in range(3))
        err = (np.abs(out[5 + t1] - u[5 + t1])[sl[1:]].mean() + np.abs(out[5 + t2] - u[5 + t2])[sl[1:]].mean())
        errs.append(err)
        # the normal field stays exactly uniform
        assert np.abs(out[5 + axis][sl[1:]] - 1.0).max() < 1e-12
    errs = np.array(errs)
    orders = np.log2(errs[:-1] / errs[1:])
    assert errs[-1] < 1e-3, errs
    assert np.all(orders > 1.85), (errs, orders)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_outflow_axis_uniformity_preserved(axis):
    """A state uniform along an axis with outflow walls stays uniform along
    it, bit for bit, through two full flux-CD steps: boundaryI_ef's outflow
    branch (flux_cd_module.f90:194-235) copies the cell next to the wall, so
    a flipped or shifted E ghost would bend the field at the wall.  (Closed
    walls do not keep such a state uniform even without flux-CD: magnetic
    tension drives the wall-normal velocity, which the wall reflects.)"""
    other = [a for a in range(3) if a != axis]
    n = 8
    bc = [BC_PERIODIC] * 6
    bc[2 * axis] = bc[2 * axis + 1] = BC_OUTFLOW
    cfg = _cfg(n, other[0], True, bc=bc, nt=n)
    shp = (n + 4,) * 3

    def fld(a, b):
        # smooth in the two other axes, constant along `axis`
        ii = np.arange(n + 4)[:, None]
        jj = np.arange(n + 4)[None, :]
        f = (a * np.sin(2 * np.pi * (ii - 1.5) / n + b) + 0.5 * a * np.cos(2 * np.pi * (jj - 1.5) / n - b))
        return np.broadcast_to(np.expand_dims(f, axis), shp).copy()
    rho = 1.0 + fld(0.2, 0.3)
    p = 1.0 + fld(0.15, 1.1)
    v = np.stack([0.2 + fld(0.3, k) for k in (0.2, 0.9, 1.7)])
    B = np.stack([0.6 + fld(0.25, k) for k in (2.1, 0.4, 1.3)])
    u = np.asfortranarray(_conserved(rho, v, p, B, cfg.cv))
    O = Oracle(cfg, nthreads=1)
    O.set_u(u)
    O.boundaryI()
    O.calcprim_u()
    for it in (11, 12):
        dt, _ = timestep_finish(it, 10, 0.4, O.cfl_dtp(), 0.0, 10.0)
        O.tstep(dt)
    out = np.moveaxis(O.get_u()[:, 2:-2, 2:-2, 2:-2], axis + 1, 1)
    np.testing.assert_array_equal(out, np.broadcast_to(out[:, :1], out.shape))
    # the field did evolve
    start = np.moveaxis(u[:, 2:-2, 2:-2, 2:-2], axis + 1, 1)
    assert np.abs(out[5:8] - start[5:8]).max() > 1e-3
    O.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_closed_wall_efield_ghosts_read_as_outflow(axis):
    """boundaryI_ef's closed branch (flux_cd_module.f90:142-192) differs from
    the outflow branch only in the one component it negates (e1 on x walls,
    e2 on y walls, nothing on z walls), which flux_cd_update never reads at
    that wall's neighbours (:311-321): with the same U ghosts, a closed-wall
    and an outflow-wall flux-CD step give the same bits."""
    n = 7
    rng = np.random.default_rng(40 + axis)
    shp = (n + 4,) * 3
    rho = rng.uniform(0.8, 1.2, shp)
    p = rng.uniform(0.8, 1.2, shp)
    v = rng.uniform(-0.5, 0.5, (3,) + shp)
    B = rng.uniform(-0.7, 0.7, (3,) + shp)
    u = np.asfortranarray(_conserved(rho, v, p, B, 1.5))
    outs = []
    for kind in (BC_CLOSED, BC_OUTFLOW):
        bc = [BC_PERIODIC] * 6
        bc[2 * axis] = bc[2 * axis + 1] = kind
        cfg = _cfg(n, 0, True, bc=bc, nt=n)
        O = Oracle(cfg, nthreads=1)
        if kind == BC_CLOSED:
            O.set_u(u)
            O.boundaryI()
            u_closed = O.get_u()
        else:
            O.set_u(u_closed)          # the closed wall's U ghosts
        O.calcprim_u()
        O.fluxes(2)
        O.step(0.01)
        outs.append((O.get_up()[:, 2:-2, 2:-2, 2:-2], O.get_e()))
        O.close()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    # the E ghosts themselves differ in the negated component on x / y walls
    e_c, e_o = outs[0][1], outs[1][1]
    if axis < 2:
        inner = [slice(2, -2)] * 3
        inner[axis] = 1                     # ghost layer 0
        sel = (axis,) + tuple(inner)
        np.testing.assert_array_equal(e_c[sel], -e_o[sel])
