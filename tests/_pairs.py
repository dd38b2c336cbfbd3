"""Helpers shared by the GPU parity tests: run the same seeded set-up
through the device step (libguacho_amd.so) and the C restatement (oracle)
with the reference's host logic, step by step."""
import numpy as np

from guacho_amd import _capi
from guacho_amd.host import timestep_clamp
from oracle.oracle import Config, Oracle


def oracle_config(params: _capi.GmParams, cfl: float = 0.4, user_src: bool = False,
                  user_bc: bool = False, eos=(1, 1.0)) -> Config:
    return Config(nx=params.nx, ny=params.ny, nz=params.nz, dx=params.dx, dy=params.dy, dz=params.dz,
                  cv=params.cv, riemann_solver=params.riemann_solver, slope_limiter=params.slope_limiter,
                  bc=list(params.bc), mhd=bool(params.mhd), pmhd=bool(params.pmhd),
                  enable_flux_cd=bool(params.enable_flux_cd), cfl=cfl,
                  eight_wave=bool(params.eight_wave), user_src=user_src, user_bc=user_bc, eta=params.eta,
                  eq_of_state=eos[0], tempsc=eos[1], npas=params.neq - 8)


def random_state(rng, nx, ny, nz, cv=1.5, contrast=0.3, vmax=1.5, bmax=1.0, npas=0):
    shape = (nx + 4, ny + 4, nz + 4)
    rho = rng.uniform(1 - contrast, 1 + contrast, shape)
    v = rng.uniform(-vmax, vmax, (3,) + shape)
    B = rng.uniform(-bmax, bmax, (3,) + shape)
    p = rng.uniform(1 - contrast, 1 + contrast, shape)
    u = np.zeros((8 + npas,) + shape, order="F")
    u[0] = rho
    u[1:4] = rho * v
    u[4] = 0.5 * rho * (v ** 2).sum(0) + cv * p + 0.5 * (B ** 2).sum(0)
    u[5:8] = B
    if npas:
        # passive densities below rho (a neutral density under EOS_H_RATE)
        u[8:] = rng.uniform(0.05, 0.95, (npas,) + shape) * rho
    return u


def run_pair(params: _capi.GmParams, u0: np.ndarray, nsteps: int, cfl: float = 0.4, check_each=False,
             user_sources=None, user_bc=None, eos=(1, 1.0), bc_first=False):
    """Advance u0 nsteps on both paths with dt from each path's own CFL
    reduction; asserts the reductions agree bitwise every step.
    user_sources / user_bc: (code object, kernel) of plugins equal to the
    oracle's user source / user boundary hook; both sides see the same
    globals::time.  eos: (eq_of_state, Tempsc) of u2prim on both paths.
    bc_first: register the boundary plugin before the source plugin.
    Returns (u_device, u_oracle, dts)."""
    O = Oracle(oracle_config(params, cfl, user_src=user_sources is not None, user_bc=user_bc is not None, eos=eos))
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    dts = []
    with _capi.Subdomain(params) as dev:
        dev.set_eos(*eos)
        if bc_first and user_bc is not None:
            dev.set_user_bc(*user_bc)
        if user_sources is not None:
            dev.set_user_sources(*user_sources)
        if user_bc is not None and not bc_first:
            dev.set_user_bc(*user_bc)
        dev.upload_u(u0)
        dev.boundary_u()
        t = 0.0
        for it in range(1, nsteps + 1):
            dtp_o = O.cfl_dtp()
            dtp_d = dev.cfl_dtp()
            assert dtp_o == dtp_d, f"step {it}: CFL reduction differs {dtp_o!r} vs {dtp_d!r}"
            dt, _ = timestep_clamp(it, 10, cfl, dtp_o, t, 1e30)
            O.set_time(t)
            dev.set_time(t)
            O.tstep(dt)
            dev.tstep(dt)
            t += dt
            dts.append(dt)
            if check_each:
                np.testing.assert_array_equal(dev.download_u(), O.get_u(), err_msg=f"step {it}")
        ud = dev.download_u()
    uo = O.get_u()
    O.close()
    return ud, uo, dts
