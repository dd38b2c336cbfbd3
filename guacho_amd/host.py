"""Host side of the drop-in: the reference's driver around the device step.

Mirrors the reference's user-facing interface for this path:

* ``Parameters``    -- the run-time view of the user's ``parameters`` module
  (OT/parameters.f90:48-177), same names;
* ``Globals``       -- ``globals`` (src/globals.f90:29-69): ``u`` in the layout
  u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax), ``dx``, ``time``,
  ``dt_CFL``, ``currentIteration``;
* ``initmain``      -- grid spacing, clock and allocation (src/init.f90:119-151);
* ``get_timestep``  -- device min reduction, then the CFL ramp and dump
  clamp on the host exactly as src/hydro_core.f90:677-694;
* ``run``           -- the time loop of src/main.f90:94-125, including its
  extra step past ``tmax``.

All arithmetic of the step runs on the GPU through libguacho_amd.so.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi

SOLVER_HLL, SOLVER_HLLC, SOLVER_HLLE, SOLVER_HLLD = 1, 2, 3, 4
BC_OUTFLOW, BC_CLOSED, BC_PERIODIC, BC_OTHER = 1, 2, 3, 4
LIMITER_NO_AVERAGE, LIMITER_NO_LIMIT, LIMITER_MINMOD, LIMITER_VAN_LEER = -1, 0, 1, 2
LIMITER_VAN_ALBADA, LIMITER_UMIST, LIMITER_WOODWARD, LIMITER_SUPERBEE = 3, 4, 5, 6
N_ITER_RAMP = 10  # src/main.f90:97


@dataclass
class Parameters:
    nxtot: int
    nytot: int
    nztot: int
    xmax: float = 1.0
    ymax: float = 1.0
    zmax: float = 1.0
    cv: float = 1.5
    riemann_solver: int = SOLVER_HLLD
    slope_limiter: int = LIMITER_MINMOD
    bc_left: int = BC_PERIODIC
    bc_right: int = BC_PERIODIC
    bc_bottom: int = BC_PERIODIC
    bc_top: int = BC_PERIODIC
    bc_out: int = BC_PERIODIC
    bc_in: int = BC_PERIODIC
    mhd: bool = True
    pmhd: bool = False
    enable_flux_cd: bool = True
    eight_wave: bool = False
    user_source_terms: bool = False
    bc_user: bool = False
    cooling: int = _capi.GM_COOL_NONE                # COOL_NONE / COOL_DMC / COOL_CHI
    th_cond: int = _capi.GM_TC_OFF                   # TC_OFF / TC_ISOTROPIC / TC_ANISOTROPIC
    eq_of_state: int = _capi.GM_EOS_ADIABATIC        # EOS_ADIABATIC / EOS_SINGLE_SPECIE / EOS_H_RATE
    npas: int = 0                                    # passive scalars (PASSIVES builds)
    tc_saturation: bool = False
    Tempsc: float = 1.0                              # parameters module scalings
    Psc: float = 1.0
    tsc: float = 1.0
    rsc: float = 1.0
    rhosc: float = 1.0
    mu: float = 1.0
    vsc2: float = 1.0
    bsc: float = 3.5449077018110318                  # sqrt(4 pi Psc)
    cfl: float = 0.4
    tmax: float = 0.5
    dtprint: float = 0.1
    eta: float = 0.0
    precision: int = _capi.GM_PRECISION_FP64

    @property
    def gamma(self) -> float:
        return (self.cv + 1.0) / self.cv                 # OT/parameters.f90:158

    @property
    def bc(self) -> List[int]:
        return [self.bc_left, self.bc_right, self.bc_bottom, self.bc_top, self.bc_out, self.bc_in]

    @property
    def nx(self) -> int:
        return self.nxtot

    @property
    def ny(self) -> int:
        return self.nytot

    @property
    def nz(self) -> int:
        return self.nztot


@dataclass
class Globals:
    u: np.ndarray
    dx: float
    dy: float
    dz: float
    time: float = 0.0
    dt_CFL: float = 0.0
    currentIteration: int = 1
    tprint: float = 0.0
    itprint: int = 0


def initmain(prm: Parameters) -> Globals:
    """Grid spacing, clock and the u allocation of src/init.f90:119-151 (cold start)."""
    dx = prm.xmax / prm.nxtot
    dy = prm.ymax / prm.nytot
    dz = prm.zmax / prm.nztot
    u = np.zeros((8 + prm.npas, prm.nxtot + 4, prm.nytot + 4, prm.nztot + 4), dtype=np.float64, order="F")
    return Globals(u=u, dx=dx, dy=dy, dz=dz, time=0.0, currentIteration=1, tprint=prm.dtprint, itprint=0)


def device_params(prm: Parameters, g: Globals, device: int = 0) -> _capi.GmParams:
    return _capi.make_params(prm.nxtot, prm.nytot, prm.nztot, g.dx, g.dy, g.dz, cv=prm.cv, gamma=prm.gamma,
                             riemann_solver=prm.riemann_solver, slope_limiter=prm.slope_limiter,
                             bc=prm.bc, mhd=prm.mhd, pmhd=prm.pmhd, enable_flux_cd=prm.enable_flux_cd,
                             eight_wave=prm.eight_wave, precision=prm.precision, device=device,
                             eta=prm.eta, npas=prm.npas)


def timestep_clamp(current_iter: int, n_iter: int, cfl: float, dtp: float, current_time: float,
                   tprint: float):
    """CFL ramp and dump clamp, src/hydro_core.f90:677-694 (MIN allreduce of one rank is the identity)."""
    if current_iter <= n_iter:
        dtp = (cfl * (2.0 ** (-float(n_iter + 1 - current_iter)))) * dtp
    else:
        dtp = cfl * dtp
    dt = dtp
    dump = False
    if (current_time + dt) >= tprint:
        dt = tprint - current_time
        dump = True
    return dt, dump


def get_timestep(dev: _capi.Subdomain, prm: Parameters, g: Globals, n_iter: int = N_ITER_RAMP):
    """get_timestep (src/hydro_core.f90:623-697): device min, host ramp and clamp."""
    dtp = dev.cfl_dtp()
    return timestep_clamp(g.currentIteration, n_iter, prm.cfl, dtp, g.time, g.tprint)


def read_cooling_table(path: str, kind: int):
    """cooltab(2, 41) as read_table_dmc / read_table_chianti build it from
    their .tab files (src/cooling_dmc.f90:57-66): DMC rows (a, b) become
    (10**a, 10**(-b)); CHIANTI rows are taken as they are."""
    import numpy as np
    rows = np.loadtxt(path, converters=lambda s: float(s.replace("d", "e").replace("D", "e")))[:41]
    tab = np.empty((2, 41), order="F")
    if kind == _capi.GM_COOL_DMC:
        tab[0] = [10.0 ** a for a in rows[:, 0]]
        tab[1] = [10.0 ** (-b) for b in rows[:, 1]]
    else:
        tab[0], tab[1] = rows[:, 0], rows[:, 1]
    return tab


def run(prm: Parameters, g: Globals, device: int = 0,
        write_output: Optional[Callable[[int, Globals, _capi.Subdomain], None]] = None,
        max_steps: Optional[int] = None, log: Optional[Callable[[str], None]] = None,
        user_sources: Optional[Tuple[str, str]] = None,
        user_bc: Optional[Tuple[str, str]] = None, cooltab=None) -> Globals:
    """The time loop of src/main.f90:70-125 with the step on the GPU.

    ``g.u`` holds the initial conditions (ghosts included).  boundaryI and
    calcprim at start-up run on the device.  ``write_output(itprint, g, dev)``
    is called at t=0 and at every dump, as write_output is in the reference.
    ``user_sources`` = (code object, kernel) is the device form of the
    problem's get_user_source_terms (needed when user_source_terms is set),
    ``user_bc`` that of impose_user_bc (needed when bc_user is set);
    ``cooltab`` is cooltab(2, 41) for DMC / CHIANTI cooling
    (read_cooling_table).
    Returns ``g`` with ``u`` downloaded at the end.
    """
    if prm.user_source_terms and user_sources is None:
        raise ValueError("user_source_terms needs the device plugin: pass user_sources=(code_object, kernel)")
    if prm.bc_user and user_bc is None:
        raise ValueError("bc_user needs the device plugin: pass user_bc=(code_object, kernel)")
    with _capi.Subdomain(device_params(prm, g, device)) as dev:
        dev.set_eos(prm.eq_of_state, prm.Tempsc)
        if prm.user_source_terms:
            dev.set_user_sources(*user_sources)
        if prm.bc_user:
            dev.set_user_bc(*user_bc)
        if prm.cooling != _capi.GM_COOL_NONE:
            if cooltab is None and prm.cooling != _capi.GM_COOL_H:
                raise ValueError("cooling needs its table: pass cooltab (read_cooling_table)")
            dev.set_cooling(prm.cooling, cooltab, prm.Tempsc, prm.Psc, prm.tsc)
        if prm.th_cond != _capi.GM_TC_OFF:
            dev.set_conduction(prm.th_cond, prm.tc_saturation, tempsc=prm.Tempsc, rsc=prm.rsc, rhosc=prm.rhosc,
                               mu=prm.mu, psc=prm.Psc, tsc=prm.tsc, vsc2=prm.vsc2, bsc=prm.bsc)
        dev.upload_u(g.u)
        dev.set_time(g.time)
        dev.boundary_u()                                     # main.f90:76
        if write_output is not None:                         # main.f90:84-87
            write_output(g.itprint, g, dev)
        g.itprint += 1
        nsteps = 0
        while g.time <= prm.tmax:                            # main.f90:94
            dt, dump = get_timestep(dev, prm, g)             # main.f90:97
            g.dt_CFL = dt
            if log is not None:
                log(f"Iteration {g.currentIteration} | time: {g.time:12.3e} | dt: {dt:12.3e} | tprint: {g.tprint:12.3e}")
            dev.set_time(g.time)                             # globals::time during tstep
            dev.tstep(dt)                                    # main.f90:106
            g.time = g.time + dt                             # main.f90:108
            if dump:                                         # main.f90:110-121
                if write_output is not None:
                    write_output(g.itprint, g, dev)
                g.tprint = g.tprint + prm.dtprint
                g.itprint += 1
            g.currentIteration += 1
            nsteps += 1
            if max_steps is not None and nsteps >= max_steps:
                break
        g.u = dev.download_u()
    return g
