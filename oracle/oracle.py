"""ctypes bindings to the oracle libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package guacho_amd/.

* ``Oracle``   -- oracle/liboracle.so, the C restatement of the reference's
  step (oracle/guacho_oracle.c).
* ``RefLib``   -- oracle/_ref/<variant>/libguacho_ref.so, the reference's own
  Fortran routines compiled from its sources by oracle/Makefile (present
  only where the reference tree was available at build time).

Arrays use the reference layout u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
(src/globals.f90:33-38): numpy shape (8, nx+4, ny+4, nz+4), order='F'.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NEQ = 8

SOLVER_HLL, SOLVER_HLLC, SOLVER_HLLE, SOLVER_HLLD = 1, 2, 3, 4
BC_OUTFLOW, BC_CLOSED, BC_PERIODIC, BC_OTHER = 1, 2, 3, 4
LIMITER_NO_AVERAGE, LIMITER_NO_LIMIT, LIMITER_MINMOD = -1, 0, 1
LIMITER_VAN_LEER, LIMITER_VAN_ALBADA, LIMITER_UMIST = 2, 3, 4
LIMITER_WOODWARD, LIMITER_SUPERBEE = 5, 6

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_dpF = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("cv", C.c_double), ("gamma", C.c_double), ("eta", C.c_double),
        ("solver", C.c_int), ("limiter", C.c_int), ("bc", C.c_int * 6),
        ("mhd", C.c_int), ("pmhd", C.c_int), ("flux_cd", C.c_int),
        ("eight_wave", C.c_int), ("user_src", C.c_int), ("user_bc", C.c_int), ("nthreads", C.c_int),
        ("eos", C.c_int), ("tempsc", C.c_double), ("npas", C.c_int),
    ]


@dataclass
class Config:
    """Run configuration in the reference's parameter names (OT/parameters.f90)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    cv: float = 1.5
    eta: float = 0.0
    riemann_solver: int = SOLVER_HLLD
    slope_limiter: int = LIMITER_MINMOD
    bc: Sequence[int] = field(default_factory=lambda: [BC_PERIODIC] * 6)
    mhd: bool = True
    pmhd: bool = False
    enable_flux_cd: bool = True
    eight_wave: bool = False
    user_src: bool = False          # the oracle's user gravity (oracle/ref/user_mod.f90)
    user_bc: bool = False           # the oracle's impose_user_bc (oracle/ref/user_mod.f90)
    cfl: float = 0.4
    eq_of_state: int = 1            # 2: EOS_SINGLE_SPECIE, 3: EOS_H_RATE
    tempsc: float = 1.0             # Tempsc of the EOS_SINGLE_SPECIE / EOS_H_RATE temperature
    npas: int = 0                   # passive scalars after the 8 MHD variables

    @property
    def neq(self) -> int:
        return NEQ + self.npas

    @property
    def gamma(self) -> float:
        return (self.cv + 1.0) / self.cv

    def to_orc(self, nthreads: int = 0) -> OrcParams:
        p = OrcParams()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        p.dx, p.dy, p.dz = self.dx, self.dy, self.dz
        p.cv, p.gamma, p.eta = self.cv, self.gamma, self.eta
        p.solver, p.limiter = self.riemann_solver, self.slope_limiter
        for b in range(6):
            p.bc[b] = int(self.bc[b])
        p.mhd, p.pmhd, p.flux_cd = int(self.mhd), int(self.pmhd), int(self.enable_flux_cd)
        p.eight_wave = int(self.eight_wave)
        p.user_src = int(self.user_src)
        p.user_bc = int(self.user_bc)
        p.nthreads = nthreads
        p.eos = int(self.eq_of_state)
        p.tempsc = float(self.tempsc)
        p.npas = int(self.npas)
        return p

    def shape(self):
        return (self.neq, self.nx + 4, self.ny + 4, self.nz + 4)


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load_oracle() -> C.CDLL:
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        build_oracle()
    L = C.CDLL(path)
    P = C.POINTER(OrcParams)
    L.orc_create.argtypes = [P, C.POINTER(C.c_void_p)]
    L.orc_destroy.argtypes = [C.c_void_p]
    for name in ("orc_set_u", "orc_get_u", "orc_set_up", "orc_get_up",
                 "orc_set_primit", "orc_get_primit", "orc_get_temp", "orc_get_e", "orc_divb"):
        getattr(L, name).argtypes = [C.c_void_p, _dpF]
    L.orc_get_flux.argtypes = [C.c_void_p, C.c_int, _dpF]
    for name in ("orc_boundaryI", "orc_boundaryII", "orc_calcprim_u", "orc_calcprim_up",
                 "orc_get_efield", "orc_viscous_copy"):
        getattr(L, name).argtypes = [C.c_void_p]
    L.orc_fluxes.argtypes = [C.c_void_p, C.c_int]
    L.orc_step.argtypes = [C.c_void_p, C.c_double]
    L.orc_tstep.argtypes = [C.c_void_p, C.c_double]
    L.orc_cfl_dtp.argtypes = [C.c_void_p]
    L.orc_cfl_dtp.restype = C.c_double
    L.orc_source.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.orc_set_time.argtypes = [C.c_void_p, C.c_double]
    L.orc_set_cooling.argtypes = [C.c_void_p, C.c_int, _dpF, C.c_double, C.c_double, C.c_double]
    L.orc_cooling.argtypes = [C.c_void_p, C.c_double]
    L.orc_set_conduction.argtypes = [C.c_void_p, C.c_int, C.c_int] + [C.c_double] * 8
    L.orc_thermal.argtypes = [C.c_void_p, C.c_double]
    L.orc_thermal.restype = C.c_int
    L.orc_dt_cond.argtypes = [C.c_void_p]
    L.orc_dt_cond.restype = C.c_double
    L.orc_timestep_finish.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.orc_u2prim.argtypes = [P, _dp, _dp, C.POINTER(C.c_double)]
    L.orc_prim2u.argtypes = [P, _dp, _dp]
    L.orc_prim2f.argtypes = [P, _dp, _dp]
    L.orc_swapy.argtypes = [_dp]
    L.orc_swapz.argtypes = [_dp]
    L.orc_csound.argtypes = [P, C.c_double, C.c_double]
    L.orc_csound.restype = C.c_double
    L.orc_cfast.argtypes = [P] + [C.c_double] * 5 + [C.POINTER(C.c_double)] * 3
    L.orc_cfastX.argtypes = [P, _dp]
    L.orc_cfastX.restype = C.c_double
    L.orc_limiter.argtypes = [P, _dp, _dp, _dp, _dp]
    for name in ("orc_riemann", "orc_hlld", "orc_hlle", "orc_hll", "orc_hllc"):
        getattr(L, name).argtypes = [P, _dp, _dp, _dp]
    return L


_ORC = None


def orc_lib() -> C.CDLL:
    global _ORC
    if _ORC is None:
        _ORC = _load_oracle()
    return _ORC


def timestep_finish(current_iter, n_iter, cfl, dtp, time, tprint):
    """get_timestep's ramp and dump clamp (hydro_core.f90:677-694)."""
    dt = C.c_double()
    dump = C.c_int(0)
    orc_lib().orc_timestep_finish(current_iter, n_iter, cfl, dtp, time, tprint,
                                  C.byref(dt), C.byref(dump))
    return dt.value, bool(dump.value)


class Oracle:
    """Whole-domain restatement of the reference's state and step."""

    def __init__(self, cfg: Config, nthreads: int = 0):
        self.cfg = cfg
        self.L = orc_lib()
        self.p = cfg.to_orc(nthreads)
        h = C.c_void_p()
        rc = self.L.orc_create(C.byref(self.p), C.byref(h))
        if rc != 0:
            raise ValueError(f"orc_create failed ({rc}) for {cfg}")
        self.h = h

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _new(self, nvar=None):
        sh = self.cfg.shape()
        return np.zeros((nvar or sh[0],) + sh[1:], dtype=np.float64, order="F")

    def set_u(self, u):
        self.L.orc_set_u(self.h, np.asfortranarray(u, dtype=np.float64))

    def get_u(self):
        a = self._new(); self.L.orc_get_u(self.h, a); return a

    def set_up(self, u):
        self.L.orc_set_up(self.h, np.asfortranarray(u, dtype=np.float64))

    def get_up(self):
        a = self._new(); self.L.orc_get_up(self.h, a); return a

    def set_primit(self, u):
        self.L.orc_set_primit(self.h, np.asfortranarray(u, dtype=np.float64))

    def get_primit(self):
        a = self._new(); self.L.orc_get_primit(self.h, a); return a

    def get_temp(self):
        a = self._new(1); self.L.orc_get_temp(self.h, a); return a[0]

    def get_flux(self, axis):
        a = self._new(); self.L.orc_get_flux(self.h, axis, a); return a

    def get_e(self):
        a = self._new(3); self.L.orc_get_e(self.h, a); return a

    def divb(self):
        c = self.cfg
        a = np.zeros((c.nx, c.ny, c.nz), order="F")
        self.L.orc_divb(self.h, a)
        return a

    def boundaryI(self): self.L.orc_boundaryI(self.h)
    def boundaryII(self): self.L.orc_boundaryII(self.h)
    def calcprim_u(self): self.L.orc_calcprim_u(self.h)
    def calcprim_up(self): self.L.orc_calcprim_up(self.h)
    def get_efield(self): self.L.orc_get_efield(self.h)
    def viscous_copy(self): self.L.orc_viscous_copy(self.h)

    def fluxes(self, choice):
        if self.L.orc_fluxes(self.h, choice) != 0:
            raise FloatingPointError("Riemann solver fall-through (non-finite state)")

    def step(self, dt): self.L.orc_step(self.h, dt)

    def tstep(self, dt):
        if self.L.orc_tstep(self.h, dt) != 0:
            raise FloatingPointError("Riemann solver fall-through (non-finite state)")

    def cfl_dtp(self) -> float:
        return self.L.orc_cfl_dtp(self.h)

    def set_cooling(self, kind: int, cooltab, tempsc: float = 1.0, psc: float = 1.0, tsc: float = 1.0) -> None:
        """Table cooling (3: DMC, 4: CHIANTI) after each step; cooltab is (2, 41)."""
        tab = np.asfortranarray(cooltab if cooltab is not None else np.zeros((2, 41)), dtype=np.float64)
        self.L.orc_set_cooling(self.h, int(kind), tab, tempsc, psc, tsc)

    def cooling(self, dt: float) -> None:
        self.L.orc_cooling(self.h, float(dt))

    def set_conduction(self, kind: int, saturation: bool = False, tempsc: float = 1.0, rsc: float = 1.0,
                       rhosc: float = 1.0, mu: float = 1.0, psc: float = 1.0, tsc: float = 1.0,
                       vsc2: float = 1.0, bsc: float = 1.0) -> None:
        """Thermal conduction after each step (1 isotropic, 2 anisotropic, 0 off)."""
        self.L.orc_set_conduction(self.h, int(kind), int(saturation), tempsc, rsc, rhosc, mu, psc, tsc, vsc2, bsc)

    def thermal(self, dt: float) -> int:
        """thermal_conduction with dt_CFL = dt; returns the sub-step count."""
        return self.L.orc_thermal(self.h, float(dt))

    def dt_cond(self) -> float:
        return self.L.orc_dt_cond(self.h)

    def set_time(self, time: float) -> None:
        """globals::time, as the user hook sees it."""
        self.L.orc_set_time(self.h, float(time))

    def source(self, i: int, j: int, k: int, prim) -> np.ndarray:
        """source(i,j,k,prim,s) with the 8-wave term, from the state's primit."""
        sv = np.zeros(self.cfg.neq)
        self.L.orc_source(self.h, i, j, k, np.ascontiguousarray(prim, dtype=np.float64), sv)
        return sv


# ---------------------------------------------------------------------------
# per-function helpers (8-vectors), used by the vector parity tests
# ---------------------------------------------------------------------------
class OracleFunctions:
    def __init__(self, cfg: Config):
        self.L = orc_lib()
        self.p = cfg.to_orc()
        self.neq = cfg.neq

    def u2prim(self, uu):
        prim = np.zeros(self.neq); t = C.c_double()
        self.L.orc_u2prim(C.byref(self.p), np.ascontiguousarray(uu, dtype=np.float64), prim, C.byref(t))
        return prim, t.value

    def prim2u(self, prim):
        out = np.zeros(self.neq); self.L.orc_prim2u(C.byref(self.p), np.ascontiguousarray(prim), out); return out

    def prim2f(self, prim):
        out = np.zeros(self.neq); self.L.orc_prim2f(C.byref(self.p), np.ascontiguousarray(prim), out); return out

    def cfast(self, p, d, bx, by, bz):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.L.orc_cfast(C.byref(self.p), p, d, bx, by, bz, C.byref(a), C.byref(b), C.byref(c))
        return np.array([a.value, b.value, c.value])

    def cfastx(self, prim):
        return self.L.orc_cfastX(C.byref(self.p), np.ascontiguousarray(prim))

    def csound(self, p, d):
        return self.L.orc_csound(C.byref(self.p), p, d)

    def limiter(self, pll, pl, pr, prr):
        pl = np.array(pl, dtype=np.float64); pr = np.array(pr, dtype=np.float64)
        self.L.orc_limiter(C.byref(self.p), np.ascontiguousarray(pll), pl, pr, np.ascontiguousarray(prr))
        return pl, pr

    def cooling_table(self, kind):
        """cooltab(2, 41) as the reference reads it (3: DMC, 4: CHIANTI)."""
        t = np.zeros((2, 41), order="F"); self.L.ref_cooling_table(kind, t); return t

    def cooling(self, kind, dt):
        self.L.ref_cooling(kind, dt)

    def thermal(self, dt):
        """thermal_conduction with dt_CFL = dt on the current globals."""
        self.L.ref_thermal(float(dt))

    def dt_cond(self):
        d = C.c_double(); self.L.ref_dt_cond(C.byref(d)); return d.value

    def source(self, i, j, k, prim):
        """source(i,j,k,prim,s) of this variant (sources.f90:188-220), from the
        primit array last set."""
        sv = np.zeros(self.neq)
        self.L.ref_source(i, j, k, np.ascontiguousarray(prim, dtype=np.float64), sv)
        return sv

    def riemann(self, pl, pr):
        ff = np.zeros(self.neq)
        rc = self.L.orc_riemann(C.byref(self.p), np.ascontiguousarray(pl), np.ascontiguousarray(pr), ff)
        return ff, rc


# ---------------------------------------------------------------------------
# the reference's own routines (oracle/_ref)
# ---------------------------------------------------------------------------
REF_DIR = os.path.join(HERE, "_ref")


def ref_variants():
    if not os.path.isdir(REF_DIR):
        return []
    return sorted(d for d in os.listdir(REF_DIR)
                  if os.path.exists(os.path.join(REF_DIR, d, "libguacho_ref.so")))


class RefLib:
    """One compiled configuration of the reference (see oracle/Makefile VARIANTS)."""

    def __init__(self, variant: str):
        path = os.path.join(REF_DIR, variant, "libguacho_ref.so")
        # RTLD_LOCAL + a private copy per variant: every variant exports the
        # same symbol names and owns its own module state.
        self.L = C.CDLL(path, mode=os.RTLD_LOCAL)
        L = self.L
        cfg = (C.c_int * 14)()
        r = (C.c_double * 3)()
        L.ref_config(cfg, r)
        c = list(cfg)
        self.nx, self.ny, self.nz, self.neq = c[0:4]
        self.npas = self.neq - NEQ
        self.riemann_solver, self.slope_limiter = c[4], c[5]
        self.bc = c[6:12]
        self.mhd, self.pmhd = bool(c[12]), bool(c[13])
        self.cv, self.gamma, self.cfl = list(r)
        L.ref_init.argtypes = [C.c_double] * 3
        for name in ("ref_u2prim",):
            getattr(L, name).argtypes = [_dp, _dp, C.POINTER(C.c_double)]
        L.ref_prim2u.argtypes = [_dp, _dp]
        L.ref_prim2f.argtypes = [_dp, _dp]
        L.ref_cfast.argtypes = [C.c_double] * 5 + [_dp]
        L.ref_cfastx.argtypes = [_dp, C.POINTER(C.c_double)]
        L.ref_csound.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.ref_limiter.argtypes = [_dp, _dp, _dp, _dp]
        L.ref_swap.argtypes = [C.c_int, _dp]
        L.ref_riemann.argtypes = [_dp, _dp, _dp]
        L.ref_set_array.argtypes = [C.c_int, _dpF]
        L.ref_get_array.argtypes = [C.c_int, _dpF]
        L.ref_get_temp.argtypes = [_dpF]
        L.ref_boundary.argtypes = [C.c_int]
        L.ref_calcprim.argtypes = [C.c_int]
        L.ref_fluxes.argtypes = [C.c_int]
        L.ref_get_timestep.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.ref_source.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.ref_set_time.argtypes = [C.c_double]
        L.ref_cooling_table.argtypes = [C.c_int, _dpF]
        L.ref_cooling.argtypes = [C.c_int, C.c_double]
        L.ref_thermal.argtypes = [C.c_double]
        L.ref_dt_cond.argtypes = [C.POINTER(C.c_double)]
        tk, ts, tt = C.c_int(), C.c_int(), C.c_double()
        L.ref_conduction_config(C.byref(tk), C.byref(ts), C.byref(tt))
        self.th_cond, self.tc_saturation, self.tempsc = tk.value, bool(ts.value), tt.value
        sw = (C.c_int * 3)()
        L.ref_switches(sw)
        self.eight_wave, self.user_source_terms, self.bc_user = bool(sw[0]), bool(sw[1]), bool(sw[2])
        ev = C.c_int()
        L.ref_eos(C.byref(ev))
        self.eq_of_state = ev.value
        # step variants (oracle/Makefile, flux_cd = 0) also carry the
        # reference's own tstep / step / viscous_copy
        self.has_tstep = hasattr(L, "ref_tstep")
        self.enable_flux_cd, self.eta = True, 0.0
        if self.has_tstep:
            fcd, etav = C.c_int(), C.c_double()
            L.ref_step_config(C.byref(fcd), C.byref(etav))
            self.enable_flux_cd, self.eta = bool(fcd.value), etav.value
            L.ref_tstep.argtypes = [C.c_double]
            L.ref_step.argtypes = [C.c_double]

    def config(self, dx=None, dy=None, dz=None) -> Config:
        dx = dx if dx is not None else 1.0 / self.nx
        dy = dy if dy is not None else 1.0 / self.ny
        dz = dz if dz is not None else 1.0 / self.nz
        return Config(nx=self.nx, ny=self.ny, nz=self.nz, dx=dx, dy=dy, dz=dz, cv=self.cv,
                      riemann_solver=self.riemann_solver, slope_limiter=self.slope_limiter, bc=list(self.bc),
                      mhd=self.mhd, pmhd=self.pmhd, enable_flux_cd=self.enable_flux_cd, eta=self.eta,
                      cfl=self.cfl,
                      eight_wave=self.eight_wave, user_src=self.user_source_terms, user_bc=self.bc_user,
                      eq_of_state=self.eq_of_state, tempsc=self.tempsc, npas=self.npas)

    def init(self, dx, dy, dz):
        self.L.ref_init(dx, dy, dz)

    def shape(self):
        return (self.neq, self.nx + 4, self.ny + 4, self.nz + 4)

    def set_array(self, which, a):
        self.L.ref_set_array(which, np.asfortranarray(a, dtype=np.float64))

    def get_array(self, which):
        a = np.zeros(self.shape(), order="F"); self.L.ref_get_array(which, a); return a

    def get_temp(self):
        a = np.zeros(self.shape()[1:], order="F"); self.L.ref_get_temp(a); return a

    def set_time(self, t): self.L.ref_set_time(float(t))

    def tstep(self, dt):
        """tstep() of hydro_solver.f90:134-229 with dt_CFL = dt (step variants)."""
        self.L.ref_tstep(float(dt))

    def step(self, dt):
        """step(dt) of hydro_solver.f90:77-127 on the current globals (step variants)."""
        self.L.ref_step(float(dt))

    def viscous_copy(self):
        self.L.ref_viscous_copy()

    def boundary(self, order): self.L.ref_boundary(order)
    def calcprim(self, which): self.L.ref_calcprim(which)
    def fluxes(self, choice): self.L.ref_fluxes(choice)

    def get_timestep(self, it, n_iter, t, tprint):
        dt = C.c_double(); d = C.c_int()
        self.L.ref_get_timestep(it, n_iter, t, tprint, C.byref(dt), C.byref(d))
        return dt.value, bool(d.value)

    def u2prim(self, uu):
        prim = np.zeros(self.neq); t = C.c_double()
        self.L.ref_u2prim(np.ascontiguousarray(uu, dtype=np.float64), prim, C.byref(t))
        return prim, t.value

    def prim2u(self, prim):
        out = np.zeros(self.neq); self.L.ref_prim2u(np.ascontiguousarray(prim), out); return out

    def prim2f(self, prim):
        out = np.zeros(self.neq); self.L.ref_prim2f(np.ascontiguousarray(prim), out); return out

    def cfast(self, p, d, bx, by, bz):
        out = np.zeros(3); self.L.ref_cfast(p, d, bx, by, bz, out); return out

    def cfastx(self, prim):
        c = C.c_double(); self.L.ref_cfastx(np.ascontiguousarray(prim), C.byref(c)); return c.value

    def csound(self, p, d):
        c = C.c_double(); self.L.ref_csound(p, d, C.byref(c)); return c.value

    def limiter(self, pll, pl, pr, prr):
        pl = np.array(pl, dtype=np.float64); pr = np.array(pr, dtype=np.float64)
        self.L.ref_limiter(np.ascontiguousarray(pll), pl, pr, np.ascontiguousarray(prr))
        return pl, pr

    def swap(self, axis, v):
        v = np.array(v, dtype=np.float64); self.L.ref_swap(axis, v); return v

    def cooling_table(self, kind):
        """cooltab(2, 41) as the reference reads it (3: DMC, 4: CHIANTI)."""
        t = np.zeros((2, 41), order="F"); self.L.ref_cooling_table(kind, t); return t

    def cooling(self, kind, dt):
        self.L.ref_cooling(kind, dt)

    def thermal(self, dt):
        """thermal_conduction with dt_CFL = dt on the current globals."""
        self.L.ref_thermal(float(dt))

    def dt_cond(self):
        d = C.c_double(); self.L.ref_dt_cond(C.byref(d)); return d.value

    def source(self, i, j, k, prim):
        """source(i,j,k,prim,s) of this variant (sources.f90:188-220), from the
        primit array last set."""
        sv = np.zeros(self.neq)
        self.L.ref_source(i, j, k, np.ascontiguousarray(prim, dtype=np.float64), sv)
        return sv

    def riemann(self, pl, pr):
        ff = np.zeros(self.neq)
        self.L.ref_riemann(np.ascontiguousarray(pl), np.ascontiguousarray(pr), ff)
        return ff
