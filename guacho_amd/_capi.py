"""ctypes binding of the C ABI in include/guacho_amd.h.

The product path: every compute call goes to guacho_amd/libguacho_amd.so.
There is no fallback; a missing library raises at load time.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# the library; GUACHO_AMD_LIB selects another build of it (the host-sanitizer
# build of `make -C guacho_amd asan`)
LIB_PATH = os.environ.get("GUACHO_AMD_LIB") or os.path.join(_HERE, "libguacho_amd.so")

GM_OK, GM_ERR_CONFIG, GM_ERR_HIP, GM_ERR_NONFINITE, GM_ERR_OOM, GM_ERR_ARG, GM_ERR_BUSY = 0, 1, 2, 3, 4, 5, 6
GM_PRECISION_FP64, GM_PRECISION_MIXED = 0, 1
GM_COOL_NONE, GM_COOL_DMC, GM_COOL_CHI = 0, 3, 4
GM_TC_OFF, GM_TC_ISOTROPIC, GM_TC_ANISOTROPIC = 0, 1, 2
GM_COOL_H = 1
GM_EOS_ADIABATIC, GM_EOS_SINGLE_SPECIE, GM_EOS_H_RATE = 1, 2, 3
GM_NKCLASS = 15

# exported entry points, in the order include/guacho_amd.h declares them
EXPORTS = (
    "gm_create", "gm_destroy", "gm_last_error", "gm_upload_u", "gm_download_u",
    "gm_download_primit", "gm_boundary_u", "gm_cfl_dtp", "gm_tstep", "gm_sync", "gm_tstep_noalloc",
    "gm_download_divb", "gm_set_timing", "gm_kernel_times", "gm_device_bytes",
    "gm_set_time", "gm_set_user_sources", "gm_set_user_bc", "gm_set_cooling", "gm_set_conduction",
    "gm_conduction_substeps", "gm_set_eos",
    "gm_batch_create", "gm_batch_destroy", "gm_batch_last_error", "gm_batch_box_bytes",
    "gm_batch_free_bytes", "gm_batch_add", "gm_batch_remove", "gm_batch_advance", "gm_batch_sync",
    "gm_batch_query", "gm_batch_clock", "gm_batch_download_u", "gm_batch_tuning", "gm_batch_active",
)


class GmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class GmParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("neq", C.c_int32), ("nghost", C.c_int32),
        ("riemann_solver", C.c_int32), ("slope_limiter", C.c_int32),
        ("bc", C.c_int32 * 6),
        ("mhd", C.c_int32), ("pmhd", C.c_int32),
        ("enable_flux_cd", C.c_int32), ("eight_wave", C.c_int32),
        ("precision", C.c_int32), ("device", C.c_int32), ("reserved", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("cv", C.c_double), ("gamma", C.c_double), ("eta", C.c_double),
    ]


class GmConductionParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("saturation", C.c_int32),
                ("tempsc", C.c_double), ("rsc", C.c_double), ("rhosc", C.c_double), ("mu", C.c_double),
                ("psc", C.c_double), ("tsc", C.c_double), ("vsc2", C.c_double), ("bsc", C.c_double)]


class GmBatchParams(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("riemann_solver", C.c_int32), ("enable_flux_cd", C.c_int32),
        ("device", C.c_int32), ("max_boxes", C.c_int32), ("steps_per_launch", C.c_int32),
        ("autotune", C.c_int32), ("reserved", C.c_int32), ("memory_budget", C.c_int64),
    ]


_LIB: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libguacho_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C guacho_amd` "
                           "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
    L.gm_create.argtypes = [C.POINTER(GmParams), C.POINTER(vp)]
    L.gm_destroy.argtypes = [vp]
    L.gm_last_error.argtypes = [vp]
    L.gm_last_error.restype = C.c_char_p
    L.gm_upload_u.argtypes = [vp, dp]
    L.gm_download_u.argtypes = [vp, dp]
    L.gm_download_primit.argtypes = [vp, dp, C.c_void_p]
    L.gm_boundary_u.argtypes = [vp]
    L.gm_cfl_dtp.argtypes = [vp, C.POINTER(C.c_double)]
    L.gm_tstep.argtypes = [vp, C.c_double]
    L.gm_sync.argtypes = [vp]
    L.gm_tstep_noalloc.argtypes = [C.POINTER(GmParams), C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
    L.gm_download_divb.argtypes = [vp, dp]
    L.gm_set_timing.argtypes = [vp, C.c_int]
    L.gm_kernel_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_char_p)]
    L.gm_device_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.gm_set_time.argtypes = [vp, C.c_double]
    L.gm_set_user_sources.argtypes = [vp, C.c_char_p, C.c_char_p]
    L.gm_set_user_bc.argtypes = [vp, C.c_char_p, C.c_char_p]
    L.gm_set_cooling.argtypes = [vp, C.c_int32, C.c_void_p, C.c_double, C.c_double, C.c_double]
    L.gm_set_conduction.argtypes = [vp, C.POINTER(GmConductionParams)]
    L.gm_conduction_substeps.argtypes = [vp, C.POINTER(C.c_int32)]
    L.gm_set_eos.argtypes = [vp, C.c_int32, C.c_double]
    P = C.POINTER(GmParams)
    L.gm_batch_create.argtypes = [C.POINTER(GmBatchParams), C.POINTER(vp)]
    L.gm_batch_destroy.argtypes = [vp]
    L.gm_batch_last_error.argtypes = [vp]
    L.gm_batch_last_error.restype = C.c_char_p
    L.gm_batch_box_bytes.argtypes = [vp, P, C.POINTER(C.c_int64)]
    L.gm_batch_free_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.gm_batch_add.argtypes = [vp, P, dp, C.c_double, C.c_double, C.c_int64, C.POINTER(C.c_int32)]
    L.gm_batch_remove.argtypes = [vp, C.c_int32]
    L.gm_batch_advance.argtypes = [vp, C.c_int32]
    L.gm_batch_sync.argtypes = [vp]
    L.gm_batch_query.argtypes = [vp, C.POINTER(C.c_int32)]
    L.gm_batch_clock.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int32)]
    L.gm_batch_download_u.argtypes = [vp, C.c_int32, dp]
    L.gm_batch_tuning.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_double)]
    L.gm_batch_active.argtypes = [vp, C.POINTER(C.c_int32)]
    for name in EXPORTS:
        if name not in ("gm_last_error", "gm_batch_last_error"):
            getattr(L, name).restype = C.c_int
    _LIB = L
    return L


def make_params(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float, *, cv: float = 1.5,
                gamma: Optional[float] = None, riemann_solver: int = 4, slope_limiter: int = 1,
                bc: Sequence[int] = (3, 3, 3, 3, 3, 3), mhd: bool = True, pmhd: bool = False,
                enable_flux_cd: bool = True, eight_wave: bool = False,
                precision: int = GM_PRECISION_FP64, device: int = 0, eta: float = 0.0,
                npas: int = 0) -> GmParams:
    p = GmParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.neq, p.nghost = 8 + npas, 2
    p.riemann_solver, p.slope_limiter = riemann_solver, slope_limiter
    for b in range(6):
        p.bc[b] = int(bc[b])
    p.mhd, p.pmhd = int(mhd), int(pmhd)
    p.enable_flux_cd, p.eight_wave = int(enable_flux_cd), int(eight_wave)
    p.precision, p.device, p.reserved = precision, device, 0
    p.dx, p.dy, p.dz = dx, dy, dz
    p.cv = cv
    p.gamma = (cv + 1.0) / cv if gamma is None else gamma
    p.eta = eta
    return p


class Subdomain:
    """One device-resident subdomain (a gm_ctx)."""

    def __init__(self, params: GmParams):
        self.L = lib()
        self.params = params
        h = C.c_void_p()
        rc = self.L.gm_create(C.byref(params), C.byref(h))
        if rc != GM_OK:
            raise GmError(rc, self.L.gm_last_error(None).decode())
        self.h = h
        self.shape = (params.neq, params.nx + 4, params.ny + 4, params.nz + 4)

    def _check(self, rc: int):
        if rc != GM_OK:
            raise GmError(rc, self.L.gm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.gm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_u(self, u: np.ndarray):
        u = np.asfortranarray(u, dtype=np.float64)
        if u.shape != self.shape:
            raise ValueError(f"u has shape {u.shape}, expected {self.shape}")
        self._check(self.L.gm_upload_u(self.h, u))

    def download_u(self) -> np.ndarray:
        u = np.zeros(self.shape, dtype=np.float64, order="F")
        self._check(self.L.gm_download_u(self.h, u))
        return u

    def download_primit(self, with_temp: bool = False):
        p = np.zeros(self.shape, dtype=np.float64, order="F")
        t = np.zeros(self.shape[1:], dtype=np.float64, order="F") if with_temp else None
        self._check(self.L.gm_download_primit(self.h, p, t.ctypes.data if t is not None else None))
        return (p, t) if with_temp else p

    def boundary_u(self):
        self._check(self.L.gm_boundary_u(self.h))

    def cfl_dtp(self) -> float:
        v = C.c_double()
        self._check(self.L.gm_cfl_dtp(self.h, C.byref(v)))
        return v.value

    def tstep(self, dt: float):
        self._check(self.L.gm_tstep(self.h, float(dt)))

    def sync(self):
        self._check(self.L.gm_sync(self.h))

    def divb(self) -> np.ndarray:
        p = self.params
        d = np.zeros((p.nx, p.ny, p.nz), dtype=np.float64, order="F")
        self._check(self.L.gm_download_divb(self.h, d))
        return d

    def set_timing(self, enable: bool):
        self._check(self.L.gm_set_timing(self.h, int(enable)))

    def set_time(self, time: float) -> None:
        """Simulation time the user plugins see (globals::time)."""
        self._check(self.L.gm_set_time(self.h, float(time)))

    def set_user_sources(self, code_object: Optional[str], kernel: Optional[str] = None) -> None:
        """Load the user's source-term kernel from a gfx950 code object
        (None removes it)."""
        co = code_object.encode() if code_object else None
        kn = kernel.encode() if kernel else None
        self._check(self.L.gm_set_user_sources(self.h, co, kn))

    def set_user_bc(self, code_object: Optional[str], kernel: Optional[str] = None) -> None:
        """Load the user's boundary hook (impose_user_bc) from a gfx950 code
        object (None removes it); ghosts then live in memory."""
        co = code_object.encode() if code_object else None
        kn = kernel.encode() if kernel else None
        self._check(self.L.gm_set_user_bc(self.h, co, kn))

    def set_cooling(self, kind: int, cooltab=None, tempsc: float = 1.0, psc: float = 1.0, tsc: float = 1.0) -> None:
        """Cooling after each step (GM_COOL_DMC / GM_COOL_CHI with cooltab, the
        reference's cooltab(2, 41); GM_COOL_H without a table; 0 off)."""
        if kind in (GM_COOL_NONE, GM_COOL_H):
            self._check(self.L.gm_set_cooling(self.h, int(kind), None, tempsc, psc, tsc))
            return
        tab = np.asfortranarray(cooltab, dtype=np.float64)
        if tab.shape != (2, 41):
            raise ValueError("cooltab must be (2, 41)")
        self._check(self.L.gm_set_cooling(self.h, int(kind), tab.ctypes.data_as(C.c_void_p), tempsc, psc, tsc))

    def set_conduction(self, kind: int, saturation: bool = False, tempsc: float = 1.0, rsc: float = 1.0,
                       rhosc: float = 1.0, mu: float = 1.0, psc: float = 1.0, tsc: float = 1.0,
                       vsc2: float = 1.0, bsc: float = 1.0) -> None:
        """Thermal conduction after each step (GM_TC_ISOTROPIC / GM_TC_ANISOTROPIC; 0 off)."""
        tp = GmConductionParams(int(kind), int(bool(saturation)), tempsc, rsc, rhosc, mu, psc, tsc, vsc2, bsc)
        self._check(self.L.gm_set_conduction(self.h, C.byref(tp)))

    def conduction_substeps(self) -> int:
        n = C.c_int32()
        self._check(self.L.gm_conduction_substeps(self.h, C.byref(n)))
        return n.value

    def set_eos(self, eq_of_state: int, tempsc: float = 1.0) -> None:
        """Equation of state of u2prim (GM_EOS_ADIABATIC / GM_EOS_SINGLE_SPECIE) and its Tempsc."""
        self._check(self.L.gm_set_eos(self.h, int(eq_of_state), float(tempsc)))

    def kernel_times(self) -> Dict[str, tuple]:
        ms = (C.c_double * GM_NKCLASS)()
        n = (C.c_int64 * GM_NKCLASS)()
        names = (C.c_char_p * GM_NKCLASS)()
        self._check(self.L.gm_kernel_times(self.h, ms, n, names))
        return {names[k].decode(): (ms[k], n[k]) for k in range(GM_NKCLASS)}

    def device_bytes(self) -> int:
        v = C.c_int64()
        self._check(self.L.gm_device_bytes(self.h, C.byref(v)))
        return v.value


class Batch:
    """A batch of independent boxes advanced together (gm_batch_*)."""

    def __init__(self, precision: int = GM_PRECISION_MIXED, riemann_solver: int = 4, enable_flux_cd: bool = True,
                 max_boxes: int = 256, steps_per_launch: int = 1, autotune: bool = False,
                 memory_budget: int = 0, device: int = 0):
        self.L = lib()
        bp = GmBatchParams()
        bp.precision, bp.riemann_solver, bp.enable_flux_cd = precision, riemann_solver, int(enable_flux_cd)
        bp.device, bp.max_boxes, bp.steps_per_launch = device, max_boxes, steps_per_launch
        bp.autotune, bp.reserved, bp.memory_budget = int(autotune), 0, memory_budget
        self.bp = bp
        h = C.c_void_p()
        rc = self.L.gm_batch_create(C.byref(bp), C.byref(h))
        if rc != GM_OK:
            raise GmError(rc, self.L.gm_batch_last_error(None).decode())
        self.h = h
        self.shapes = {}

    def _check(self, rc: int):
        if rc != GM_OK:
            raise GmError(rc, self.L.gm_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.gm_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def box_bytes(self, params: GmParams) -> int:
        v = C.c_int64()
        self._check(self.L.gm_batch_box_bytes(self.h, C.byref(params), C.byref(v)))
        return v.value

    def free_bytes(self) -> int:
        v = C.c_int64()
        self._check(self.L.gm_batch_free_bytes(self.h, C.byref(v)))
        return v.value

    def add(self, params: GmParams, u: np.ndarray, cfl: float = 0.4, t_end: float = float("inf"),
            max_steps: int = 0) -> int:
        u = np.asfortranarray(u, dtype=np.float64)
        shape = (params.neq, params.nx + 4, params.ny + 4, params.nz + 4)
        if u.shape != shape:
            raise ValueError(f"u has shape {u.shape}, expected {shape}")
        i = C.c_int32()
        self._check(self.L.gm_batch_add(self.h, C.byref(params), u, cfl, t_end, max_steps, C.byref(i)))
        self.shapes[i.value] = shape
        return i.value

    def remove(self, box: int):
        self._check(self.L.gm_batch_remove(self.h, box))
        self.shapes.pop(box, None)

    def advance(self, nlaunch: int = 1):
        self._check(self.L.gm_batch_advance(self.h, nlaunch))

    def sync(self):
        self._check(self.L.gm_batch_sync(self.h))

    def query(self) -> bool:
        """True once the queued launches have finished (clocks refreshed);
        never waits."""
        idle = C.c_int32()
        self._check(self.L.gm_batch_query(self.h, C.byref(idle)))
        return bool(idle.value)

    def clock(self, box: int):
        t, n, d = C.c_double(), C.c_int64(), C.c_int32()
        self._check(self.L.gm_batch_clock(self.h, box, C.byref(t), C.byref(n), C.byref(d)))
        return t.value, n.value, bool(d.value)

    def download_u(self, box: int) -> np.ndarray:
        u = np.zeros(self.shapes[box], dtype=np.float64, order="F")
        self._check(self.L.gm_batch_download_u(self.h, box, u))
        return u

    def tuning(self):
        """(column tile width, z-pencil length, slab planes (0: two-kernel
        stage; positive: slab kernel for the predictor; negative: for both
        stages, -slab planes), autotuned ms per step)"""
        b, k, sl, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        self._check(self.L.gm_batch_tuning(self.h, C.byref(b), C.byref(k), C.byref(sl), C.byref(ms)))
        return b.value, k.value, sl.value, ms.value

    def active(self) -> int:
        v = C.c_int32()
        self._check(self.L.gm_batch_active(self.h, C.byref(v)))
        return v.value
