"""guacho_amd -- MI355X-native Godunov MHD step of Guacho-3D.

The step (predictor/corrector with limited reconstruction, HLLD/HLLE/HLL/HLLC
fluxes, flux-CD, boundaries, CFL reduction) runs as hand-written gfx950 HIP
kernels in libguacho_amd.so behind the C ABI of include/guacho_amd.h.
This package is the host side: the ctypes binding (_capi), the time loop
mirroring the reference's main loop (host), problem set-ups (problems) and
the BIN dump format (binio).
"""
from . import _capi, binio, host, problems  # noqa: F401
from ._capi import GmError, Subdomain, make_params  # noqa: F401

__all__ = ["GmError", "Subdomain", "make_params", "host", "problems", "binio"]
