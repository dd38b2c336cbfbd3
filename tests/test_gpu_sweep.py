"""A seeded sweep of random configurations against the oracle, bit for bit.

Grid sizes of 2 to 45 cells per axis (ragged against the 32 x 16 plane
tiles, the 32 x 8 column tiles and the z pencils), solver, limiter,
boundary kinds per axis and flux-CD on or off are drawn from one seed.
Each case runs 3 steps through each stage-kernel form: the size-aware
default, which runs these small grids as slab kernels for both stages; the
slab predictor with the two-kernel corrector, as large grids run; and the
two-kernel stage (plane tiles, then z pencils) for both.  The CFL minimum
must agree every step and the final state bit for bit.
"""
import sys

import numpy as np
import pytest

from guacho_amd import _capi
from guacho_amd.host import timestep_clamp
from tests._pairs import random_state, run_pair

pytestmark = pytest.mark.gpu

P, O_, C_ = 3, 1, 2  # periodic, outflow, closed
SOLVERS = [(4, True, False), (3, True, False), (1, True, False), (2, False, True), (3, False, True)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cases(n=32, seed=2026):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx, ny, nz = (int(v) for v in rng.integers(2, 46, 3))
        solver, mhd, pmhd = SOLVERS[int(rng.integers(len(SOLVERS)))]
        lim = int(rng.integers(-1, 7))
        bc = []
        for _ in range(3):
            k = int(rng.choice([P, O_, C_]))
            bc += [P, P] if k == P else [k, int(rng.choice([O_, C_]))]  # periodic pairs only
        fcd = bool(rng.integers(2))
        out.append((i, nx, ny, nz, solver, mhd, pmhd, lim, tuple(bc), fcd))
    return out


@pytest.mark.parametrize("form", ["default", "slab-predictor", "two-kernel"])
@pytest.mark.parametrize("case", cases(), ids=lambda c: f"c{c[0]}")
def test_random_configuration_bitwise(case, form, monkeypatch):
    i, nx, ny, nz, solver, mhd, pmhd, lim, bc, fcd = case
    if form == "two-kernel":
        monkeypatch.setenv("GUACHO_AMD_SLAB_KCHUNK", "0")
    elif form == "slab-predictor":
        monkeypatch.setenv("GUACHO_AMD_SLAB_STAGES", "1")
    say(f"case {i} ({form}): {nx}x{ny}x{nz} solver {solver} mhd {mhd} pmhd {pmhd} limiter {lim} "
        f"bc {bc} flux-CD {fcd}, 3 steps")
    u0 = random_state(np.random.default_rng(100 + i), nx, ny, nz)
    p = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, riemann_solver=solver, mhd=mhd, pmhd=pmhd,
                          slope_limiter=lim, bc=bc, enable_flux_cd=fcd)
    ud, uo, _ = run_pair(p, u0, 3)
    np.testing.assert_array_equal(ud, uo)


@pytest.mark.parametrize("form", ["default", "two-kernel"])
@pytest.mark.parametrize("case", cases()[:16], ids=lambda c: f"c{c[0]}")
def test_random_configuration_mixed_within_tolerance(case, form, monkeypatch):
    """The mixed form on the same cases, against the FP64 device path with
    the same dt sequence: density within 1e-5 relative L1, every conserved
    variable within 1e-4, after 3 steps."""
    i, nx, ny, nz, solver, mhd, pmhd, lim, bc, fcd = case
    if form == "two-kernel":
        monkeypatch.setenv("GUACHO_AMD_SLAB_KCHUNK", "0")
    say(f"case {i} mixed ({form}): {nx}x{ny}x{nz} solver {solver} limiter {lim} bc {bc} flux-CD {fcd}, 3 steps")
    u0 = random_state(np.random.default_rng(100 + i), nx, ny, nz)
    kw = dict(riemann_solver=solver, mhd=mhd, pmhd=pmhd, slope_limiter=lim, bc=bc, enable_flux_cd=fcd)
    p64 = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, precision=_capi.GM_PRECISION_FP64, **kw)
    p32 = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, precision=_capi.GM_PRECISION_MIXED, **kw)
    with _capi.Subdomain(p64) as d64, _capi.Subdomain(p32) as d32:
        for d in (d64, d32):
            d.upload_u(u0)
            d.boundary_u()
        t = 0.0
        for it in range(1, 4):
            dt, _ = timestep_clamp(it, 10, 0.4, d64.cfl_dtp(), t, 1e30)
            d32.cfl_dtp()
            d64.tstep(dt)
            d32.tstep(dt)
            t += dt
        a, b = d64.download_u(), d32.download_u()
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    a, b = a[inner], b[inner]
    assert np.all(np.isfinite(b))
    rel = np.abs(a - b).sum(axis=(1, 2, 3)) / np.maximum(np.abs(a).sum(axis=(1, 2, 3)), 1e-300)
    assert rel[0] <= 1e-5 and np.all(rel <= 1e-4), rel
