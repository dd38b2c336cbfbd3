"""The 2.5-D stage kernel (k_slab: x, y and z faces and the whole update in
one pass marching up z) against the two-kernel stage (k_xy plane tiles,
then k_z pencils through W): the same bits after several steps, every
step's CFL minimum included, with the slab kernel for both stages at slab
lengths from one plane to the whole grid and with the library's default
(on these grids, which do not fill the GPU, the slab kernel for both stages
with one or two planes per slab; at 64^3 and up the slab predictor and the
two-kernel corrector), partial tiles in x and y, every
boundary kind, the four solvers, passive scalars, FP64 and mixed precision.
(The FP64 path is also checked against the oracle bit for bit by the other
GPU tests, which run the default.)"""
import contextlib
import os

import numpy as np
import pytest

from guacho_amd import _capi, problems
from guacho_amd.host import timestep_clamp

pytestmark = pytest.mark.gpu

ENV = "GUACHO_AMD_SLAB_KCHUNK"


ENV_STAGES = "GUACHO_AMD_SLAB_STAGES"


@contextlib.contextmanager
def slab_kchunk(v):
    """v planes per slab for both stages (0: the two-kernel stage for both;
    None: the library's default choice)."""
    old = {e: os.environ.get(e) for e in (ENV, ENV_STAGES)}
    for e in (ENV, ENV_STAGES):
        os.environ.pop(e, None)
    if v is not None:
        os.environ[ENV] = str(v)
        os.environ[ENV_STAGES] = "12"
    try:
        yield
    finally:
        for e, val in old.items():
            if val is None:
                os.environ.pop(e, None)
            else:
                os.environ[e] = val


def run(prm, u0, steps, kchunk):
    with slab_kchunk(kchunk):
        with _capi.Subdomain(prm) as dev:
            dev.upload_u(u0)
            t, dts = 0.0, []
            for it in range(1, steps + 1):
                dtp = dev.cfl_dtp()
                dts.append(dtp)
                dt, _ = timestep_clamp(it, 10, 0.4, dtp, t, 1e30)
                dev.tstep(dt)
                t += dt
            dts.append(dev.cfl_dtp())
            return dev.download_u(), dts


def random_state(nx, ny, nz, npas, seed):
    rng = np.random.default_rng(seed)
    shp = (nx + 4, ny + 4, nz + 4)
    rho = rng.uniform(0.6, 1.6, shp)
    p = rng.uniform(0.4, 1.4, shp)
    v = rng.uniform(-0.6, 0.6, (3,) + shp)
    B = rng.uniform(-0.8, 0.8, (3,) + shp)
    u = np.zeros((8 + npas,) + shp)
    u[0] = rho
    u[1:4] = rho * v
    u[4] = 0.5 * rho * (v * v).sum(0) + 1.5 * p + 0.5 * (B * B).sum(0)
    u[5:8] = B
    for a in range(npas):
        u[8 + a] = rho * rng.uniform(0.0, 1.0, shp)
    return np.asfortranarray(u)


CASES = [
    # name, nx, ny, nz, solver, limiter, bc, flux_cd, npas
    ("hlld-minmod-periodic", 45, 19, 13, 4, 1, (3, 3, 3, 3, 3, 3), True, 0),
    ("hlle-vanleer-outflow", 33, 9, 11, 3, 2, (1, 1, 1, 1, 1, 1), True, 0),
    ("hll-umist-closed", 31, 17, 7, 1, 4, (2, 2, 2, 2, 2, 2), False, 0),
    ("hlld-woodward-mixed-bc", 70, 12, 9, 4, 5, (2, 2, 1, 1, 3, 3), True, 0),
    ("hllc-superbee-pmhd", 24, 10, 10, 2, 6, (3, 3, 3, 3, 3, 3), True, 0),
    ("hlld-minmod-2passives", 40, 16, 12, 4, 1, (1, 1, 3, 3, 2, 2), True, 2),
]


@pytest.mark.parametrize("precision", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_slab_equals_two_kernel_stage(case, precision):
    name, nx, ny, nz, solver, lim, bc, fcd, npas = case
    print(f"case: slab vs two-kernel stage, {name} {nx}x{ny}x{nz}, precision {precision}", flush=True)
    pm = solver == 2
    prm = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, riemann_solver=solver, slope_limiter=lim,
                            bc=bc, enable_flux_cd=fcd, mhd=not pm, pmhd=pm, precision=precision, npas=npas)
    u0 = random_state(nx, ny, nz, npas, seed=nx * 7 + ny)
    ref_u, ref_dt = run(prm, u0, 3, 0)
    for k in (1, 4, nz, None):
        u, dts = run(prm, u0, 3, k)
        assert dts == ref_dt, (name, k)
        np.testing.assert_array_equal(u, ref_u, err_msg=f"{name} slab length {k}")


def test_slab_turbulence_64_bitwise():
    print("case: slab vs two-kernel stage, turbulence 64^3 FP64 and mixed, 5 steps", flush=True)
    u0, dx, dy, dz = problems.turbulence(64, seed=5)
    for prec in (_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED):
        prm = _capi.make_params(64, 64, 64, dx, dy, dz, precision=prec)
        a, da = run(prm, u0, 5, 0)
        b, db = run(prm, u0, 5, None)
        assert da == db
        np.testing.assert_array_equal(a, b)
