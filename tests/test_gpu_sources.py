"""Source terms of step() on the device against the C restatement.

Powell 8-wave div B terms (src/sources.f90:124-175) and a user source
plugin (the device form of get_user_source_terms, src/sources.f90:206),
applied as up = up + dt*s after each stage's update
(src/hydro_solver.f90:115-121).  The restatement's source() is pinned
bit for bit against the reference's own sources.f90 (tests/golden
fixtures ref_hlld_minmod_8wave, ref_hlle_minmod_usersrc_8wave,
ref_hlld_minmod_usersrc); here the FP64 device path must equal the
restatement bit for bit, every step.  The test plugin
(tests/plugins/gravity.hip) computes the same gravity as the oracle's
Fortran user_mod.
"""
import math
import zlib
import os
import sys

import numpy as np
import pytest

from guacho_amd import _capi, problems
from guacho_amd.host import timestep_clamp
from tests._pairs import random_state, run_pair

pytestmark = pytest.mark.gpu

PLUGIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plugins", "gravity.hsaco")
GRAVITY = (PLUGIN, "gm_test_gravity")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # name, (nx, ny, nz), solver, limiter, bc, flux_cd, pmhd
    ("hlld-minmod-periodic-cd", (16, 12, 10), 4, 1, (3,) * 6, True, False),
    ("hlle-vanleer-outflow", (12, 10, 8), 3, 2, (1,) * 6, False, False),
    ("hlld-woodward-mixed-cd", (10, 14, 8), 4, 5, (2, 2, 1, 1, 3, 3), True, False),
    ("hllc-superbee-pmhd", (12, 8, 10), 2, 6, (3, 3, 2, 2, 1, 1), False, True),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_eight_wave_bitwise(case):
    name, (nx, ny, nz), solver, lim, bc, fcd, pmhd = case
    say(f"case: eight_wave {name} {nx}x{ny}x{nz}, 6 steps, checked each step")
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    u0 = random_state(rng, nx, ny, nz)
    p = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, riemann_solver=solver, slope_limiter=lim,
                          bc=bc, enable_flux_cd=fcd, eight_wave=True, mhd=not pmhd, pmhd=pmhd)
    ud, uo, _ = run_pair(p, u0, 6, check_each=True)
    np.testing.assert_array_equal(ud, uo)


@pytest.mark.parametrize("eight", [False, True], ids=["user", "user+8wave"])
def test_user_source_plugin_bitwise(eight):
    say(f"case: user gravity plugin (eight_wave={eight}) 16x12x10 outflow/closed, 6 steps")
    rng = np.random.default_rng(7 + int(eight))
    nx, ny, nz = 16, 12, 10
    u0 = random_state(rng, nx, ny, nz)
    p = _capi.make_params(nx, ny, nz, 1.0 / nx, 1.0 / ny, 1.0 / nz, bc=(1, 1, 2, 2, 3, 3),
                          eight_wave=eight)
    ud, uo, _ = run_pair(p, u0, 6, check_each=True, user_sources=GRAVITY)
    np.testing.assert_array_equal(ud, uo)


def test_sources_change_the_result():
    """The source pass is not a no-op: 8-wave on a state with div B != 0
    differs from the plain step."""
    say("case: 8-wave vs plain step differ on a random state")
    rng = np.random.default_rng(3)
    u0 = random_state(rng, 12, 12, 12)
    outs = []
    for eight in (False, True):
        p = _capi.make_params(12, 12, 12, 1 / 12, 1 / 12, 1 / 12, eight_wave=eight)
        with _capi.Subdomain(p) as dev:
            dev.upload_u(u0)
            dev.tstep(1e-3)
            outs.append(dev.download_u())
    assert not np.array_equal(outs[0], outs[1])


def test_eight_wave_mixed_against_fp64():
    say("case: eight_wave mixed vs FP64, turbulence 32^3, 10 steps")
    u0, dx, dy, dz = problems.turbulence(32, seed=11)
    res = {}
    for prec in (_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED):
        p = _capi.make_params(32, 32, 32, dx, dy, dz, eight_wave=True, enable_flux_cd=False, precision=prec)
        with _capi.Subdomain(p) as dev:
            dev.upload_u(u0)
            t = 0.0
            for it in range(1, 11):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
                dev.tstep(dt)
                t += dt
            res[prec] = dev.download_u()[:, 2:-2, 2:-2, 2:-2]
    a, b = res[_capi.GM_PRECISION_FP64], res[_capi.GM_PRECISION_MIXED]
    rel = np.abs(a - b).sum(axis=(1, 2, 3)) / np.maximum(np.abs(a).sum(axis=(1, 2, 3)), 1e-300)
    assert rel[0] <= 1e-5 and np.all(rel <= 1e-4), rel


def test_batch_eight_wave_equals_single_path():
    """Boxes with and without 8-wave terms in one batch: each equals its
    single-box run bit for bit."""
    say("case: batch of 3 boxes (8-wave on two), 5 steps, vs single-box path")
    rng = np.random.default_rng(5)
    boxes = []
    for n, eight, bc in ((10, True, (1,) * 6), (12, False, (3,) * 6), (8, True, (2, 2, 3, 3, 1, 1))):
        u0 = random_state(rng, n, n, n)
        boxes.append((_capi.make_params(n, n, n, 1 / n, 1 / n, 1 / n, bc=bc, eight_wave=eight), u0))
    with _capi.Batch(precision=_capi.GM_PRECISION_FP64, max_boxes=4, steps_per_launch=5) as B:
        ids = [B.add(p, u, max_steps=5) for p, u in boxes]
        B.advance(1)
        B.sync()
        got = [B.download_u(i) for i in ids]
    for (p, u0), g in zip(boxes, got):
        with _capi.Subdomain(p) as dev:
            dev.upload_u(u0)
            t = 0.0
            for it in range(1, 6):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, math.inf)
                dev.tstep(dt)
                t += dt
            np.testing.assert_array_equal(g, dev.download_u())


def test_user_sources_load_errors():
    say("case: plugin load errors")
    p = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8)
    with _capi.Subdomain(p) as dev:
        with pytest.raises(_capi.GmError) as ei:
            dev.set_user_sources(PLUGIN + ".missing", "gm_test_gravity")
        assert ei.value.code == _capi.GM_ERR_CONFIG
        with pytest.raises(_capi.GmError) as ei:
            dev.set_user_sources(PLUGIN, "no_such_kernel")
        assert ei.value.code == _capi.GM_ERR_CONFIG
        dev.set_user_sources(*GRAVITY)
        dev.set_user_sources(None)
