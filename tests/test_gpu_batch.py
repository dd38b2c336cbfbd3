"""Batches of independent boxes (gm_batch_*): each box advanced in a batch
equals, bit for bit, the same box advanced alone through the single-box
path with the reference's host time loop -- for FP64 (hence equal to the
restatement of the reference) and for the mixed form; a box's result does
not depend on the other boxes in the batch, on the graph length or on the
autotuned launch shape."""
import math
import sys

import numpy as np
import pytest

from guacho_amd import _capi, problems
from guacho_amd.host import timestep_clamp
from tests._pairs import random_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def single(params, u0, nsteps=None, t_end=math.inf, cfl=0.4):
    """The single-box path: host ramp/clamp; stop after nsteps or at the
    dump clamp onto t_end.  Returns (u, steps, time)."""
    with _capi.Subdomain(params) as dev:
        dev.upload_u(u0)
        t, it = 0.0, 1
        while True:
            dt, dump = timestep_clamp(it, 10, cfl, dev.cfl_dtp(), t, t_end)
            dev.tstep(dt)
            t = t + dt
            if (nsteps is not None and it >= nsteps) or dump:
                break
            it += 1
        return dev.download_u(), it, t


def boxes(prec):
    """A small uneven mix: blasts (outflow), OT (periodic), a random closed box."""
    out = []
    for n, seed in ((12, 1), (20, 2), (16, 3)):
        u, dx, dy, dz = problems.blast(n, seed=seed)
        out.append((_capi.make_params(n, n, n, dx, dy, dz, bc=(1,) * 6, precision=prec), u))
    u, dx, dy, dz = problems.orszag_tang(24, 24, 4)
    out.append((_capi.make_params(24, 24, 4, dx, dy, dz, precision=prec), u))
    u, dx, dy, dz = problems.turbulence(16, seed=4)
    out.append((_capi.make_params(16, 16, 16, dx, dy, dz, bc=(2, 2, 1, 1, 3, 3), precision=prec), u))
    # Powell 8-wave: batches add the terms in a pass of their own, single
    # boxes in the stage kernels
    u, dx, dy, dz = problems.blast(14, seed=5)
    out.append((_capi.make_params(14, 14, 14, dx, dy, dz, bc=(1,) * 6, precision=prec, eight_wave=True), u))
    u, dx, dy, dz = problems.turbulence(12, seed=6)
    out.append((_capi.make_params(12, 12, 12, dx, dy, dz, precision=prec, eight_wave=True), u))
    return out


@pytest.mark.parametrize("prec", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("spl", [1, 4])
def test_batch_equals_single_box_path(prec, spl):
    say(f"case: batch of 7 uneven boxes, precision {prec}, steps_per_launch {spl}, 8 steps")
    bx = boxes(prec)
    with _capi.Batch(precision=prec, max_boxes=8, steps_per_launch=spl) as B:
        ids = [B.add(p, u, max_steps=8) for p, u in bx]
        B.advance(8 // spl)
        B.sync()
        got = [B.download_u(i) for i in ids]
        clocks = [B.clock(i) for i in ids]
    for (p, u), g, (t, n, done) in zip(bx, got, clocks):
        ref, steps, tref = single(p, u, nsteps=8)
        np.testing.assert_array_equal(g, ref)
        assert (n, done) == (8, True) and t == tref


@pytest.mark.parametrize("prec", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED], ids=["fp64", "mixed"])
def test_eight_wave_batch_without_flux_cd(prec):
    say(f"case: batch without flux-CD, 8-wave and plain boxes, precision {prec}, 6 steps")
    bx = []
    for n, seed, eight in ((14, 7, True), (10, 8, False), (12, 9, True)):
        u, dx, dy, dz = problems.blast(n, seed=seed)
        bx.append((_capi.make_params(n, n, n, dx, dy, dz, bc=(1,) * 6, precision=prec, eight_wave=eight,
                                     enable_flux_cd=False), u))
    with _capi.Batch(precision=prec, max_boxes=4, enable_flux_cd=False) as B:
        ids = [B.add(p, u, max_steps=6) for p, u in bx]
        B.advance(6)
        B.sync()
        got = [B.download_u(i) for i in ids]
    for (p, u), g in zip(bx, got):
        np.testing.assert_array_equal(g, single(p, u, nsteps=6)[0])


def test_alone_equals_in_mix_mixed():
    say("case: one blast box alone vs among four others (mixed), 10 steps")
    bx = boxes(_capi.GM_PRECISION_MIXED)
    p, u = bx[1]
    with _capi.Batch(precision=_capi.GM_PRECISION_MIXED, max_boxes=8) as B:
        i = B.add(p, u, max_steps=10)
        B.advance(10)
        B.sync()
        alone = B.download_u(i)
    with _capi.Batch(precision=_capi.GM_PRECISION_MIXED, max_boxes=8) as B:
        ids = [B.add(q, v, max_steps=10 + k) for k, (q, v) in enumerate(bx)]
        B.advance(10)
        B.sync()
        mixed = B.download_u(ids[1])
    np.testing.assert_array_equal(alone, mixed)


@pytest.mark.parametrize("prec", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("autotune", [False, True], ids=["fixed", "autotuned"])
def test_viscous_boxes_equal_single_box_path(prec, autotune):
    """Boxes with artificial viscosity of their own (viscous_copy,
    src/hydro_solver.f90:47-65, 188) beside one without, in one batch: each
    equals the single-box path bit for bit.  The graph runs three steps past
    the boxes' last (5 steps in launches of 4), and the autotuner times
    shapes with dt = 0 steps; neither may diffuse a box further."""
    say(f"case: viscous batch, eta 0.02 / 0 / 0.05 / 0.01, precision {prec}, autotune {autotune}, 5 steps")
    bx = []
    for n, seed, eta in ((12, 11, 0.02), (16, 12, 0.0), (10, 13, 0.05)):
        u, dx, dy, dz = problems.blast(n, seed=seed)
        bx.append((_capi.make_params(n, n, n, dx, dy, dz, bc=(1,) * 6, precision=prec, eta=eta), u))
    u, dx, dy, dz = problems.turbulence(12, seed=14)
    bx.append((_capi.make_params(12, 12, 12, dx, dy, dz, bc=(2, 2, 1, 1, 3, 3), precision=prec, eta=0.01), u))
    with _capi.Batch(precision=prec, max_boxes=8, steps_per_launch=4, autotune=autotune) as B:
        ids = [B.add(p, u, max_steps=5) for p, u in bx]
        B.advance(2)
        B.sync()
        got = [B.download_u(i) for i in ids]
        clocks = [B.clock(i) for i in ids]
    for (p, u), g, (t, n, done) in zip(bx, got, clocks):
        ref, steps, tref = single(p, u, nsteps=5)
        np.testing.assert_array_equal(g, ref)
        assert (n, done) == (5, True) and t == tref


@pytest.mark.parametrize("slab", [0, 8, -2], ids=["two-kernel", "slab-predictor-8", "slab-both-2"])
@pytest.mark.parametrize("npas", [1, 2])
@pytest.mark.parametrize("prec", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED], ids=["fp64", "mixed"])
def test_passive_scalar_boxes_equal_single_box_path(prec, npas, slab, monkeypatch):
    """Boxes carrying passive scalars (PASSIVES builds, neq = 8 + npas): the
    batch advances them, 8-wave terms and viscosity included, through each
    stage-kernel form, and each box equals the single-box path bit for bit.
    A box with another passive-scalar count is refused while they run."""
    say(f"case: batch with {npas} passive scalars, stage kernels {slab}, precision {prec}, 6 steps")
    monkeypatch.setenv("GUACHO_AMD_BATCH_SLAB", str(slab))
    rng = np.random.default_rng(20 + npas)
    bx = []
    for (nx, ny, nz), bc, extra in (((12, 10, 8), (1,) * 6, {}),
                                    ((16, 16, 6), (3,) * 6, {"eta": 0.02}),
                                    ((10, 14, 12), (2, 2, 1, 1, 3, 3), {"eight_wave": True})):
        u = random_state(rng, nx, ny, nz, npas=npas)
        bx.append((_capi.make_params(nx, ny, nz, 1 / nx, 1 / ny, 1 / nz, bc=bc, precision=prec, npas=npas, **extra), u))
    with _capi.Batch(precision=prec, max_boxes=4, steps_per_launch=3) as B:
        ids = [B.add(p, u, max_steps=6) for p, u in bx]
        p8 = _capi.make_params(8, 8, 8, 1 / 8, 1 / 8, 1 / 8, precision=prec)
        with pytest.raises(_capi.GmError) as ei:
            B.add(p8, random_state(rng, 8, 8, 8), max_steps=1)
        assert ei.value.code == _capi.GM_ERR_CONFIG
        B.advance(2)
        B.sync()
        got = [B.download_u(i) for i in ids]
    for (p, u), g in zip(bx, got):
        assert g.shape == u.shape
        np.testing.assert_array_equal(g, single(p, u, nsteps=6)[0])


def test_passive_scalar_count_follows_the_boxes_held():
    """A batch takes the passive-scalar count of the boxes it holds: once
    its last box of one count is removed, a box of another count is admitted
    and stepped (its graph captured anew) like the single-box path."""
    say("case: batch empties of npas = 1 boxes, then takes an npas = 0 box, 4 steps each")
    rng = np.random.default_rng(31)
    p1 = _capi.make_params(10, 12, 6, 0.1, 1 / 12, 1 / 6, npas=1)
    u1 = random_state(rng, 10, 12, 6, npas=1)
    p0 = _capi.make_params(12, 8, 10, 1 / 12, 0.125, 0.1, bc=(1,) * 6)
    u0 = random_state(rng, 12, 8, 10)
    with _capi.Batch(precision=_capi.GM_PRECISION_FP64, max_boxes=2, steps_per_launch=4) as B:
        a = B.add(p1, u1, max_steps=4)
        B.advance(1)
        B.sync()
        g1 = B.download_u(a)
        B.remove(a)
        b = B.add(p0, u0, max_steps=4)
        B.advance(1)
        B.sync()
        g0 = B.download_u(b)
    np.testing.assert_array_equal(g1, single(p1, u1, nsteps=4)[0])
    np.testing.assert_array_equal(g0, single(p0, u0, nsteps=4)[0])


def test_t_end_completion_matches_host_loop():
    say("case: OT 32x32x4 to t_end = 0.05 in a batch vs the host loop")
    u, dx, dy, dz = problems.orszag_tang(32, 32, 4)
    p = _capi.make_params(32, 32, 4, dx, dy, dz)
    ref, steps, tref = single(p, u, t_end=0.05)
    with _capi.Batch(precision=_capi.GM_PRECISION_FP64, max_boxes=4) as B:
        i = B.add(p, u, t_end=0.05)
        while True:
            B.advance(1)
            B.sync()
            t, n, done = B.clock(i)
            if done:
                break
            assert n < 10000
        got = B.download_u(i)
        # completed boxes stay put on further advances
        B.advance(3)
        B.sync()
        assert B.clock(i) == (t, n, True)
        np.testing.assert_array_equal(B.download_u(i), got)
    assert (n, t) == (steps, tref) and t == 0.05
    np.testing.assert_array_equal(got, ref)


def test_autotune_keeps_results_and_picks_a_shape():
    say("case: autotuned batch (dt = 0 timing steps) vs single path, 6 steps")
    bx = boxes(_capi.GM_PRECISION_FP64)
    with _capi.Batch(precision=_capi.GM_PRECISION_FP64, max_boxes=8, autotune=True) as B:
        ids = [B.add(p, u, max_steps=6) for p, u in bx]
        B.advance(6)
        B.sync()
        block, kchunk, slab, ms = B.tuning()
        got = [B.download_u(i) for i in ids]
    assert block in (64, 128, 256) and kchunk in (8, 16, 32) and slab in (0, 4, 8, 16, 32, -1, -2, -4, -8, -16, -32) and ms > 0
    for (p, u), g in zip(bx, got):
        np.testing.assert_array_equal(g, single(p, u, nsteps=6)[0])


def test_budget_and_removal():
    say("case: memory budget refuses a box, removal frees room")
    u, dx, dy, dz = problems.blast(16, seed=1)
    p = _capi.make_params(16, 16, 16, dx, dy, dz, bc=(1,) * 6, precision=_capi.GM_PRECISION_MIXED)
    with _capi.Batch(precision=_capi.GM_PRECISION_MIXED, max_boxes=4, memory_budget=1) as B0:
        with pytest.raises(_capi.GmError) as ei:
            B0.add(p, u, max_steps=1)
        assert ei.value.code == _capi.GM_ERR_BUSY
    with _capi.Batch(precision=_capi.GM_PRECISION_MIXED, max_boxes=4) as B:
        need = B.box_bytes(p)
    with _capi.Batch(precision=_capi.GM_PRECISION_MIXED, max_boxes=4, memory_budget=need + need // 2) as B:
        a = B.add(p, u, max_steps=2)
        with pytest.raises(_capi.GmError) as ei:
            B.add(p, u, max_steps=2)
        assert ei.value.code == _capi.GM_ERR_BUSY
        B.remove(a)
        b = B.add(p, u, max_steps=2)
        B.advance(2)
        B.sync()
        assert B.clock(b)[2]
        assert B.active() == 1


@pytest.mark.parametrize("slab", [8, -2], ids=["predictor-8", "both-2"])
@pytest.mark.parametrize("prec", [_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED], ids=["fp64", "mixed"])
def test_batch_slab_kernel_equals_single_box_path(prec, slab, monkeypatch):
    """The batch with the 2.5-D stage kernel (kb_slab) for the predictor (8
    planes per slab: boxes of one to three slabs) or for both stages (2
    planes per slab) equals the single-box path bit for bit."""
    say(f"case: batch with the slab stage kernel ({slab}) vs single path, precision {prec}, 5 steps")
    monkeypatch.setenv("GUACHO_AMD_BATCH_SLAB", str(slab))
    bx = boxes(prec)
    with _capi.Batch(precision=prec, max_boxes=8, steps_per_launch=5) as B:
        ids = [B.add(p, u, max_steps=5) for p, u in bx]
        assert B.tuning()[2] == slab
        B.advance(1)
        B.sync()
        got = [B.download_u(i) for i in ids]
    for (p, u), g in zip(bx, got):
        np.testing.assert_array_equal(g, single(p, u, nsteps=5)[0])
