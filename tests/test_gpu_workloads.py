"""The benchmark configurations at their own workloads, not reduced sizes:

* 64 MHD blasts of uneven cubic size n ~ U{32..128}, outflow, HLLD, mixed
  precision, 100 steps in one graph launch: every box bit for bit equal to
  the same box advanced alone through the single-box mixed path; the same
  64 boxes as an FP64 batch, whose three smallest boxes equal the oracle
  (the restatement of the reference) bit for bit;
* the serving stream at its stated mix (Brio-Wu 1024x4x4 to t = 0.1,
  Orszag-Tang 256^2 x 4 to t = 0.05, blasts 64^3..192^3 for 40 steps):
  every request's result the same alone and in the mix, bit for bit, and
  within 1e-5 relative L1 in density of the plain FP64 solve;
* the seeded turbulence box at 256^3, 50 steps, mixed against FP64:
  relative L1 per field, mass and energy drift and max |div B| within the
  stated bounds.
"""
import sys

import numpy as np
import pytest

from guacho_amd import _capi, problems, serving
from guacho_amd.host import timestep_clamp
from oracle.oracle import BC_OUTFLOW, Config, Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def single(params, u0, nsteps, cfl=0.4):
    with _capi.Subdomain(params) as dev:
        dev.upload_u(u0)
        t = 0.0
        for it in range(1, nsteps + 1):
            dt, _ = timestep_clamp(it, 10, cfl, dev.cfl_dtp(), t, 1e30)
            dev.tstep(dt)
            t = t + dt
        return dev.download_u()


def blasts64(seed=2024):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(32, 129, size=64)
    return [(int(n),) + problems.blast(int(n), seed=seed * 1000 + b) for b, n in enumerate(sizes)]


def run_batch(boxes, prec, steps):
    with _capi.Batch(precision=prec, max_boxes=len(boxes), steps_per_launch=steps, autotune=True) as B:
        ids = [B.add(_capi.make_params(n, n, n, h, h, h, bc=(1,) * 6, precision=prec), u, max_steps=steps)
               for n, u, h, _, _ in boxes]
        B.advance(1)
        B.sync()
        assert all(B.clock(i)[1:] == (steps, True) for i in ids)
        return [B.download_u(i) for i in ids]


def test_64_blasts_mixed_batch_equals_each_box_alone():
    say("case: 64 blasts n~U{32..128}^3, mixed, 100 steps in one graph launch, vs each box alone")
    boxes = blasts64()
    mp = _capi.GM_PRECISION_MIXED
    got = run_batch(boxes, mp, 100)
    for b, ((n, u0, h, _, _), ub) in enumerate(zip(boxes, got)):
        alone = single(_capi.make_params(n, n, n, h, h, h, bc=(1,) * 6, precision=mp), u0, 100)
        np.testing.assert_array_equal(ub, alone, err_msg=f"box {b} ({n}^3)")


def test_64_blasts_fp64_batch_smallest_boxes_equal_oracle():
    say("case: 64 blasts n~U{32..128}^3, FP64 batch, 100 steps; the three smallest boxes vs the oracle")
    boxes = blasts64()
    got = run_batch(boxes, _capi.GM_PRECISION_FP64, 100)
    order = sorted(range(len(boxes)), key=lambda b: boxes[b][0])[:3]
    for b in order:
        n, u0, h, _, _ = boxes[b]
        say(f"  oracle: box {b} ({n}^3), 100 steps")
        O = Oracle(Config(nx=n, ny=n, nz=n, dx=h, dy=h, dz=h, bc=[BC_OUTFLOW] * 6))
        O.set_u(u0)
        O.boundaryI()
        O.calcprim_u()
        t = 0.0
        for it in range(1, 101):
            dt, _ = timestep_clamp(it, 10, 0.4, O.cfl_dtp(), t, 1e30)
            O.tstep(dt)
            t = t + dt
        np.testing.assert_array_equal(got[b], O.get_u(), err_msg=f"box {b} ({n}^3)")
        O.close()


def test_serving_stated_mix_alone_equals_mix_and_fp64_tolerance():
    say("case: serving 6 requests of the stated mix (Brio-Wu 1024x4x4, OT 256^2, blasts 64..192^3)")
    rng = np.random.default_rng(11)
    mp = _capi.GM_PRECISION_MIXED
    reqs = []
    for i, kind in enumerate(("briowu", "ot", "blast", "blast", "ot", "briowu")):
        r = serving.make_request(i, kind, rng, mp)
        r.arrival = 0.01 * i
        reqs.append(r)
    stats = serving.serve(reqs, steps_per_launch=4, keep_results=True)
    assert stats["requests"] == 6 and stats["rejected"] == 0
    for r in reqs:
        with _capi.Batch(precision=mp, max_boxes=1, steps_per_launch=1) as B:
            box = B.add(r.params, r.u0, cfl=r.cfl, t_end=r.t_end, max_steps=r.max_steps)
            while True:
                B.advance(1)
                B.sync()
                _, steps, done = B.clock(box)
                if done:
                    break
            alone = B.download_u(box)
        assert steps == r.steps, (r.kind, steps, r.steps)
        np.testing.assert_array_equal(r.result, alone, err_msg=f"request {r.rid} ({r.kind})")
    # the same requests through the plain FP64 path
    fp = [serving.Request(r.rid, r.kind, r.params, r.u0, t_end=r.t_end, max_steps=r.max_steps) for r in reqs]
    serving.serve_sequential_fp64(fp, keep_results=True)
    for r, f in zip(reqs, fp):
        a, b = r.result[0, 2:-2, 2:-2, 2:-2], f.result[0, 2:-2, 2:-2, 2:-2]
        l1 = np.abs(a - b).sum() / np.abs(b).sum()
        say(f"  request {r.rid} ({r.kind} {r.params.nx}x{r.params.ny}x{r.params.nz}): rho rel L1 vs FP64 {l1:.2e}")
        assert l1 <= 1e-5, (r.kind, l1)


def test_turbulence_256_mixed_vs_fp64_50_steps():
    say("case: turbulence 256^3, 50 steps, mixed vs FP64")
    n = 256
    u0, dx, dy, dz = problems.turbulence(n, seed=2024)
    out = {}
    for prec in (_capi.GM_PRECISION_FP64, _capi.GM_PRECISION_MIXED):
        prm = _capi.make_params(n, n, n, dx, dy, dz, precision=prec)
        with _capi.Subdomain(prm) as dev:
            dev.upload_u(u0)
            t = 0.0
            for it in range(1, 51):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
                dev.tstep(dt)
                t = t + dt
            out[prec] = (dev.download_u()[:, 2:-2, 2:-2, 2:-2], float(np.abs(dev.divb()).max()))
    (a, divb_m), (b, divb_f) = out[_capi.GM_PRECISION_MIXED], out[_capi.GM_PRECISION_FP64]
    l1 = [float(np.abs(a[q] - b[q]).sum() / np.abs(b[q]).sum()) for q in range(8)]
    say(f"  rel L1 per field {['%.2e' % v for v in l1]}, max|divB| fp64 {divb_f:.2e} mixed {divb_m:.2e}")
    assert l1[0] <= 1e-5                      # the stated density bound
    assert max(l1) <= 1e-5
    s0 = u0[:, 2:-2, 2:-2, 2:-2].sum(axis=(1, 2, 3))
    for q in (0, 4):                          # mass and energy of the periodic box
        assert abs(a[q].sum() - s0[q]) / abs(s0[q]) <= 1e-9
        assert abs(b[q].sum() - s0[q]) / abs(s0[q]) <= 1e-11
    assert divb_f <= 1e-10                    # flux-CD keeps the dump formula's div B at round-off
    assert divb_m <= 1e-2                     # FP32 storage of B: round-off of single precision
