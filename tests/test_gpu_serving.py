"""Continuous-batching server (guacho_amd.serving): requests that arrive
while others run are admitted as memory frees up, and every result is
bitwise the same as the same request solved alone."""
import sys

import numpy as np
import pytest

from guacho_amd import _capi, problems, serving

pytestmark = pytest.mark.gpu


def small_stream(seed=3):
    rng = np.random.default_rng(seed)
    mp = _capi.GM_PRECISION_MIXED
    reqs = []
    t = 0.0
    for i in range(9):
        t += float(rng.exponential(0.01))
        k = i % 3
        if k == 0:
            u, dx, dy, dz = problems.brio_wu(128, 4, 4)
            p = _capi.make_params(128, 4, 4, dx, dy, dz, cv=1.0, bc=(1, 1, 3, 3, 3, 3), precision=mp)
            r = serving.Request(i, "briowu", p, u, t_end=0.02)
        elif k == 1:
            u, dx, dy, dz = problems.orszag_tang(32, 32, 4)
            p = _capi.make_params(32, 32, 4, dx, dy, dz, precision=mp)
            r = serving.Request(i, "ot", p, u, t_end=0.01)
        else:
            n = int(rng.integers(12, 25))
            u, dx, dy, dz = problems.blast(n, seed=i)
            p = _capi.make_params(n, n, n, dx, dy, dz, bc=(1,) * 6, precision=mp)
            r = serving.Request(i, "blast", p, u, max_steps=int(rng.integers(5, 15)))
        r.arrival = t
        reqs.append(r)
    return reqs


def test_results_same_alone_and_in_the_mix():
    print("case: 9 requests (Brio-Wu, OT, blasts) through the server vs each alone", flush=True)
    reqs = small_stream()
    # a budget that holds only a few boxes at once forces queueing on memory
    with _capi.Batch(precision=_capi.GM_PRECISION_MIXED, max_boxes=4) as B:
        per = max(B.box_bytes(r.params) for r in reqs)
    stats = serving.serve(reqs, steps_per_launch=3, memory_budget=3 * per, keep_results=True)
    assert stats["requests"] == 9 and all(r.finished >= r.arrival for r in reqs)
    for r in reqs:
        with _capi.Batch(precision=_capi.GM_PRECISION_MIXED, max_boxes=1, steps_per_launch=1) as B:
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
    sys.stdout.flush()
