"""One-core cost of the C restatement against the reference's own compiled
time step (test infrastructure; needs oracle/_ref, built from the reference
tree by `make -C oracle ref`).

The reference's tstep (src/hydro_solver.f90:134-229) builds serially only
with flux-CD off (src/flux_cd_module.f90:120 leaves an `if` open without
MPI), so both sides run that configuration: HLLD, minmod, periodic, on
Orszag-Tang 128x128x4 and on the seeded turbulence box at 64^3.  Each side
runs the reference's main-loop iteration (get_timestep, then tstep) from
the same initial state, single-threaded, after one untimed iteration; the
two end states must agree bit for bit.  Prints a markdown table of the cost
per zone-update and the ratio.

    python3 oracle/calibrate.py [--steps N]   (from the repository root)
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[0] = ROOT  # not oracle/, whose oracle.py would shadow the package

from guacho_amd import problems  # noqa: E402
from oracle.oracle import Oracle, RefLib, timestep_finish  # noqa: E402


def run_ref(R, u0, dx, dy, dz, steps):
    R.init(dx, dy, dz)
    R.set_array(0, u0)
    R.boundary(1)
    R.calcprim(0)
    t, it = 0.0, 1
    dt, _ = R.get_timestep(it, 10, t, 1e30)
    R.tstep(dt)
    t += dt
    t0 = time.perf_counter()
    for it in range(2, steps + 2):
        dt, _ = R.get_timestep(it, 10, t, 1e30)
        R.tstep(dt)
        t += dt
    el = time.perf_counter() - t0
    return el, R.get_array(0)


def run_port(cfg, u0, steps):
    O = Oracle(cfg, nthreads=1)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    t = 0.0
    dt, _ = timestep_finish(1, 10, cfg.cfl, O.cfl_dtp(), t, 1e30)
    O.tstep(dt)
    t += dt
    t0 = time.perf_counter()
    for it in range(2, steps + 2):
        dt, _ = timestep_finish(it, 10, cfg.cfl, O.cfl_dtp(), t, 1e30)
        O.tstep(dt)
        t += dt
    el = time.perf_counter() - t0
    u = O.get_u()
    O.close()
    return el, u


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20)
    a = ap.parse_args()
    cases = (
        ("Orszag-Tang 128x128x4", "calib_ot128", lambda: problems.orszag_tang(128, 128, 4)),
        ("turbulence 64^3 (seed 2024)", "calib_turb64", lambda: problems.turbulence(64, seed=2024)),
    )
    print("| workload | steps | reference (amdflang -O2) us/zone-update | restatement (gcc -O2, 1 thread) "
          "us/zone-update | restatement / reference | end states bitwise equal |")
    print("|---|---|---|---|---|---|")
    for name, variant, make in cases:
        u0, dx, dy, dz = make()
        R = RefLib(variant)
        cfg = R.config(dx, dy, dz)
        zones = cfg.nx * cfg.ny * cfg.nz * a.steps
        tr, ur = run_ref(R, u0, dx, dy, dz, a.steps)
        tp, up = run_port(cfg, u0, a.steps)
        same = bool(np.array_equal(ur, up))
        print(f"| {name} | {a.steps} | {1e6 * tr / zones:.3f} | {1e6 * tp / zones:.3f} | {tp / tr:.3f} | {same} |",
              flush=True)


if __name__ == "__main__":
    main()
