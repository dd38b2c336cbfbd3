"""Benchmark of the Godunov MHD step on one MI355X.

Workload: the seeded 3D MHD turbulence box, 256^3 periodic, plain FP64,
HLLD with minmod reconstruction and flux-CD, CFL 0.4 -- 20 steps from the
seeded initial state.  A "step" is one iteration of the reference's time
loop (src/main.f90:94-125): the CFL reduction (get_timestep) and one tstep,
state resident in HBM.

Prints one JSON line: zone-updates/s over the timed steps, the HBM roofline
of the dominant kernel class from HIP events on the library's stream, the
step-level roofline at the 128 B/zone-update compulsory traffic, the
mixed-precision form on the 512^3 box against the FP64 path (speed, per-field
error, drift, div B), the optional terms (8-wave sources, user plugins), the
batch and serving forms, and the C restatement of the reference's CPU path
timed on a bounded sample on the host cores.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
BYTES_PER_ZONE_FP64 = 128      # read 8 + write 8 conserved doubles per zone-update


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="cells per side of the turbulence box")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision section")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-boxes section")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving section")
    ap.add_argument("--no-ot", action="store_true", help="skip the Orszag-Tang 128x128x4 section")
    ap.add_argument("--no-extensions", action="store_true",
                    help="skip the 8-wave / user-plugin section")
    ap.add_argument("--serving-requests", type=int, default=24)
    ap.add_argument("--batch-boxes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--mixed-n", type=int, default=512)
    ap.add_argument("--mixed-steps", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def host_physical_cores():
    """Physical cores of the host (distinct core, socket pairs of lscpu -p)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=Core,Socket"], capture_output=True, text=True, timeout=10).stdout
        pairs = {ln.strip() for ln in out.splitlines() if ln and not ln.startswith("#")}
        return len(pairs) or None
    except (OSError, subprocess.SubprocessError):
        return None


def cpu_baseline(u0, dx, dy, dz, n: int, seconds: float):
    """The C restatement of the reference's CPU path (oracle/) on the host
    cores, on the headline workload itself (the seeded turbulence box, FP64
    HLLD + minmod + flux-CD): one untimed step, then steps until `seconds`
    of wall time have passed (at most 200).  Threads: the host share this
    process is given (OMP_NUM_THREADS, which the GPU host sets to its CPU
    share per GPU), else every physical core."""
    from guacho_amd.host import timestep_clamp
    from oracle.oracle import Config, Oracle

    phys = host_physical_cores()
    env_threads = os.environ.get("OMP_NUM_THREADS")
    threads = int(env_threads) if env_threads else (phys or os.cpu_count() or 1)
    O = Oracle(Config(nx=n, ny=n, nz=n, dx=dx, dy=dy, dz=dz), nthreads=threads)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    t = 0.0
    dt, _ = timestep_clamp(1, 10, 0.4, O.cfl_dtp(), t, 1e30)
    O.tstep(dt)
    t += dt
    t0 = time.perf_counter()
    steps = 0
    while steps < 200:
        dt, _ = timestep_clamp(steps + 2, 10, 0.4, O.cfl_dtp(), t, 1e30)
        O.tstep(dt)
        t += dt
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    O.close()
    value = n ** 3 * steps / el
    out = {"value": value, "unit": "zone-updates/s", "cores": threads, "kind": "port",
           "host_physical_cores": phys,
           "threads_from": "OMP_NUM_THREADS (the host's CPU share per GPU)" if env_threads else "lscpu physical cores",
           "sample": f"C restatement (oracle/guacho_oracle.c, gcc -O3, OpenMP {threads} threads) of the headline "
                     f"workload, turbulence {n}^3 FP64 HLLD minmod flux-CD: {steps} steps in {el:.2f} s after one "
                     f"untimed step"}
    try:
        with open(os.path.join(ROOT, "profiles", "cpu_calibration.json")) as fh:
            cal = json.load(fh)
        r = float(cal["ratio_3d"])
        out["calibration"] = {"restatement_over_reference_1core_3d": r, "source": cal["source"]}
        out["reference_equivalent_value"] = value * r
    except (OSError, KeyError, ValueError):
        pass
    return out


# Compulsory bytes one launch of each kernel class moves (FP64, flux-CD),
# per interior cell: every array the kernel must read or write, once.
def algorithmic_bytes(cls: str, n: int) -> float:
    N = float(n) ** 3
    return {
        # read U^n (8); write the x/y partials (5 hydro + 2 E) and E3
        "xy_predictor": 128.0 * N,
        # read U* (8) and the hydro part of U^n (5); write partials (7) and E3
        "xy_corrector": 168.0 * N,
        # read U (8) and the partials (7); write the hydro update (5), E1, E2
        "z_predictor": 176.0 * N,
        "z_corrector": 176.0 * N,
        # read B (3) and E (3); write B (3)
        "bfield": 72.0 * N,
        # the corrector's field update plus the CFL candidates of the new
        # state: also read its hydro part (5)
        "bfield_cfl": 112.0 * N,
        # read U (8)
        "cfl_dt": 64.0 * N,
        # the 2.5-D stage kernel (xy and z faces in one pass): read U^n (8);
        # write U*'s hydro part (5) and E (3)
        "slab_predictor": 128.0 * N,
        # read U* (8) and U^n's hydro part (5); write the hydro part (5) and E (3)
        "slab_corrector": 168.0 * N,
    }.get(cls, 0.0)


def measured_profile(cls: str):
    """Calibrated HBM bytes and FP64 VALU instructions per launch of `cls`
    from the committed rocprofv3 summary of this workload (FETCH_SIZE,
    WRITE_SIZE, SQ_INSTS_VALU), if one exists."""
    path = os.path.join(ROOT, "profiles", "traffic_fp64_turb256.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        return (d["per_launch_bytes"].get(cls), d.get("valu_instr_per_launch", {}).get(cls),
                os.path.relpath(path, ROOT))
    except (OSError, KeyError, ValueError):
        return None, None, None


N_SIMD = 1024        # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9     # MI355X peak engine clock
VALU_CYCLES = 4      # a wave64 FP64 VALU instruction occupies a 16-lane SIMD 4 cycles


def dominant_roofline(kt, n, steps):
    """Roofline of the kernel class with the largest total time.  `achieved`
    is the path's algorithmic bytes, 128 B per zone-update (read and write
    the 8 conserved doubles), times the zones one launch advances, over the
    kernel's average launch time from HIP events on the library's stream.
    The kernel is bounded by FP64 VALU issue, not HBM: `valu_issue_frac`
    is its VALU instructions per launch (a property of the kernel and the
    workload, from the committed SQ counters) x 4 cycles over the live
    launch time x 1024 SIMDs x 2.4 GHz."""
    cls = max((k for k in kt if kt[k][1] and algorithmic_bytes(k, n) > 0), key=lambda k: kt[k][0])
    ms, launches = kt[cls]
    avg_s = ms / launches / 1e3
    zone_bytes = float(BYTES_PER_ZONE_FP64) * n ** 3
    ach = zone_bytes / avg_s / 1e9
    io = algorithmic_bytes(cls, n)
    traffic, valu_instr, src = measured_profile(cls)
    valu = round(valu_instr * VALU_CYCLES / (avg_s * N_SIMD * CLOCK_HZ), 3) if valu_instr else None
    return {"bound": "valu", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": cls,
            "frac_basis": f"{BYTES_PER_ZONE_FP64} B per zone-update x {n}^3 zones per launch",
            "launches_per_step": launches // steps, "avg_launch_ms": round(ms / launches, 4),
            "algorithmic_bytes_per_launch": zone_bytes,
            # the kernel's own compulsory I/O (every array it reads or writes, once)
            "kernel_io_bytes_per_launch": io, "kernel_io_achieved": round(io / avg_s / 1e9, 1),
            "kernel_io_frac": round(io / avg_s / 1e9 / HBM_PEAK_GBS, 4),
            "valu_issue_frac": valu, "valu_instr_per_launch": valu_instr,
            "traffic_source": src}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_steps(dev, u0, steps, warmup, cfl=0.4):
    """warm-up, re-upload the seeded state, then `steps` timed steps; returns seconds."""
    from guacho_amd.host import timestep_clamp
    dev.upload_u(u0)
    dev.boundary_u()
    t = 0.0
    for it in range(1, warmup + 1):
        dt, _ = timestep_clamp(it, 10, cfl, dev.cfl_dtp(), t, 1e30)
        dev.tstep(dt)
        t += dt
    dev.sync()
    dev.upload_u(u0)
    dev.boundary_u()
    dev.sync()
    t = 0.0
    t0 = time.perf_counter()
    for it in range(1, steps + 1):
        dt, _ = timestep_clamp(it, 10, cfl, dev.cfl_dtp(), t, 1e30)
        dev.tstep(dt)
        t += dt
    dev.sync()
    return time.perf_counter() - t0


def mixed_section(n: int, steps: int, seed: int):
    """The mixed-precision form on the seeded turbulence box (FP32 states,
    reconstruction and fluxes; FP64 pressure recovery, update accumulation
    and reductions) against the FP64 path: zone-updates/s of both, per-field
    relative L1 error, mass and energy drift, max |div B|."""
    import numpy as np
    from guacho_amd import _capi, problems
    log(f"mixed section: generating turbulence {n}^3")
    u0, dx, dy, dz = problems.turbulence(n, seed=seed)
    res = {}
    for name, prec in (("fp64", _capi.GM_PRECISION_FP64), ("mixed", _capi.GM_PRECISION_MIXED)):
        log(f"mixed section: {name} {steps} steps")
        prm = _capi.make_params(n, n, n, dx, dy, dz, precision=prec)
        with _capi.Subdomain(prm) as dev:
            el = run_steps(dev, u0, steps, warmup=2)
            res[name] = {"value": n ** 3 * steps / el, "ms_per_step": 1e3 * el / steps,
                         "u": dev.download_u(), "divb": float(np.abs(dev.divb()).max())}
    a = res["mixed"]["u"][:, 2:-2, 2:-2, 2:-2]
    b = res["fp64"]["u"][:, 2:-2, 2:-2, 2:-2]
    names = ["rho", "mx", "my", "mz", "E", "Bx", "By", "Bz"]
    l1 = {names[q]: float(np.abs(a[q] - b[q]).sum() / np.abs(b[q]).sum()) for q in range(8)}
    s0 = u0[:, 2:-2, 2:-2, 2:-2].sum(axis=(1, 2, 3))
    drift = {k: {"mass": float(abs(res[k]["u"][0, 2:-2, 2:-2, 2:-2].sum() - s0[0]) / abs(s0[0])),
                 "energy": float(abs(res[k]["u"][4, 2:-2, 2:-2, 2:-2].sum() - s0[4]) / abs(s0[4]))}
             for k in ("fp64", "mixed")}
    return {
        "workload": f"turbulence-{n}^3-hlld-minmod-fluxcd-cfl0.4, {steps} steps",
        "value": round(res["mixed"]["value"], 1), "unit": "zone-updates/s",
        "ms_per_step": round(res["mixed"]["ms_per_step"], 4),
        "fp64_value": round(res["fp64"]["value"], 1), "fp64_ms_per_step": round(res["fp64"]["ms_per_step"], 4),
        "speedup_vs_fp64": round(res["mixed"]["value"] / res["fp64"]["value"], 3),
        "rel_l1_vs_fp64": l1, "drift": drift,
        "max_abs_divb": {"fp64": res["fp64"]["divb"], "mixed": res["mixed"]["divb"]},
        "bytes_per_zone_update": 64,
        "step_roofline_frac": round(64.0 * n ** 3 / (res["mixed"]["ms_per_step"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
    }


def batch_section(nbox: int, steps: int, seed: int):
    """64 independent MHD blasts of uneven cubic size n ~ U{32..128}, outflow,
    HLLD, `steps` steps each: all boxes in one batch whose steps are captured
    in one HIP graph launch, launch shape autotuned, in mixed precision (the
    configuration) and in FP64; against the same boxes solved one at a time
    with the single-box path in both precisions.  Every leg's timed region
    starts with the boxes uploaded and ends with their last step done: no
    leg times an upload or allocation."""
    import numpy as np
    from guacho_amd import _capi, problems
    from guacho_amd.host import timestep_clamp
    rng = np.random.default_rng(seed)
    sizes = rng.integers(32, 129, size=nbox)
    log(f"batch section: {nbox} blasts, sizes {sizes.min()}..{sizes.max()}, {steps} steps")
    boxes = []
    for b, n in enumerate(sizes):
        u, dx, dy, dz = problems.blast(int(n), seed=seed * 1000 + b)
        boxes.append((int(n), u, dx))
    zones = float(sum(n ** 3 for n, _, _ in boxes)) * steps

    def batched(prec):
        def admit(B):
            return [B.add(_capi.make_params(n, n, n, h, h, h, bc=(1,) * 6, precision=prec), u, max_steps=steps)
                    for n, u, h in boxes]
        with _capi.Batch(precision=prec, max_boxes=nbox, steps_per_launch=steps, autotune=True) as B:
            ids = admit(B)
            B.advance(1)        # first use: autotune (dt = 0 steps), tables, graph capture, the run
            B.sync()
            block, kchunk, slab, tuned_ms = B.tuning()
            u_first = B.download_u(ids[0])
            for i in ids:
                B.remove(i)
            ids = admit(B)      # same boxes again: same geometry, the captured graph is reused
            B.sync()
            t0 = time.perf_counter()
            B.advance(1)        # all boxes x `steps` steps in one graph launch
            B.sync()
            el = time.perf_counter() - t0
            assert all(B.clock(i) == (B.clock(i)[0], steps, True) for i in ids)
            same = bool(np.array_equal(B.download_u(ids[0]), u_first))
        return el, (block, kchunk, slab, tuned_ms), same

    def sequential(prec):
        devs = []
        try:
            for n, u, h in boxes:
                d = _capi.Subdomain(_capi.make_params(n, n, n, h, h, h, bc=(1,) * 6, precision=prec))
                devs.append(d)
                d.upload_u(u)
                d.sync()
            t0 = time.perf_counter()
            for d in devs:
                t = 0.0
                for it in range(1, steps + 1):
                    dt, _ = timestep_clamp(it, 10, 0.4, d.cfl_dtp(), t, 1e30)
                    d.tstep(dt)
                    t += dt
                d.sync()
            return time.perf_counter() - t0
        finally:
            for d in devs:
                d.close()

    mp, fp = _capi.GM_PRECISION_MIXED, _capi.GM_PRECISION_FP64
    log("batch section: batched, mixed")
    el_m, tune_m, same_m = batched(mp)
    log("batch section: batched, FP64")
    el_f, tune_f, same_f = batched(fp)
    log("batch section: one box at a time, mixed")
    seq_m = sequential(mp)
    log("batch section: one box at a time, FP64")
    seq_f = sequential(fp)
    return {"workload": f"{nbox} MHD blasts n~U{{32..128}}^3, outflow, HLLD, mixed, {steps} steps, one graph launch",
            "value": round(zones / el_m, 1), "unit": "zone-updates/s", "seconds": round(el_m, 4),
            "zone_updates": zones,
            "tuned_shape": {"block": tune_m[0], "kchunk": tune_m[1], "slab": tune_m[2],
                            "ms_per_step": round(tune_m[3], 4)},
            "tuned_shape_fp64": {"block": tune_f[0], "kchunk": tune_f[1], "slab": tune_f[2],
                                 "ms_per_step": round(tune_f[3], 4)},
            "batched_fp64_value": round(zones / el_f, 1), "batched_fp64_seconds": round(el_f, 4),
            "sequential_mixed_value": round(zones / seq_m, 1), "sequential_mixed_seconds": round(seq_m, 4),
            "sequential_fp64_value": round(zones / seq_f, 1), "sequential_fp64_seconds": round(seq_f, 4),
            # batching's own gain, precision held fixed
            "batching_speedup_mixed": round(seq_m / el_m, 3), "batching_speedup_fp64": round(seq_f / el_f, 3),
            "speedup_vs_sequential_fp64": round(seq_f / el_m, 3),
            "timed_region": "boxes resident in HBM; last step done (sync); no uploads or allocations in any leg",
            "repeat_bitwise_equal": same_m and same_f}


def ot_section(cpu_threads: int):
    """Orszag-Tang 128x128x4 to t = 0.5 (HLLD, minmod, flux-CD, CFL 0.4,
    dumps every 0.1 as the reference's driver takes them; the dumps are
    not written): the reference's host time loop over the single-box path
    (one CFL read-back and a handful of launches per step: launch-latency
    bound at 65,536 cells), the same box as a one-box batch whose clock
    (ramp and t_end clamp) runs on the device, 64 steps per graph launch
    (its only clamp is onto t = 0.5, so its step count can differ from the
    host loop's by the dump clamps; autotuned on a first run, the second
    timed with the tuned shape and captured graph), and the C restatement
    on the host cores with the host loop.  FP64 throughout."""
    import numpy as np
    from guacho_amd import _capi, problems
    from guacho_amd.host import timestep_clamp
    from oracle.oracle import Config, Oracle
    n, nz, tmax, dtprint = 128, 4, 0.5, 0.1
    u0, dx, dy, dz = problems.orszag_tang(n, n, nz)
    prm = _capi.make_params(n, n, nz, dx, dy, dz)

    def host_loop(step_fn, dtp_fn):
        t, tprint, it = 0.0, dtprint, 1
        while t <= tmax:
            dt, dump = timestep_clamp(it, 10, 0.4, dtp_fn(), t, tprint)
            step_fn(dt)
            t = t + dt
            if dump:
                tprint = tprint + dtprint
            it += 1
        return it - 1
    log("OT 128x128x4 section: host loop, single-box path")
    with _capi.Subdomain(prm) as dev:
        dev.upload_u(u0)
        host_loop(dev.tstep, dev.cfl_dtp)          # warm-up run
        dev.upload_u(u0)
        dev.sync()
        t0 = time.perf_counter()
        steps = host_loop(dev.tstep, dev.cfl_dtp)
        dev.sync()
        el_host = time.perf_counter() - t0
        u_host = dev.download_u()
    log("OT 128x128x4 section: one-box batch, device clock, 64 steps per graph launch")
    el_b = None
    # the first run autotunes the stage kernels and captures the graph; the
    # second, timed, reuses both (same box and tile counts)
    with _capi.Batch(precision=_capi.GM_PRECISION_FP64, max_boxes=1, steps_per_launch=64, autotune=True) as B:
        for rep in range(2):
            box = B.add(prm, u0, cfl=0.4, t_end=tmax)
            t0 = time.perf_counter()
            while True:
                B.advance(1)
                B.sync()
                _, bsteps, done = B.clock(box)
                if done:
                    break
            el_b = time.perf_counter() - t0
            B.remove(box)
        ot_shape = B.tuning()
    log(f"OT 128x128x4 section: C restatement, {cpu_threads} threads")
    O = Oracle(Config(nx=n, ny=n, nz=nz, dx=dx, dy=dy, dz=dz), nthreads=cpu_threads)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    t0 = time.perf_counter()
    csteps = host_loop(O.tstep, O.cfl_dtp)
    el_cpu = time.perf_counter() - t0
    same = bool(np.array_equal(u_host[:, 2:-2, 2:-2, 2:-2], O.get_u()[:, 2:-2, 2:-2, 2:-2]))
    O.close()
    zones = float(n * n * nz)
    return {"workload": "Orszag-Tang 128x128x4 to t=0.5, FP64 HLLD minmod flux-CD, CFL 0.4", "unit": "zone-updates/s",
            "steps": steps, "host_loop": {"value": round(zones * steps / el_host, 1),
                                          "ms_per_step": round(1e3 * el_host / steps, 4), "seconds": round(el_host, 4)},
            "batch_graph": {"value": round(zones * bsteps / el_b, 1), "ms_per_step": round(1e3 * el_b / bsteps, 4),
                            "seconds": round(el_b, 4), "steps": bsteps, "steps_per_launch": 64,
                            "tuned_shape": {"block": ot_shape[0], "kchunk": ot_shape[1], "slab": ot_shape[2]}},
            "cpu_restatement": {"value": round(zones * csteps / el_cpu, 1), "seconds": round(el_cpu, 3),
                                "threads": cpu_threads, "steps": csteps},
            "gpu_equals_restatement_at_end": same}


def extensions_section(u0, dx, dy, dz, n: int, steps: int):
    """The step with the reference's optional terms, FP64 on the seeded
    turbulence box, against the same box without them: Powell 8-wave
    source terms (no flux-CD, as the 8-wave scheme is run), the user
    plugins -- a source-term kernel (uniform + point-mass gravity) and a
    boundary hook (inflow jet + interior obstacle; ghosts then live in
    memory: boundaryII + hook between the stages, boundaryI + hook after)
    -- on outflow faces, DMC table cooling (T ~ 1e6 K), and isotropic
    thermal conduction (one sub-step per step at these scalings).  Per
    variant: zone-updates/s and the per-class kernel times."""
    from guacho_amd import _capi
    from guacho_amd.host import timestep_clamp
    plugins = os.path.join(ROOT, "tests", "plugins")
    gravity = (os.path.join(plugins, "gravity.hsaco"), "gm_test_gravity")
    jet = (os.path.join(plugins, "jet.hsaco"), "gm_test_jet")
    import numpy as np
    dmc = np.load(os.path.join(ROOT, "tests", "golden", "ref_cooling.npz"))["dmc_table"]
    variants = (
        ("no_flux_cd", dict(enable_flux_cd=False), None, None),
        ("eight_wave", dict(enable_flux_cd=False, eight_wave=True), None, None),
        ("outflow", dict(bc=(1,) * 6), None, None),
        ("outflow_user_plugins", dict(bc=(1,) * 6), gravity, jet),
        ("cooling_dmc", dict(), None, None),
        ("conduction_iso", dict(), None, None),
    )
    res = {}
    for name, kw, src, hook in variants:
        log(f"extensions section: {name}, {steps} steps")
        prm = _capi.make_params(n, n, n, dx, dy, dz, **kw)
        with _capi.Subdomain(prm) as dev:
            if src:
                dev.set_user_sources(*src)
            if hook:
                dev.set_user_bc(*hook)
            if name == "cooling_dmc":
                dev.set_cooling(_capi.GM_COOL_DMC, dmc, 1e6, 1e-22, 10.0)
            if name == "conduction_iso":
                dev.set_conduction(_capi.GM_TC_ISOTROPIC)
            run_steps(dev, u0, 2, warmup=0)
            dev.upload_u(u0)
            dev.boundary_u()
            dev.sync()
            dev.kernel_times()
            dev.set_timing(True)
            t = 0.0
            t0 = time.perf_counter()
            for it in range(1, steps + 1):
                dt, _ = timestep_clamp(it, 10, 0.4, dev.cfl_dtp(), t, 1e30)
                dev.set_time(t)
                dev.tstep(dt)
                t += dt
            dev.sync()
            el = time.perf_counter() - t0
            kt = dev.kernel_times()
        res[name] = {"value": round(n ** 3 * steps / el, 1), "ms_per_step": round(1e3 * el / steps, 4),
                     "kernel_ms_per_step": {k: round(v[0] / steps, 4) for k, v in kt.items() if v[1]}}
    src_ms = res["eight_wave"]["kernel_ms_per_step"].get("sources", 0.0)
    # the source pass per launch: read the stage input (8) and the updated state (8), write it (8)
    src_gbs = 192.0 * n ** 3 / (src_ms / 2 / 1e3) / 1e9 if src_ms else None
    return {"workload": f"turbulence-{n}^3-fp64-hlld-minmod, {steps} steps per variant", "unit": "zone-updates/s",
            "variants": res,
            "eight_wave_overhead": round(res["eight_wave"]["ms_per_step"] / res["no_flux_cd"]["ms_per_step"] - 1, 4),
            "sources_pass_achieved_GBs": round(src_gbs, 1) if src_gbs else None,
            "user_plugins_overhead": round(res["outflow_user_plugins"]["ms_per_step"]
                                           / res["outflow"]["ms_per_step"] - 1, 4),
            "cooling_ms_per_step": res["cooling_dmc"]["kernel_ms_per_step"].get("cooling"),
            "conduction_ms_per_step": res["conduction_iso"]["kernel_ms_per_step"].get("conduction")}


def serving_section(nreq: int, seed: int, fractions=(0.25, 0.5, 1.0, 2.0, 4.0)):
    """Continuous batching of a Poisson stream of mixed solves (Brio-Wu
    1024x4x4, Orszag-Tang 256^2, blasts 64^3-192^3) against serving the same
    stream one request at a time with the plain FP64 path; arrival rate
    swept as fractions of the sequential FP64 service rate (at 0.25x to 1x
    arrivals bound both; from 2x on, and saturated, the GPU does)."""
    from guacho_amd import serving
    log(f"serving section: {nreq} requests, sequential FP64 service rate")
    base = serving.make_stream(nreq, rate=1e9, seed=seed)        # all present at t ~ 0
    seq0 = serving.serve_sequential_fp64(base)
    mu = nreq / seq0["makespan_s"]
    sweep = []
    for f in list(fractions) + [None]:
        lam = f * mu if f is not None else 1e9          # None: saturated, every request present at t ~ 0
        log(f"serving section: rate {'saturated' if f is None else f'{f:.2f} x sequential'} ({lam:.2f}/s)")
        cb = serving.serve(serving.make_stream(nreq, rate=lam, seed=seed + 1), steps_per_launch=4)
        sq = serving.serve_sequential_fp64(serving.make_stream(nreq, rate=lam, seed=seed + 1))
        sweep.append({"rate_fraction": f if f is not None else "saturated", "rate_per_s": round(lam, 3),
                      "batched_mixed": {k: round(v, 4) for k, v in cb.items()},
                      "sequential_fp64": {k: round(v, 4) for k, v in sq.items()},
                      "throughput_ratio": round(cb["zone_updates_per_s"] / sq["zone_updates_per_s"], 3)})
    return {"workload": f"{nreq} Poisson requests: Brio-Wu 1024x4x4 to t=0.1, OT 256^2x4 to t=0.05, "
                        f"blasts n^3 n~U{{64..192}} 40 steps; HLLD minmod flux-CD",
            "sequential_fp64_service_rate_per_s": round(mu, 3),
            "sequential_fp64_zone_updates_per_s": round(seq0["zone_updates_per_s"], 1),
            "sweep": sweep}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library runs one subdomain per GPU and no "
              f"multi-GPU path exists (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    from guacho_amd import _capi, problems
    from guacho_amd.host import timestep_clamp

    n = args.n
    u0, dx, dy, dz = problems.turbulence(n, seed=args.seed)
    prm = _capi.make_params(n, n, n, dx, dy, dz)
    cfl = 0.4
    log(f"FP64 turbulence {n}^3: {args.warmup} warm-up + {args.steps} timed steps")
    with _capi.Subdomain(prm) as dev:
        dev.kernel_times()                      # reset counters
        # warm-up runs untimed inside run_steps; kernel timing covers the timed steps only
        dev.upload_u(u0)
        dev.boundary_u()
        t = 0.0
        for it in range(1, args.warmup + 1):
            dt, _ = timestep_clamp(it, 10, cfl, dev.cfl_dtp(), t, 1e30)
            dev.tstep(dt)
            t += dt
        dev.sync()
        dev.upload_u(u0)
        dev.boundary_u()
        dev.sync()
        dev.kernel_times()
        dev.set_timing(True)
        t = 0.0
        t0 = time.perf_counter()
        for it in range(1, args.steps + 1):
            dt, _ = timestep_clamp(it, 10, cfl, dev.cfl_dtp(), t, 1e30)
            dev.tstep(dt)
            t += dt
        dev.sync()
        el = time.perf_counter() - t0
        kt = dev.kernel_times()
        dev.set_timing(False)

    zones = n ** 3 * args.steps
    value = zones / el
    ms_per_step = 1e3 * el / args.steps
    roof = dominant_roofline(kt, n, args.steps)
    step_ach = BYTES_PER_ZONE_FP64 * n ** 3 / (el / args.steps) / 1e9
    out = {
        "metric": "zone-updates/s", "value": round(value, 1), "unit": "zone-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"turbulence-{n}^3-fp64-hlld-minmod-fluxcd-cfl0.4", "grid": [n, n, n],
                   "solver": "HLLD", "limiter": "minmod", "flux_cd": True, "precision": "fp64",
                   "boundaries": "periodic", "seed": args.seed},
        "roofline": roof,
        "step_roofline": {"bytes_per_zone_update": BYTES_PER_ZONE_FP64, "achieved": round(step_ach, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_ach / HBM_PEAK_GBS, 4)},
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kt.items() if v[1]},
    }
    if not args.no_ot:
        out["ot128"] = ot_section(int(os.environ.get("OMP_NUM_THREADS") or host_physical_cores() or 1))
    if not args.no_extensions:
        out["extensions"] = extensions_section(u0, dx, dy, dz, n, 10)
    if not args.no_serving:
        out["serving"] = serving_section(args.serving_requests, args.seed)
    if not args.no_batch:
        out["batch"] = batch_section(args.batch_boxes, args.batch_steps, args.seed)
    if not args.no_mixed:
        out["mixed"] = mixed_section(args.mixed_n, args.mixed_steps, args.seed)
    if not args.no_cpu_baseline:
        log("CPU baseline (C restatement on the host cores)")
        out["cpu_baseline"] = cpu_baseline(u0, dx, dy, dz, n, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
