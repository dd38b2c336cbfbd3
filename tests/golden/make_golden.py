"""Write tests/golden/ref_<variant>.npz from the reference's own routines.

Runs in the build container only (needs oracle/_ref, compiled from the
reference's sources by `make -C oracle ref`).  Each fixture holds seeded
inputs and the reference's outputs for
  - per-interface solves (prim2fhlld / hlle / hll / hllc of the variant),
  - the slope limiter, u2prim, prim2u, prim2f, cfast, cfastX,
  - one boundaryI + calcprim + first-order sweep on a random state,
  - one boundaryII + calcprim + limited sweep on a random state,
  - get_timestep's ramp and clamp (the boundary fills include the
    oracle's impose_user_bc in bc_user builds, at globals::time = 0.3),
  - (ref_cooling.npz) coolingdmc and coolingchi with the reference's tables,
  - (ref_cooling_h.npz) coolingh, the H-rate cooling with the neutral
    density as the first passive scalar, under EOS_H_RATE,
  - thermal_conduction (isotropic / anisotropic, with and without
    saturation) on a static state with a random temperature field,
  - source(i,j,k,prim,s) of sources.f90 at every interior cell of a random
    primitive array (zero, 8-wave, the oracle's user gravity, or both,
    per the variant's switches),
  - (tstep_<variant>.npz, the step variants) the reference's own main-loop
    iteration, get_timestep then tstep (hydro_solver.f90:134-229), four
    times from a random state: the dt sequence, a CRC of u after every
    step and the final u and primit, ghosts included.
The fixtures are data (inputs and expected outputs); tests replay the
inputs through the C restatement and require identical bits.
"""
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import RefLib, ref_variants  # noqa: E402

NVEC = 600

# seeds of the fixtures already committed (sorted variant order of the
# first build); later variants take a seed from their name
SEEDS = {"hll_umist_closed": 1000, "hllc_superbee_pmhd": 1001, "hlld_minmod_periodic": 1002,
         "hlld_nolimit_periodic": 1003, "hlld_vanalbada_mixed": 1004, "hlld_woodward_mixed": 1005,
         "hlle_noaverage_closed": 1006, "hlle_vanleer_outflow": 1007}


def seed_of(variant):
    return SEEDS.get(variant, 2000 + zlib.crc32(variant.encode()) % 100000)


def random_prims(rng, n, npas=0):
    p = np.empty((n, 8 + npas))
    p[:, 0] = rng.uniform(0.1, 2.0, n)
    p[:, 1:4] = rng.uniform(-3.0, 3.0, (n, 3))
    p[:, 4] = rng.uniform(0.02, 2.0, n)
    p[:, 5:8] = rng.uniform(-2.0, 2.0, (n, 3))
    if npas:
        # passive densities below rho (a neutral density under EOS_H_RATE)
        p[:, 8:] = rng.uniform(0.05, 0.95, (n, npas)) * p[:, :1]
    return p


def random_state(rng, shape, cv):
    _, NX, NY, NZ = shape
    # moderate contrasts: unlimited reconstruction (LIMITER_NO_LIMIT) must
    # keep interface density and pressure positive
    rho = rng.uniform(0.7, 1.3, (NX, NY, NZ))
    v = rng.uniform(-1.5, 1.5, (3, NX, NY, NZ))
    B = rng.uniform(-1.0, 1.0, (3, NX, NY, NZ))
    p = rng.uniform(0.7, 1.3, (NX, NY, NZ))
    u = np.zeros(shape, order="F")
    u[0] = rho
    u[1:4] = rho * v
    u[4] = 0.5 * rho * (v ** 2).sum(0) + cv * p + 0.5 * (B ** 2).sum(0)
    u[5:8] = B
    if shape[0] > 8:
        u[8:] = rng.uniform(0.05, 0.95, (shape[0] - 8, NX, NY, NZ)) * rho
    return u


def make(variant, seed):
    R = RefLib(variant)
    rng = np.random.default_rng(seed)
    out = {}
    pl = random_prims(rng, NVEC, R.npas)
    pr = random_prims(rng, NVEC, R.npas)
    pr[::7, 5] = pl[::7, 5]            # continuous normal field
    pl[::11, 5] = 0.0
    pr[::11, 5] = 0.0                  # Bx = 0: single-star regions only
    pr[::13] = pl[::13]                # consistency F(U,U) = F(U)
    out["riem_pl"], out["riem_pr"] = pl, pr
    out["riem_ff"] = np.array([R.riemann(a, b) for a, b in zip(pl, pr)])
    pll = pl + rng.normal(0, 0.3, pl.shape)
    prr = pr + rng.normal(0, 0.3, pr.shape)
    pll[::5] = pl[::5]                 # flat left slope
    out["lim_in"] = np.stack([pll, pl, pr, prr])
    out["lim_out"] = np.array([np.concatenate(R.limiter(a, b, c, d)) for a, b, c, d in zip(pll, pl, pr, prr)])
    uu = np.array([np.concatenate([[a[0]], a[0] * a[1:4], [a[4] * 2.0], a[5:8], a[8:]]) for a in pl])
    uu[::17, 0] = 1e-20                # density floor
    uu[::19, 4] = -1.0                 # pressure floor
    out["u2p_in"] = uu
    out["u2p_out"] = np.array([np.concatenate([R.u2prim(a)[0], [R.u2prim(a)[1]]]) for a in uu])
    out["p2u_out"] = np.array([R.prim2u(a) for a in pl])
    out["p2f_out"] = np.array([R.prim2f(a) for a in pl])
    out["cfast_out"] = np.array([R.cfast(a[4], a[0], a[5], a[6], a[7]) for a in pl])
    out["cfastx_out"] = np.array([R.cfastx(a) for a in pl])
    # sweeps on a random state
    dx, dy, dz = 0.1, 0.07, 0.13
    R.init(dx, dy, dz)
    hook_time = 0.3 if R.bc_user else 0.0  # globals::time the user boundary hook reads
    R.set_time(hook_time)
    u = random_state(rng, R.shape(), R.cv)
    out["sweep_u_in"] = u
    R.set_array(0, u)
    R.boundary(1)
    out["sweep_u_bc1"] = R.get_array(0)
    R.calcprim(0)
    out["sweep_primit1"] = R.get_array(2)
    out["sweep_temp1"] = R.get_temp()
    R.fluxes(1)
    out["sweep_f1"], out["sweep_g1"], out["sweep_h1"] = (R.get_array(3 + a) for a in range(3))
    up = random_state(rng, R.shape(), R.cv)
    out["sweep_up_in"] = up
    R.set_array(1, up)
    R.boundary(2)
    out["sweep_up_bc2"] = R.get_array(1)
    R.calcprim(1)
    R.fluxes(2)
    out["sweep_f2"], out["sweep_g2"], out["sweep_h2"] = (R.get_array(3 + a) for a in range(3))
    ts = []
    for it in (1, 4, 10, 11, 25):
        for t, tp in ((0.0, 10.0), (0.3, 0.3000001)):
            dt, dump = R.get_timestep(it, 10, t, tp)
            ts.append([it, t, tp, dt, float(dump)])
    out["timestep"] = np.array(ts)
    if R.th_cond:
        # thermal_conduction after boundaryI + calcprim, on a static state
        # with a random temperature field (T ~ 10^5.5..10^6.5 K), at 0.5, 3
        # and 20 conduction timesteps (1, 2 and 6 super-steps)
        sh = R.shape()
        rho = rng.uniform(0.5, 2.0, sh[1:])
        temp = 10.0 ** rng.uniform(5.5, 6.5, sh[1:])
        tu = np.zeros(sh, order="F")
        tu[0] = rho
        tu[5:8] = np.array([0.3, 0.2, 0.1])[:, None, None, None] + rng.uniform(-0.1, 0.1, (3,) + sh[1:])
        tu[4] = R.cv * rho * temp / R.tempsc + 0.5 * (tu[5:8] ** 2).sum(0)
        out["tc_u_in"] = tu
        R.set_array(0, tu)
        R.boundary(1)
        R.calcprim(0)
        dtc = R.dt_cond()
        facs = (0.5, 3.0) if R.th_cond == 2 and not R.tc_saturation else (0.5, 3.0, 20.0)
        out["tc_dt"] = np.array([f * dtc for f in facs])
        for n, f in enumerate(facs):
            R.set_array(0, tu)
            R.boundary(1)
            R.calcprim(0)
            R.thermal(f * dtc)
            out[f"tc_u_{n}"] = R.get_array(0)
            out[f"tc_primit_{n}"] = R.get_array(2)
        out["tc_config"] = np.array([R.th_cond, int(R.tc_saturation), R.tempsc])
    if R.eight_wave or R.user_source_terms:
        prim = np.asfortranarray(rng.uniform(0.5, 1.5, R.shape()))
        prim[1:4] -= 1.0
        prim[5:8] -= 1.0
        R.set_array(2, prim)
        src = np.zeros((R.nx, R.ny, R.nz, 8))
        for k in range(1, R.nz + 1):
            for j in range(1, R.ny + 1):
                for i in range(1, R.nx + 1):
                    src[i - 1, j - 1, k - 1] = R.source(i, j, k, prim[:, i + 1, j + 1, k + 1])
        out["src_primit"], out["src_out"] = prim, src
    out["switches"] = np.array([int(R.eight_wave), int(R.user_source_terms), int(R.bc_user)])
    if R.eq_of_state != 1:
        out["eos_config"] = np.array([R.eq_of_state, R.tempsc])
    if R.npas:
        out["npas"] = np.array([R.npas])
    out["hook_time"] = np.array([hook_time])
    out["dxyz"] = np.array([dx, dy, dz])
    cfg = np.array([R.nx, R.ny, R.nz, R.riemann_solver, R.slope_limiter, *R.bc, int(R.mhd), int(R.pmhd)])
    out["config_int"] = cfg
    out["config_real"] = np.array([R.cv, R.gamma, R.cfl])
    return out


def make_cooling(seed=77):
    """coolingdmc / coolingchi (cooling_dmc.f90:115-156, cooling_chi.f90) of
    the reference on a random state whose temperatures span the table, the
    sqrt law above 1e8 K and the uncooled range below 1e4 K; the tables as
    the reference reads them from its own src/DMClib and src/CHIANTIlib."""
    R = RefLib("hlld_minmod_periodic")
    dx, dy, dz = 0.1, 0.07, 0.13
    R.init(dx, dy, dz)
    rng = np.random.default_rng(seed)
    sh = R.shape()
    rho = rng.uniform(0.5, 2.0, sh[1:])
    v = rng.uniform(-1.0, 1.0, (3,) + sh[1:])
    B = rng.uniform(-1.0, 1.0, (3,) + sh[1:])
    pr = rho * 10.0 ** rng.uniform(3.5, 8.5, sh[1:])
    u = np.zeros(sh, order="F")
    u[0] = rho
    u[1:4] = rho * v
    u[4] = 0.5 * rho * (v ** 2).sum(0) + R.cv * pr + 0.5 * (B ** 2).sum(0)
    u[5:8] = B
    out = {"u_in": u, "dt": np.array([1e26]), "dxyz": np.array([dx, dy, dz]),
           "config_real": np.array([R.cv, R.gamma, R.cfl])}
    for kind, name in ((3, "dmc"), (4, "chi")):
        out[f"{name}_table"] = R.cooling_table(kind)
        R.set_array(0, u)
        R.calcprim(0)
        R.cooling(kind, 1e26)
        out[f"{name}_u"] = R.get_array(0)
        out[f"{name}_primit"] = R.get_array(2)
    return out


def make_cooling_h(seed=78):
    """coolingh (cooling_h.f90:41-67: atomic and ALOSS per cell) of the
    reference under EOS_H_RATE with two passive scalars, the first the
    neutral density: temperatures 5e3..2e5 K span ALOSS's no-loss, blended
    and high-temperature branches; dt = 1e10 s (tsc = 1) so the neutral
    fraction and the temperature move."""
    R = RefLib("hlld_minmod_pas2_hrate")
    dx, dy, dz = 0.1, 0.07, 0.13
    R.init(dx, dy, dz)
    rng = np.random.default_rng(seed)
    sh = R.shape()
    rho = rng.uniform(0.5, 2.0, sh[1:])
    v = rng.uniform(-1.0, 1.0, (3,) + sh[1:])
    B = rng.uniform(-1.0, 1.0, (3,) + sh[1:])
    y0 = rng.uniform(0.02, 0.98, sh[1:])
    t = 10.0 ** rng.uniform(np.log10(5e3), np.log10(2e5), sh[1:])
    u = np.zeros(sh, order="F")
    u[0] = rho
    u[1:4] = rho * v
    u[8] = y0 * rho
    u[9] = rng.uniform(0.0, 1.0, sh[1:]) * rho
    pr = (2.0 * rho - u[8]) * t / R.tempsc
    u[4] = 0.5 * rho * (v ** 2).sum(0) + R.cv * pr + 0.5 * (B ** 2).sum(0)
    u[5:8] = B
    R.set_array(0, u)
    R.cooling(1, 1e10)
    return {"u_in": u, "u_out": R.get_array(0), "dt": np.array([1e10]), "tempsc": np.array([R.tempsc]),
            "tsc": np.array([1.0]), "dxyz": np.array([dx, dy, dz]),
            "config_real": np.array([R.cv, R.gamma, R.cfl])}


# main-loop iterations of the tstep fixtures: two in the CFL ramp
# (hydro_core.f90:677-682), two at the full CFL number
TSTEP_ITERS = (1, 2, 11, 12)


def tstep_state(R, rng):
    """A random state for the tstep fixtures; for conduction builds the
    thermal energy is raised to ~1e6 K so conduction sub-steps."""
    u = random_state(rng, R.shape(), R.cv)
    if R.th_cond:
        temp = 10.0 ** rng.uniform(5.8, 6.2, R.shape()[1:])
        u[4] = u[4] + R.cv * u[0] * temp / R.tempsc
    return u


def make_tstep(variant, seed):
    """get_timestep + tstep of the reference (step variants, flux-CD off),
    as main.f90:94-125 calls them, on a random state."""
    R = RefLib(variant)
    assert R.has_tstep and not R.enable_flux_cd
    dx, dy, dz = 0.1, 0.07, 0.13
    R.init(dx, dy, dz)
    rng = np.random.default_rng(seed)
    t = 0.3 if R.bc_user else 0.0
    R.set_time(t)
    u = tstep_state(rng=rng, R=R)
    R.set_array(0, u)
    R.boundary(1)
    R.calcprim(0)
    dts, crcs = [], []
    for it in TSTEP_ITERS:
        dt, _ = R.get_timestep(it, 10, t, 10.0)
        R.tstep(dt)
        dts.append(dt)
        crcs.append(zlib.crc32(np.ascontiguousarray(R.get_array(0)).tobytes()))
    cfg = np.array([R.nx, R.ny, R.nz, R.riemann_solver, R.slope_limiter, *R.bc, int(R.mhd), int(R.pmhd)])
    return {"u_in": u, "iters": np.array(TSTEP_ITERS), "dt": np.array(dts), "u_crc": np.array(crcs, dtype=np.uint32),
            "u_out": R.get_array(0), "primit_out": R.get_array(2), "time": np.array([t]),
            "dxyz": np.array([dx, dy, dz]), "config_int": cfg, "config_real": np.array([R.cv, R.gamma, R.cfl]),
            "switches": np.array([int(R.eight_wave), int(R.user_source_terms), int(R.bc_user)]),
            "eos_config": np.array([R.eq_of_state, R.tempsc]), "npas": np.array([R.npas]),
            "eta": np.array([R.eta]), "tc_config": np.array([R.th_cond, int(R.tc_saturation), R.tempsc])}


def main():
    vs = ref_variants()
    if not vs:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    rewrite = "--all" in sys.argv
    for v in vs:
        if v.startswith("step_"):
            path = os.path.join(HERE, f"tstep_{v[5:]}.npz")
            if rewrite or not os.path.exists(path):
                np.savez_compressed(path, **make_tstep(v, seed_of(v)))
                print("wrote", v)
            continue
        path = os.path.join(HERE, f"ref_{v}.npz")
        if os.path.exists(path) and not rewrite:
            continue
        np.savez_compressed(path, **make(v, seed_of(v)))
        print("wrote", v)
    path = os.path.join(HERE, "ref_cooling.npz")
    if rewrite or not os.path.exists(path):
        np.savez_compressed(path, **make_cooling())
        print("wrote cooling")
    path = os.path.join(HERE, "ref_cooling_h.npz")
    if rewrite or not os.path.exists(path):
        np.savez_compressed(path, **make_cooling_h())
        print("wrote cooling_h")


if __name__ == "__main__":
    main()
