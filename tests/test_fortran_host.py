"""The Fortran host (host/, amdflang over ISO_C_BINDING) keeps the
reference's problem interface: the reference's own Orszag-Tang problem
files compile against it unchanged, and its OT 128x128x4 run to t = 0.5
writes BIN dumps equal, bit for bit, to the C restatement of the
reference driven by the same time loop (src/main.f90:94-125)."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from guacho_amd import binio
from guacho_amd.host import timestep_clamp
from oracle.oracle import Config, Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_OT = "/root/reference/OT"
HOST_BIN = os.path.join(ROOT, "host", "build", "ot", "guacho_amd")
HOST_WARM_BIN = os.path.join(ROOT, "host", "build", "ot_warm", "guacho_amd")


@pytest.mark.skipif(not os.path.isdir(REF_OT), reason="reference tree not present")
def test_reference_problem_files_compile_against_host():
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "host"), f"BUILD_ROOT={tmp}",
                            f"PROBLEM={REF_OT}", "NAME=refot", f"PROBLEM_EXTRA={REF_OT}/orzag_tang.f90"],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
        assert os.path.exists(os.path.join(tmp, "refot", "guacho_amd"))


def oracle_loop(u0, cfg, tmax, dtprint, cfl, n_iter=10, itprint0=None):
    """The reference time loop on the restatement; returns {itprint: u}.
    With itprint0 it starts warm from u0 as initmain and initflow do
    (src/init.f90:124-142,432-477): time = itprint0 dtprint, the next dump
    one dtprint later, currentIteration = 1 (the CFL ramp restarts)."""
    O = Oracle(cfg)
    O.set_u(u0)
    O.boundaryI()
    O.calcprim_u()
    if itprint0 is None:
        dumps = {0: O.get_u()}
        t, tprint, itprint, it = 0.0, dtprint, 1, 1
    else:
        dumps = {}
        t = float(itprint0) * dtprint
        tprint, itprint, it = t + dtprint, itprint0 + 1, 1
    while t <= tmax:
        dt, dump = timestep_clamp(it, n_iter, cfl, O.cfl_dtp(), t, tprint)
        O.tstep(dt)
        t = t + dt
        if dump:
            dumps[itprint] = O.get_u()
            tprint = tprint + dtprint
            itprint += 1
        it += 1
    return dumps


@pytest.mark.gpu
def test_orszag_tang_128_dumps_match_restatement():
    print("case: Fortran host OT 128x128x4 HLLD minmod flux-CD CFL 0.4 to t=0.5", flush=True)
    assert os.path.exists(HOST_BIN), "build the host first (make -C host)"
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([HOST_BIN], cwd=tmp, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        files = sorted(os.listdir(os.path.join(tmp, "OT128", "BIN")))
        assert "points005.bin" in files, files
        hdr, u0 = binio.read_bin(os.path.join(tmp, "OT128", "BIN", "points000.bin"))
        dumps = {k: binio.read_bin(os.path.join(tmp, "OT128", "BIN", f"points{k:03d}.bin"))[1] for k in range(1, 7)
                 if os.path.exists(os.path.join(tmp, "OT128", "BIN", f"points{k:03d}.bin"))}
        divb5 = np.fromfile(os.path.join(tmp, "OT128", "BIN", "divB-005.bin"), dtype=np.uint8)
    assert (hdr.nx, hdr.ny, hdr.nz, hdr.neq, hdr.nghost) == (128, 128, 4, 8, 2)
    cfg = Config(nx=128, ny=128, nz=4, dx=hdr.dx, dy=hdr.dy, dz=hdr.dz, cv=1.5)
    ref = oracle_loop(u0, cfg, tmax=0.5, dtprint=0.1, cfl=0.4)
    assert sorted(ref) == [0] + sorted(dumps), (sorted(ref), sorted(dumps))
    for k, u in dumps.items():
        np.testing.assert_array_equal(u, ref[k], err_msg=f"dump {k}")
    # relative L1 per conserved variable at t = 0.5 (dump 5); rho*w and Bz stay exactly 0
    a, b = dumps[5][:, 2:-2, 2:-2, 2:-2], ref[5][:, 2:-2, 2:-2, 2:-2]
    for q in range(8):
        den = np.abs(b[q]).sum()
        if den == 0:
            assert np.all(a[q] == 0)
        else:
            assert np.abs(a[q] - b[q]).sum() / den <= 1e-12
    assert divb5.size > 0


@pytest.mark.gpu
def test_orszag_tang_warm_start_from_dump_3():
    print("case: Fortran host OT 128x128x4 cold to t=0.5, then warm start from dump 3 to t=0.5", flush=True)
    assert os.path.exists(HOST_BIN) and os.path.exists(HOST_WARM_BIN), "build the hosts first (build())"
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([HOST_BIN], cwd=tmp, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        bdir = os.path.join(tmp, "OT128", "BIN")
        cold = {k: binio.read_bin(os.path.join(bdir, f"points{k:03d}.bin"))[1] for k in range(0, 7)
                if os.path.exists(os.path.join(bdir, f"points{k:03d}.bin"))}
        for k in (4, 5, 6):
            if os.path.exists(os.path.join(bdir, f"points{k:03d}.bin")):
                os.remove(os.path.join(bdir, f"points{k:03d}.bin"))
        r = subprocess.run([HOST_WARM_BIN], cwd=tmp, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "points003.bin" in r.stdout
        hdr, _ = binio.read_bin(os.path.join(bdir, "points003.bin"))
        warm = {k: binio.read_bin(os.path.join(bdir, f"points{k:03d}.bin"))[1] for k in range(4, 7)
                if os.path.exists(os.path.join(bdir, f"points{k:03d}.bin"))}
    cfg = Config(nx=128, ny=128, nz=4, dx=hdr.dx, dy=hdr.dy, dz=hdr.dz, cv=1.5)
    ref = oracle_loop(cold[3], cfg, tmax=0.5, dtprint=0.1, cfl=0.4, itprint0=3)
    assert sorted(ref) == sorted(warm), (sorted(ref), sorted(warm))
    for k, u in warm.items():
        np.testing.assert_array_equal(u, ref[k], err_msg=f"warm dump {k}")
    # the restarted ramp takes other steps than the cold run: the states differ
    assert not np.array_equal(warm[4], cold[4])
