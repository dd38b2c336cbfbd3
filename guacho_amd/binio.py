"""BIN dump format of write_BIN (src/Out_BIN_Module.f90:40-248), the
reference's output and warm-start wire format, written and read on the host.

Layout of points<rank>.<itprint>.bin (stream access, native endianness):
  ASCII header lines, each ending in LF, then achar(255), LF
  'd' (8-byte reals)
  int32 nx, ny, nz
  f64   dx, dy, dz
  int32 block origin (coords*n) x3, MPI blocks x3
  int32 neq, neqdyn
  int32 nghost
  f64   rsc, vsc, rhosc
  f64   cv
  f64   u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax), column-major
The warm-start reader (src/init.f90:451-467) skips to achar(255) and reads
the rest; the ASCII part is informational.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np


@dataclass
class BinHeader:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    origin: tuple = (0, 0, 0)
    mpi_blocks: tuple = (1, 1, 1)
    neq: int = 8
    neqdyn: int = 8
    nghost: int = 2
    rsc: float = 1.0
    vsc: float = 1.0
    rhosc: float = 1.0
    cv: float = 1.5


def _es(x: float, w: int = 10, d: int = 3) -> str:
    return f"{x:.{d}E}".rjust(w)


def write_bin(path: str, hdr: BinHeader, u: np.ndarray) -> None:
    u = np.asfortranarray(u, dtype=np.float64)
    lines = [
        "**************** Output for Guacho v1.3****************",
        f"Dimensions    : {hdr.nx} {hdr.ny} {hdr.nz}",
        "Spacings      : " + "".join(_es(v) for v in (hdr.dx, hdr.dy, hdr.dz)),
        "Block Origin, cells    : " + " ".join(str(v) for v in hdr.origin),
        "MPI blocks (X, Y, Z)   : " + " ".join(str(v) for v in hdr.mpi_blocks),
        f"Number of Equations/dynamical ones  {hdr.neq}/{hdr.neqdyn}",
        f"Number of Ghost Cells  {hdr.nghost}",
        "Scalings",
        f"r_sc: {_es(hdr.rsc)} v_sc: {_es(hdr.vsc)} rho_sc: {_es(hdr.rhosc)}",
        f"Specfic heat at constant volume Cv: {hdr.cv:7.2f}",
        "Double precision 8 byte floats",
        "*******************************************************",
    ]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as fh:
        for ln in lines:
            fh.write(ln.encode("ascii") + b"\n")
        fh.write(b"\xff\n")
        fh.write(b"d")
        fh.write(np.array([hdr.nx, hdr.ny, hdr.nz], dtype=np.int32).tobytes())
        fh.write(np.array([hdr.dx, hdr.dy, hdr.dz], dtype=np.float64).tobytes())
        fh.write(np.array(hdr.origin, dtype=np.int32).tobytes())
        fh.write(np.array(hdr.mpi_blocks, dtype=np.int32).tobytes())
        fh.write(np.array([hdr.neq, hdr.neqdyn], dtype=np.int32).tobytes())
        fh.write(np.array([hdr.nghost], dtype=np.int32).tobytes())
        fh.write(np.array([hdr.rsc, hdr.vsc, hdr.rhosc], dtype=np.float64).tobytes())
        fh.write(np.array([hdr.cv], dtype=np.float64).tobytes())
        fh.write(u.tobytes(order="F"))


def read_bin(path: str):
    """Warm-start read (src/init.f90:451-467): returns (BinHeader, u)."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = data.index(b"\xff") + 1
    pos += 1  # the LF after achar(255)
    kind = data[pos:pos + 1]
    pos += 1
    if kind != b"d":
        raise ValueError("only 8-byte real dumps are supported")

    def take(dtype, count):
        nonlocal pos
        a = np.frombuffer(data, dtype=dtype, count=count, offset=pos)
        pos += a.nbytes
        return a

    nx, ny, nz = (int(v) for v in take(np.int32, 3))
    dx, dy, dz = (float(v) for v in take(np.float64, 3))
    origin = tuple(int(v) for v in take(np.int32, 3))
    blocks = tuple(int(v) for v in take(np.int32, 3))
    neq, neqdyn = (int(v) for v in take(np.int32, 2))
    nghost = int(take(np.int32, 1)[0])
    rsc, vsc, rhosc = (float(v) for v in take(np.float64, 3))
    cv = float(take(np.float64, 1)[0])
    shape = (neq, nx + 2 * nghost, ny + 2 * nghost, nz + 2 * nghost)
    u = take(np.float64, int(np.prod(shape))).reshape(shape, order="F").copy(order="F")
    return BinHeader(nx, ny, nz, dx, dy, dz, origin, blocks, neq, neqdyn, nghost, rsc, vsc, rhosc, cv), u
