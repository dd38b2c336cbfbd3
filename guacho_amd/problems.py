"""Initial conditions of the benchmark problems (the user_mod role,
OT/user_mod.f90:58-66): each returns u(neq, nxmin:nxmax, ...) with every
cell, ghosts included, set, as impose_ot does (OT/orzag_tang.f90:32-68).

Host-side data generation only; the step itself runs on the GPU.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np


def prim_to_cons(rho, vx, vy, vz, p, bx, by, bz, cv):
    """prim2u for full MHD (src/hydro_core.f90:349-373), elementwise in the same order."""
    u = np.empty((8,) + np.shape(rho), dtype=np.float64)
    u[0] = rho
    u[1] = rho * vx
    u[2] = rho * vy
    u[3] = rho * vz
    u[4] = ((0.5 * rho) * ((vx * vx + vy * vy) + vz * vz) + cv * p) + 0.5 * ((bx * bx + by * by) + bz * bz)
    u[5] = bx
    u[6] = by
    u[7] = bz
    return u


def orszag_tang(nx: int, ny: int, nz: int, cv: float = 1.5) -> Tuple[np.ndarray, float, float, float]:
    """Orszag-Tang vortex on [0,1]^2 (OT/orzag_tang.f90:14-70): returns (u, dx, dy, dz).

    dz = dx (zmax = nz/nx), so the z sweep does not change dt.  The sines
    go through the C library's sin, like the reference's, cell by cell.
    """
    pi = math.acos(-1.0)
    twopi = 2.0 * pi
    dx, dy = 1.0 / nx, 1.0 / ny
    dz = (nz / nx) / nz
    rho = 25.0 / (36.0 * pi)
    p = 5.0 / (12.0 * pi)
    s4pi = math.sqrt(4 * pi)
    NX, NY, NZ = nx + 4, ny + 4, nz + 4
    u = np.zeros((8, NX, NY, NZ), dtype=np.float64, order="F")
    for ii in range(NX):
        i = ii - 1
        x = (float(i) + 0.5) * dx * 1.0
        vy = math.sin(x * twopi)
        by = math.sin(2.0 * x * twopi) / s4pi
        for jj in range(NY):
            j = jj - 1
            y = (float(j) + 0.5) * dy * 1.0
            vx = -math.sin(y * twopi)
            bx = -math.sin(y * twopi) / s4pi
            e = 0.5 * rho * (vx * vx + vy * vy + 0.0 * 0.0) + cv * p + 0.5 * (bx * bx + by * by + 0.0 * 0.0)
            u[:, ii, jj, :] = np.array([rho, rho * vx, rho * vy, rho * 0.0, e, bx, by, 0.0])[:, None]
    return u, dx, dy, dz


def _random_field(rng, n: int, kmax: int = 8, slope: float = -2.0) -> np.ndarray:
    """Real periodic field on n^3 cells, indexed [x, y, z] in Fortran order:
    the real part of a sum of random-phase Gaussian modes exp(2 pi i k.x)
    over integer wavevectors 1 <= |k| <= kmax with amplitude |k|^slope, x
    on the cell grid j/n.

    Only (2 kmax + 1)^3 modes are non-zero, so the sum is evaluated as three
    separable partial transforms (small matrix products) instead of a full
    n^3 FFT."""
    k = np.arange(-kmax, kmax + 1)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    amp = np.where((kk >= 1.0) & (kk <= kmax), np.maximum(kk, 1.0) ** slope, 0.0)
    coef = (rng.standard_normal(kk.shape) + 1j * rng.standard_normal(kk.shape)) * amp   # [kx, ky, kz]
    ph = np.exp(2j * np.pi * np.outer(k, np.arange(n) / n))            # (K, n)
    a = np.tensordot(coef, ph, axes=([2], [0]))                         # [kx, ky, z]
    b = np.einsum("abz,by->azy", a, ph)                                 # [kx, z, y]
    b = b.reshape(len(k), n * n)
    br = np.ascontiguousarray(b.real)
    bi = np.ascontiguousarray(b.imag)
    f = np.ascontiguousarray(br.T) @ np.ascontiguousarray(ph.real)      # [(z, y), x]
    f -= np.ascontiguousarray(bi.T) @ np.ascontiguousarray(ph.imag)
    return f.reshape(n, n, n).transpose(2, 1, 0)                        # [x, y, z], Fortran order


def _periodic_state(fields) -> np.ndarray:
    """Eight [x, y, z] interior fields -> u(8, n+4, n+4, n+4) (Fortran order)
    with two periodic ghost layers on every face."""
    n = fields[0].shape[0]
    u = np.empty((8, n + 4, n + 4, n + 4), dtype=np.float64, order="F")
    for q, f in enumerate(fields):
        v = u[q]
        v[2:-2, 2:-2, 2:-2] = f
        v[:2, 2:-2, 2:-2] = v[n:n + 2, 2:-2, 2:-2]
        v[-2:, 2:-2, 2:-2] = v[2:4, 2:-2, 2:-2]
        v[:, :2, 2:-2] = v[:, n:n + 2, 2:-2]
        v[:, -2:, 2:-2] = v[:, 2:4, 2:-2]
        v[:, :, :2] = v[:, :, n:n + 2]
        v[:, :, -2:] = v[:, :, 2:4]
    return u


def _unit_gaussian(f: np.ndarray) -> np.ndarray:
    f = f - f.mean()
    return f / math.sqrt(np.mean(f * f))


def turbulence(n: int, seed: int = 2024, cv: float = 1.5, kmax: int = 8,
               beta: float = 1.0, white_noise: bool = False) -> Tuple[np.ndarray, float, float, float]:
    """Seeded MHD turbulence box on [0,1]^3, periodic, gamma = 5/3.

    rho = exp(0.5 g - 0.125), p = exp(0.3 g' - 0.045) (lognormal, sigma 0.5
    and 0.3, means 1), where g, g' are unit-variance Gaussian random fields
    on the band 1 <= |k| <= kmax (or, with white_noise=True, independent
    N(0,1) per cell -- grid-scale contrasts that drive the scheme to its
    pressure floor at 512^3); v: random-phase modes 1 <= |k| <= 8 with
    amplitude |k|^-2, scaled to rms |v| = sqrt(gamma) (rms Mach 1 at
    <c_s> = sqrt(gamma)); B = curl A with cell-centred central differences
    of a vector potential of the same spectrum (central-difference div B = 0
    to round-off), scaled so <B^2> = 2<p>/beta.
    """
    rng = np.random.default_rng(seed)
    gamma = (cv + 1.0) / cv
    if white_noise:
        g1 = rng.standard_normal((n, n, n)).transpose(2, 1, 0)
        g2 = rng.standard_normal((n, n, n)).transpose(2, 1, 0)
    else:
        g1 = _unit_gaussian(_random_field(rng, n, kmax, slope=-1.0))
        g2 = _unit_gaussian(_random_field(rng, n, kmax, slope=-1.0))
    rho = np.exp(0.5 * g1 - 0.125)
    p = np.exp(0.3 * g2 - 0.045)
    del g1, g2
    v = [_random_field(rng, n, kmax) for _ in range(3)]
    vs = math.sqrt(gamma) / math.sqrt(np.mean(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))
    v = [c * vs for c in v]
    A = [_random_field(rng, n, kmax) for _ in range(3)]
    h = 1.0 / n

    def d(f, ax):  # central difference, periodic
        return (np.roll(f, -1, axis=ax) - np.roll(f, 1, axis=ax)) / (2.0 * h)

    B = [d(A[2], 1) - d(A[1], 2), d(A[0], 2) - d(A[2], 0), d(A[1], 0) - d(A[0], 1)]
    del A
    bs = math.sqrt(2.0 * np.mean(p) / beta / np.mean(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]))
    B = [c * bs for c in B]
    e = ((0.5 * rho) * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + cv * p) \
        + 0.5 * ((B[0] * B[0] + B[1] * B[1]) + B[2] * B[2])
    u = _periodic_state([rho, rho * v[0], rho * v[1], rho * v[2], e, B[0], B[1], B[2]])
    return u, h, h, h


def blast(n: int, seed: int = 7, cv: float = 1.5) -> Tuple[np.ndarray, float, float, float]:
    """Seeded MHD blast on [0,1]^3 with outflow boundaries: ambient rho=1,
    p=0.1; |x-1/2| < 0.1 overpressure 10^U(1,3) times ambient; uniform
    |B| ~ U(0.1, 1) along a random isotropic direction; v = 0."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    x = (np.arange(-1, n + 3) - 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2)
    over = 10.0 ** rng.uniform(1.0, 3.0)
    bmag = rng.uniform(0.1, 1.0)
    mu = rng.uniform(-1.0, 1.0)
    phi = rng.uniform(0.0, 2.0 * math.pi)
    st = math.sqrt(1.0 - mu * mu)
    bdir = (st * math.cos(phi), st * math.sin(phi), mu)
    rho = np.ones_like(r)
    p = np.where(r < 0.1, 0.1 * over, 0.1)
    z = np.zeros_like(r)
    u = prim_to_cons(rho, z, z, z, p, z + bmag * bdir[0], z + bmag * bdir[1], z + bmag * bdir[2], cv)
    return np.asfortranarray(u), h, h, h


def brio_wu(nx: int = 1024, ny: int = 4, nz: int = 4) -> Tuple[np.ndarray, float, float, float]:
    """Brio-Wu shock tube along x on [0,1] (gamma = 2, cv = 1): left rho=1,
    p=1, By=1; right rho=0.125, p=0.1, By=-1; Bx=0.75."""
    cv = 1.0
    h = 1.0 / nx
    x = (np.arange(-1, nx + 3) - 0.5) * h
    left = x < 0.5
    shp = (nx + 4, ny + 4, nz + 4)
    rho = np.broadcast_to(np.where(left, 1.0, 0.125)[:, None, None], shp)
    p = np.broadcast_to(np.where(left, 1.0, 0.1)[:, None, None], shp)
    by = np.broadcast_to(np.where(left, 1.0, -1.0)[:, None, None], shp)
    z = np.zeros(shp)
    u = prim_to_cons(rho, z, z, z, p, z + 0.75, by, z, cv)
    return np.asfortranarray(u), h, h, h
