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
    """Real periodic field on n^3 cells: random-phase Gaussian modes with
    integer wavevectors 1 <= |k| <= kmax and amplitude |k|^slope."""
    k = np.fft.fftfreq(n, d=1.0 / n)
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = np.zeros_like(kk)
    m = (kk >= 1.0) & (kk <= kmax)
    amp[m] = kk[m] ** slope
    coef = (rng.standard_normal(kk.shape) + 1j * rng.standard_normal(kk.shape)) * amp
    return np.real(np.fft.ifftn(coef))


def _fill_periodic_ghosts(interior: np.ndarray) -> np.ndarray:
    """(8, n, n, n) interior -> (8, n+4, n+4, n+4) with periodic ghosts."""
    return np.asfortranarray(np.pad(interior, ((0, 0), (2, 2), (2, 2), (2, 2)), mode="wrap"))


def turbulence(n: int, seed: int = 2024, cv: float = 1.5, kmax: int = 8,
               beta: float = 1.0) -> Tuple[np.ndarray, float, float, float]:
    """Seeded MHD turbulence box on [0,1]^3, periodic, gamma = 5/3.

    rho = exp(0.5 g - 0.125), p = exp(0.3 g' - 0.045) with g, g' ~ N(0,1)
    per cell (means 1); v: random-phase modes 1 <= |k| <= 8 with amplitude
    |k|^-2, scaled to rms |v| = sqrt(gamma) (rms Mach 1 at <c_s> = sqrt(gamma));
    B = curl A with cell-centred central differences of a vector potential
    of the same spectrum (central-difference div B = 0 to round-off),
    scaled so <B^2> = 2<p>/beta.
    """
    rng = np.random.default_rng(seed)
    gamma = (cv + 1.0) / cv
    g1 = rng.standard_normal((n, n, n))
    g2 = rng.standard_normal((n, n, n))
    rho = np.exp(0.5 * g1 - 0.125)
    p = np.exp(0.3 * g2 - 0.045)
    v = np.stack([_random_field(rng, n, kmax) for _ in range(3)])
    v *= math.sqrt(gamma) / math.sqrt(np.mean(np.sum(v * v, axis=0)))
    A = np.stack([_random_field(rng, n, kmax) for _ in range(3)])
    h = 1.0 / n

    def d(f, ax):  # central difference, periodic
        return (np.roll(f, -1, axis=ax) - np.roll(f, 1, axis=ax)) / (2.0 * h)

    B = np.stack([d(A[2], 1) - d(A[1], 2), d(A[0], 2) - d(A[2], 0), d(A[1], 0) - d(A[0], 1)])
    B *= math.sqrt(2.0 * np.mean(p) / beta / np.mean(np.sum(B * B, axis=0)))
    u = prim_to_cons(rho, v[0], v[1], v[2], p, B[0], B[1], B[2], cv)
    return _fill_periodic_ghosts(u), h, h, h


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
